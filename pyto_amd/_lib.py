"""ctypes binding of libpytohip.so (the C ABI declared in include/pyto_hip.h).

The library is loaded from this package directory.  There is no fallback:
every GPU entry point raises if the library is missing, so nothing silently
runs on the CPU.
"""
import ctypes
import os
import warnings

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpytohip.so")

PYTO_OK = 0
PYTO_EINVAL = -1
PYTO_EHIP = -2
PYTO_ENOMEM = -3
PYTO_ESTATE = -4
PYTO_ENOCONV = -5
PYTO_ENUMERIC = -6


class ConvergenceWarning(UserWarning):
    """The iterative solver stopped at its iteration limit."""


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("matvecs64", ctypes.c_int),
                ("matvecs32", ctypes.c_int), ("rel", ctypes.c_double),
                ("rel_true", ctypes.c_double), ("setup_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MmaStats(ctypes.Structure):
    _fields_ = [("newton_steps", ctypes.c_int), ("evaluations", ctypes.c_int),
                ("levels", ctypes.c_int), ("rounds", ctypes.c_int),
                ("residual_norm", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int

_SIGS = {
    "pyto_last_error": ([], ctypes.c_char_p),
    "pyto_version": ([], _i),
    "pyto_ke_hex8": ([_d, _d, _d, _d, _d, _vp], _i),
    "pyto_create": ([ctypes.POINTER(_vp), _i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _vp, _vp], _i),
    "pyto_ke_hex8_thermal": ([_d, _d, _d, _d, _vp], _i),
    "pyto_matvec_lift_f64": ([_vp, _vp, _vp, _vp], _i),
    "pyto_thermal_gradient": ([_vp, _vp, _vp, _vp], _i),
    "pyto_thermoelastic_force": ([_vp, _vp, _vp, _d, _d, _d, _d, _vp, _vp, _vp], _i),
    "pyto_thermoelastic_adjoint_rhs": ([_vp, _vp, _vp, _d, _d, _d, _vp, _vp, _vp], _i),
    "pyto_thermoelastic_bilinear": ([_vp, _vp, _vp, _d, _vp, _vp, _vp], _i),
    "pyto_mma_create": ([ctypes.POINTER(_vp), ctypes.c_int64, _i], _i),
    "pyto_mma_destroy": ([_vp], _i),
    "pyto_mma_sub": ([_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp, _d,
                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pyto_mma_kkt": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pyto_grid_arrays": ([_i, _i, _i, _d, _d, _d, _vp, _vp, _vp, _vp], _i),
    "pyto_plane_mask": ([_i, _i, _i, _i, _i, _i, _vp, _vp], _i),
    "pyto_face_load": ([_i, _i, _i, _i, _i, _i, _i, _d, _vp, _vp], _i),
    "pyto_node_groups": ([_i, _i, _i, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp], _i),
    "pyto_group_sums": ([_i, _i, _i, _d, _d, _d, _vp, _i, _vp, _vp, _vp], _i),
    "pyto_element_components": ([_i, _i, _i, _vp, _d, _vp, ctypes.POINTER(_i), _vp], _i),
    "pyto_create_thermal": ([ctypes.POINTER(_vp), _i, _i, _i, _d, _d, _d, _d, _d, _d, _vp, _vp],
                            _i),
    "pyto_destroy": ([_vp], _i),
    "pyto_info": ([_vp, _vp, _vp, _vp, _vp], _i),
    "pyto_set_density": ([_vp, _vp, _vp], _i),
    "pyto_set_density_batch": ([_i, _vp, _vp, _vp], _i),
    "pyto_set_groups": ([_vp, _vp, _i, _vp, _vp], _i),
    "pyto_matvec_f64": ([_vp, _vp, _vp, _vp], _i),
    "pyto_matvec_f32": ([_vp, _vp, _vp, _vp], _i),
    "pyto_matvec_f32_mfma": ([_vp, _vp, _vp, _vp], _i),
    "pyto_set_tile_rows": ([_vp, _i, _vp], _i),
    "pyto_get_tile_rows": ([_vp], _i),
    "pyto_diagonal_f64": ([_vp, _vp, _vp], _i),
    "pyto_restrict_f64": ([_vp, _vp, _vp, _vp], _i),
    "pyto_prolong_f64": ([_vp, _vp, _vp, _vp], _i),
    "pyto_coarse_solve_f64": ([_vp, _vp, _vp, _vp], _i),
    "pyto_coarse_matrix_f64": ([_vp, _vp, _vp], _i),
    "pyto_set_coarse_shift": ([_vp, _d, _d], _i),
    "pyto_solve_dpcg_f64": ([_vp, _vp, _vp, _d, _i, ctypes.POINTER(SolveStats), _vp], _i),
    "pyto_solve_dpcg_mixed": ([_vp, _vp, _vp, _d, _d, _i, _i, ctypes.POINTER(SolveStats), _vp], _i),
    "pyto_compliance": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pyto_element_bilinear": ([_vp, _vp, _vp, _vp, _vp], _i),
    "pyto_stress": ([_vp, _vp, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "pyto_pnorm_adjoint_rhs": ([_vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _vp, _vp], _i),
    "pyto_filter_apply": ([_vp, _d, _i, _vp, _vp, _vp, _vp], _i),
    "pyto_filter_stencil": ([_vp, _d, _i, ctypes.POINTER(_i), _vp, _vp], _i),
    "pyto_batch_create": ([ctypes.POINTER(_vp), _i, _vp, _vp], _i),
    "pyto_batch_destroy": ([_vp], _i),
    "pyto_batch_set_zchunk": ([_vp, _i, _vp], _i),
    "pyto_measure_grid_barrier": ([_i, _i, ctypes.POINTER(ctypes.c_float), _vp], _i),
    "pyto_batch_set_tile_rows": ([_vp, _i, _vp], _i),
    "pyto_batch_get_tile_rows": ([_vp], _i),
    "pyto_batch_info": ([_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)], _i),
    "pyto_batch_time_matvec": ([_vp, _i, ctypes.POINTER(ctypes.c_float), _vp], _i),
    "pyto_batch_solve": ([_vp, _vp, _vp, _d, _d, _i, _i, _vp, _vp], _i),
    "pyto_oc_update": ([_vp, _vp, _vp, ctypes.c_int64, _d, _d, _d, _d, _d, _i, _vp,
                        ctypes.POINTER(_d), ctypes.POINTER(_i), _vp], _i),
    "pyto_oc_update_noalloc": ([_vp, _vp, ctypes.c_int64, _d, _d, _d, _d, _d, _i, _vp,
                                ctypes.POINTER(_d), ctypes.POINTER(_i), _vp], _i),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libpytohip.so (raises RuntimeError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpytohip.so not found at %s; build it with "
                               "`python -m pyto_amd.csrc.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc, what=""):
    if rc == PYTO_OK:
        return rc
    msg = lib().pyto_last_error().decode(errors="replace")
    if what:
        msg = "%s: %s" % (what, msg)
    if rc == PYTO_EINVAL:
        raise ValueError(msg)
    if rc == PYTO_ENOCONV:
        warnings.warn(msg, ConvergenceWarning, stacklevel=3)
        return rc
    if rc == PYTO_ENUMERIC:
        import numpy.linalg
        raise numpy.linalg.LinAlgError(msg)
    if rc == PYTO_ESTATE:
        raise RuntimeError("invalid call order: " + msg)
    if rc == PYTO_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("pyto_amd needs a ROCm GPU (torch.cuda.is_available() is False)")


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t, dtype=None, numel=None, name="tensor"):
    """Device pointer of a contiguous CUDA tensor, with dtype/size checks."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a CUDA tensor" % name)
    if dtype is not None and t.dtype != dtype:
        raise TypeError("%s must have dtype %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if numel is not None and t.numel() != numel:
        raise ValueError("%s must have %d elements, got %d" % (name, numel, t.numel()))
    return ctypes.c_void_p(t.data_ptr())
