"""The C ABI: the library loads, exports every symbol include/pyto_hip.h
declares, and its host-only entry point (the element stiffness) equals the
oracle KE.  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

from pyto_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "pyto_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pyto_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert sorted(_lib.EXPORTED) == declared_symbols()


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.pyto_version() >= 1


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.7, 1.3)])
@pytest.mark.parametrize("E,nu", [(1.0, 0.3), (2.1e11, 0.3), (5.0, 0.45)])
def test_host_ke_equals_oracle(h, E, nu):
    from oracle import ke
    from pyto_amd.hex_element_stiffness import hex8_stiffness_matrix_structural
    A = hex8_stiffness_matrix_structural(E, nu, h)
    B = ke.hex8_stiffness(E, nu, h)
    assert np.abs(A - B).max() <= 1e-14 * np.abs(B).max()


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.7, 1.3)])
@pytest.mark.parametrize("k", [1.0, 205.0])
def test_host_thermal_ke_equals_oracle(h, k):
    from oracle import ke
    from pyto_amd.hex_element_stiffness import hex8_stiffness_matrix_thermal
    A = hex8_stiffness_matrix_thermal(k, h)
    B = ke.hex8_stiffness_thermal_exact(k, h)
    assert np.abs(A - B).max() <= 1e-14 * np.abs(B).max()


def test_bad_arguments_raise():
    L = _lib.lib()
    ke = np.zeros(576)
    rc = L.pyto_ke_hex8(1.0, 0.6, 1.0, 1.0, 1.0, ke.ctypes.data_as(ctypes.c_void_p))
    assert rc == _lib.PYTO_EINVAL
    assert b"Poisson" in L.pyto_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)


def test_mma_workspace_argument_checks():
    """Rejected before any device call: more than four constraints, m >= n."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.pyto_mma_create(ctypes.byref(h), 100, 5) == _lib.PYTO_EINVAL
    assert b"m <= 4" in L.pyto_last_error()
    assert L.pyto_mma_create(ctypes.byref(h), 3, 3) == _lib.PYTO_EINVAL
    assert b"m < n" in L.pyto_last_error()
    from pyto_amd import mmaWrapper
    z = np.zeros(10)
    with pytest.raises(ValueError):
        mmaWrapper.mmasub(5, 10, 1, z, z, z + 1, z, z, 0.0, z, np.zeros(5), np.zeros((5, 10)), z,
                          z + 1, 1.0, np.zeros(5), np.ones(5), np.ones(5), 0.2)
