"""Host object around one libpytohip handle: the matrix-free stiffness operator
of a voxel grid and everything that runs on it (solves, element kernels,
cone filter).

All vectors are torch CUDA tensors: DOF vectors are float64 of length
3*num_nodes in the reference's AoS order (num_nodes for ``ThermalOperator``),
element vectors float64 of length num_elems.  The reference-facing modules (hex_structural_fea, linear_solvers,
deflation, topopt_*) are thin layers over this class.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream_ptr

F64 = torch.float64
F32 = torch.float32


def fixed_mask_from_dofs(num_nodes, fixed_dofs):
    """uint8 per node, bit c set when DOF 3*node+c is fixed."""
    fixed = np.asarray(fixed_dofs, dtype=np.int64).ravel()
    if fixed.size and (fixed.min() < 0 or fixed.max() >= 3 * num_nodes):
        raise ValueError("fixed_dofs out of range")
    m = np.zeros(num_nodes, dtype=np.uint8)
    np.bitwise_or.at(m, fixed // 3, (1 << (fixed % 3)).astype(np.uint8))
    return m


def fixed_mask_from_nodes(num_nodes, fixed_nodes):
    """uint8 per node, bit 0 set when the node's temperature is fixed."""
    fixed = np.asarray(fixed_nodes, dtype=np.int64).ravel()
    if fixed.size and (fixed.min() < 0 or fixed.max() >= num_nodes):
        raise ValueError("fixed_dofs out of range")
    m = np.zeros(num_nodes, dtype=np.uint8)
    m[fixed] = 1
    return m


class GridOperator:
    """K(x) of a structured grid of nelx*nely*nelz brick elements.

    Replaces the assembled COO/CSR ``stiff_mtrx`` of
    ``HexStructuralFEA.solve`` (src/hex_structural_fea.py:87-102).
    """

    def __init__(self, num_elems, elem_size, E0, nu, fixed_dofs, penal=3.0, eps_void=1e-8,
                 device=None, fixed_mask=None):
        """``fixed_mask`` (a per-node uint8 CUDA tensor, bit c = DOF c fixed,
        e.g. from pyto_amd.builders.fixed_mask) replaces ``fixed_dofs``."""
        _lib.require_gpu()
        self.device = torch.device(device or "cuda")
        self.grid = tuple(int(v) for v in num_elems)
        self.elem_size = tuple(float(v) for v in elem_size)
        nelx, nely, nelz = self.grid
        self.num_nodes = (nelx + 1) * (nely + 1) * (nelz + 1)
        self.num_elems = nelx * nely * nelz
        self.num_dofs = 3 * self.num_nodes
        self.E0, self.nu, self.penal, self.eps_void = float(E0), float(nu), float(penal), float(eps_void)
        if fixed_mask is not None:
            if fixed_mask.dtype != torch.uint8 or fixed_mask.numel() != self.num_nodes:
                raise ValueError("fixed_mask must be a uint8 tensor of one entry per node")
            self.fixed_mask = fixed_mask.to(self.device).contiguous()
        else:
            mask = fixed_mask_from_dofs(self.num_nodes, fixed_dofs)
            self.fixed_mask = torch.from_numpy(mask).to(self.device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(_lib.lib().pyto_create(ctypes.byref(h), nelx, nely, nelz, *self.elem_size,
                                         self.E0, self.nu, self.penal, self.eps_void,
                                         ptr(self.fixed_mask), stream_ptr()), "pyto_create")
        self._h = h
        self.ngroups = 0
        self.x = None

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().pyto_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("operator is closed")
        return self._h

    def workspace_bytes(self):
        wb = ctypes.c_int64()
        check(_lib.lib().pyto_info(self.handle, None, None, None, ctypes.byref(wb)))
        return wb.value

    # -- state -------------------------------------------------------------
    def set_groups(self, node_group, centers):
        """Deflation groups (int32 per node) and centres (G x 3)."""
        if isinstance(node_group, torch.Tensor):
            g = node_group.to(device=self.device, dtype=torch.int32).contiguous()
        else:
            g = torch.as_tensor(np.ascontiguousarray(node_group, dtype=np.int32)).to(self.device)
        c = torch.as_tensor(np.ascontiguousarray(centers, dtype=np.float64)).to(self.device)
        ng = int(c.shape[0])
        if ng == 0:
            check(_lib.lib().pyto_set_groups(self.handle, None, 0, None, stream_ptr()))
            self.ngroups = 0
            return
        check(_lib.lib().pyto_set_groups(self.handle, ptr(g, torch.int32, self.num_nodes, "node_group"),
                                         ng, ptr(c, F64, 3 * ng, "centers"), stream_ptr()),
              "pyto_set_groups")
        self.ngroups = ng

    def set_coarse_shift(self, abs_shift=1e-10, rel_shift=0.0):
        """Coarse diagonal shift; (1e-10, 0) is the reference's, (0, 0) the
        exact Galerkin matrix.  Applies from the next set_density."""
        check(_lib.lib().pyto_set_coarse_shift(self.handle, float(abs_shift), float(rel_shift)))

    def set_density(self, x):
        """Element densities; also rebuilds Jacobi and the coarse factor."""
        x = self.as_elem(x)
        check(_lib.lib().pyto_set_density(self.handle, ptr(x, F64, self.num_elems, "x"),
                                          stream_ptr()), "pyto_set_density")
        self.x = x
        return x

    # -- conversions -------------------------------------------------------
    def as_elem(self, v):
        t = torch.as_tensor(v, dtype=F64)
        return t.to(self.device).contiguous().reshape(-1)

    def as_dof(self, v):
        t = torch.as_tensor(v, dtype=F64)
        return t.to(self.device).contiguous().reshape(-1)

    # -- operator ------------------------------------------------------------
    def matvec(self, u, out=None):
        if u.dtype == F32:
            out = torch.empty_like(u) if out is None else out
            check(_lib.lib().pyto_matvec_f32(self.handle, ptr(u, F32, self.num_dofs, "u"),
                                             ptr(out, F32, self.num_dofs, "out"), stream_ptr()))
            return out
        u = self.as_dof(u)
        out = torch.empty_like(u) if out is None else out
        check(_lib.lib().pyto_matvec_f64(self.handle, ptr(u, F64, self.num_dofs, "u"),
                                         ptr(out, F64, self.num_dofs, "out"), stream_ptr()))
        return out

    __matmul__ = matvec

    def matvec_mfma(self, u, out=None):
        """FP32 y = K_ff u with the element contraction on the matrix cores
        (the dense KE, pyto_matvec_f32_mfma): a measurement variant of
        matvec for float32 vectors."""
        out = torch.empty_like(u) if out is None else out
        check(_lib.lib().pyto_matvec_f32_mfma(self.handle, ptr(u, F32, self.num_dofs, "u"),
                                              ptr(out, F32, self.num_dofs, "out"), stream_ptr()),
              "pyto_matvec_f32_mfma")
        return out

    def lift(self, g, out=None):
        """K_fc g: the operator with g's fixed entries kept and the output's
        fixed entries zero (FP64); f - lift(g) eliminates the Dirichlet
        values g (src/bound_cond.py:57-66)."""
        g = self.as_dof(g)
        out = torch.empty_like(g) if out is None else out
        check(_lib.lib().pyto_matvec_lift_f64(self.handle, ptr(g, F64, self.num_dofs, "g"),
                                              ptr(out, F64, self.num_dofs, "out"), stream_ptr()))
        return out

    def diagonal(self):
        d = torch.empty(self.num_dofs, dtype=F64, device=self.device)
        check(_lib.lib().pyto_diagonal_f64(self.handle, ptr(d), stream_ptr()))
        return d

    @property
    def shape(self):
        return (self.num_dofs, self.num_dofs)

    # -- launch shape ----------------------------------------------------------
    @property
    def tile_rows(self):
        """Element rows of the operator's 16-column tiles (8, 16 or 32)."""
        return int(_lib.lib().pyto_get_tile_rows(self.handle))

    def set_tile_rows(self, rows):
        check(_lib.lib().pyto_set_tile_rows(self.handle, int(rows), stream_ptr()),
              "pyto_set_tile_rows")

    def autotune_tile_rows(self, candidates=(8, 16, 32), reps=20, precision="fp32"):
        """Times the operator for each tile shape on this grid (HIP events on
        the current stream, after one warm-up launch) and keeps the fastest.
        The operator's result does not depend on the shape.  Needs
        set_density first.  Returns {rows: microseconds per launch}."""
        dtype = F32 if precision == "fp32" else F64
        u = torch.ones(self.num_dofs, dtype=dtype, device=self.device)
        y = torch.empty_like(u)
        times = {}
        for rows in candidates:
            self.set_tile_rows(rows)
            self.matvec(u, y)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                self.matvec(u, y)
            e1.record()
            e1.synchronize()
            times[int(rows)] = e0.elapsed_time(e1) * 1e3 / reps
        self.set_tile_rows(min(times, key=times.get))
        return times

    # -- coarse space --------------------------------------------------------
    # coarse unknowns per group: the six rigid-body modes (one constant per
    # group on the thermal operator)
    modes_per_group = 6

    def restrict(self, v):
        t = torch.empty(self.modes_per_group * self.ngroups, dtype=F64, device=self.device)
        check(_lib.lib().pyto_restrict_f64(self.handle, ptr(self.as_dof(v)), ptr(t), stream_ptr()))
        return t

    def prolong(self, mu):
        v = torch.empty(self.num_dofs, dtype=F64, device=self.device)
        mu = torch.as_tensor(mu, dtype=F64, device=self.device).contiguous()
        check(_lib.lib().pyto_prolong_f64(self.handle, ptr(mu, F64, self.modes_per_group * self.ngroups), ptr(v),
                                          stream_ptr()))
        return v

    def coarse_solve(self, t):
        t = torch.as_tensor(t, dtype=F64, device=self.device).contiguous()
        mu = torch.empty_like(t)
        check(_lib.lib().pyto_coarse_solve_f64(self.handle, ptr(t, F64, self.modes_per_group * self.ngroups),
                                               ptr(mu), stream_ptr()))
        return mu

    def coarse_matrix(self):
        n = self.modes_per_group * self.ngroups
        E = torch.empty((n, n), dtype=F64, device=self.device)
        check(_lib.lib().pyto_coarse_matrix_f64(self.handle, ptr(E), stream_ptr()))
        return E

    # -- solves ----------------------------------------------------------------
    def solve(self, f, rtol=1e-8, max_iters=2000, precision="fp64", inner_rtol=0.1,
              max_outer=40, out=None):
        """u with K_ff u_f = f_f and u = 0 on fixed DOFs.

        precision "fp64": the reference DPCG (src/deflation.py:504-584).
        precision "mixed": FP32 DPCG inside FP64 iterative refinement; each
        refinement step lowers the residual by about inner_rtol and then
        replaces it by the FP64 one, continuing the FP32 search direction
        (reliable update).  Returns (u, stats dict).
        """
        f = self.as_dof(f)
        u = torch.empty_like(f) if out is None else out
        st = _lib.SolveStats()
        L = _lib.lib()
        if precision == "fp64":
            rc = L.pyto_solve_dpcg_f64(self.handle, ptr(f, F64, self.num_dofs, "f"),
                                       ptr(u, F64, self.num_dofs, "u"), float(rtol),
                                       int(max_iters), ctypes.byref(st), stream_ptr())
        elif precision == "mixed":
            rc = L.pyto_solve_dpcg_mixed(self.handle, ptr(f, F64, self.num_dofs, "f"),
                                         ptr(u, F64, self.num_dofs, "u"), float(rtol),
                                         float(inner_rtol), int(max_iters), int(max_outer),
                                         ctypes.byref(st), stream_ptr())
        else:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        check(rc, "solve")
        return u, st.as_dict()

    # -- element kernels -------------------------------------------------------
    def compliance(self, u, x=None, want_ce=False, want_grad=True):
        x = self.x if x is None else self.as_elem(x)
        u = self.as_dof(u)
        C = torch.empty(1, dtype=F64, device=self.device)
        ce = torch.empty(self.num_elems, dtype=F64, device=self.device) if want_ce else None
        g = torch.empty(self.num_elems, dtype=F64, device=self.device) if want_grad else None
        check(_lib.lib().pyto_compliance(self.handle, ptr(u, F64, self.num_dofs), ptr(x, F64, self.num_elems),
                                         ptr(ce), ptr(g), ptr(C), stream_ptr()))
        return C[0], g, ce

    def bilinear(self, a, b):
        out = torch.empty(self.num_elems, dtype=F64, device=self.device)
        check(_lib.lib().pyto_element_bilinear(self.handle, ptr(self.as_dof(a), F64, self.num_dofs),
                                               ptr(self.as_dof(b), F64, self.num_dofs), ptr(out),
                                               stream_ptr()))
        return out

    def stress(self, u, x=None, q=0.5, pexp=6.0, want_strain=True, want_stress=True,
               want_energy=True):
        x = self.x if x is None else self.as_elem(x)
        u = self.as_dof(u)
        ne = self.num_elems
        dev = self.device
        strain = torch.empty((ne, 6), dtype=F64, device=dev) if want_strain else None
        stress = torch.empty((ne, 6), dtype=F64, device=dev) if want_stress else None
        energy = torch.empty(ne, dtype=F64, device=dev) if want_energy else None
        vm = torch.empty(ne, dtype=F64, device=dev)
        s = torch.empty(1, dtype=F64, device=dev)
        check(_lib.lib().pyto_stress(self.handle, ptr(u, F64, self.num_dofs), ptr(x, F64, ne),
                                     float(q), float(pexp), ptr(strain), ptr(stress), ptr(vm),
                                     ptr(energy), ptr(s), stream_ptr()))
        return {"strain": strain, "stress": stress, "vm": vm, "energy": energy, "sum_vm_p": s}

    def pnorm_adjoint_rhs(self, u, x, vm, sum_vm_p, q=0.5, pexp=6.0):
        x = self.as_elem(x)
        g = torch.empty(self.num_dofs, dtype=F64, device=self.device)
        T1 = torch.empty(self.num_elems, dtype=F64, device=self.device)
        check(_lib.lib().pyto_pnorm_adjoint_rhs(self.handle, ptr(self.as_dof(u), F64, self.num_dofs),
                                                ptr(x, F64, self.num_elems), ptr(vm), ptr(sum_vm_p, F64, 1),
                                                float(q), float(pexp), ptr(g), ptr(T1), stream_ptr()))
        return g, T1

    # -- optimality-criteria update ----------------------------------------------
    def oc_update(self, x, dc, volfrac, move=0.2, xmin=0.001, xmax=1.0, max_steps=160,
                  out=None, read_back=False):
        """OC update with the reference's volume bisection run on the device
        (src/topopt_ocm.py:247-269).  Returns xnew, or (xnew, lmid, steps)
        with read_back (which synchronises)."""
        x = self.as_elem(x)
        dc = self.as_elem(dc)
        n = x.numel()
        if dc.numel() != n:
            raise ValueError("x and dc differ in length")
        xnew = torch.empty_like(x) if out is None else out
        dv = 1.0 / float(volfrac) / n
        lm = ctypes.c_double()
        st = ctypes.c_int()
        check(_lib.lib().pyto_oc_update(self.handle, ptr(x, F64, n), ptr(dc, F64, n), n, dv,
                                        float(volfrac), float(move), float(xmin), float(xmax),
                                        int(max_steps), ptr(xnew, F64, n),
                                        ctypes.byref(lm) if read_back else None,
                                        ctypes.byref(st) if read_back else None, stream_ptr()),
              "pyto_oc_update")
        if read_back:
            return xnew, lm.value, st.value
        return xnew

    # -- cone filter -------------------------------------------------------------
    def filter(self, v, rel_radius=1.5, mode=0, aux=None):
        v = self.as_elem(v) if v is not None else None
        aux = self.as_elem(aux) if aux is not None else None
        out = torch.empty(self.num_elems, dtype=F64, device=self.device)
        check(_lib.lib().pyto_filter_apply(self.handle, float(rel_radius), int(mode),
                                           ptr(v, F64, self.num_elems) if v is not None else None,
                                           ptr(aux, F64, self.num_elems) if aux is not None else None,
                                           ptr(out), stream_ptr()))
        return out

    def filter_stencil(self, rel_radius=1.5):
        n = ctypes.c_int()
        L = _lib.lib()
        check(L.pyto_filter_stencil(self.handle, float(rel_radius), 0, ctypes.byref(n), None, None))
        off = np.zeros((n.value, 3), dtype=np.int32)
        w = np.zeros(n.value, dtype=np.float64)
        check(L.pyto_filter_stencil(self.handle, float(rel_radius), n.value, ctypes.byref(n),
                                    off.ctypes.data_as(ctypes.c_void_p),
                                    w.ctypes.data_as(ctypes.c_void_p)))
        return off, w


class ThermalOperator(GridOperator):
    """K_T(x) of a structured grid: one temperature per node, the 8x8 KE_T of
    src/hex_element_stiffness.py:339-385 scaled by the conductivity law
    k_e = k0 (k_void + x^penal (1 - k_void)) (src/topopt_material_model.py:98-109).

    Replaces the assembled ``stiff_mtrx`` of ``HexThermalFEA.solve``
    (src/hex_thermal_fea.py:74-117).  matvec, diagonal, set_density, solve
    and compliance act on vectors of length num_nodes.  Deflation groups
    (``set_groups``) span one constant per group: W[n, g] = 1 for the free
    nodes n of group g, E = W^T K_T W is G x G, and solve runs the deflated
    PCG with that coarse space (Jacobi PCG without groups).
    """
    modes_per_group = 1

    def __init__(self, num_elems, elem_size, k0, fixed_nodes, penal=1.0, k_void=1e-8,
                 device=None):
        _lib.require_gpu()
        self.device = torch.device(device or "cuda")
        self.grid = tuple(int(v) for v in num_elems)
        self.elem_size = tuple(float(v) for v in elem_size)
        nelx, nely, nelz = self.grid
        self.num_nodes = (nelx + 1) * (nely + 1) * (nelz + 1)
        self.num_elems = nelx * nely * nelz
        self.num_dofs = self.num_nodes
        self.E0, self.nu, self.penal, self.eps_void = float(k0), 0.0, float(penal), float(k_void)
        mask = fixed_mask_from_nodes(self.num_nodes, fixed_nodes)
        self.fixed_mask = torch.from_numpy(mask).to(self.device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(_lib.lib().pyto_create_thermal(ctypes.byref(h), nelx, nely, nelz,
                                                 *self.elem_size, self.E0, self.penal,
                                                 self.eps_void, ptr(self.fixed_mask),
                                                 stream_ptr()), "pyto_create_thermal")
        self._h = h
        self.ngroups = 0
        self.x = None

    def gradient(self, T):
        """Centre temperature gradient in reference coordinates, 3 x num_elems
        (src/hex_thermal_fea.py:250-270)."""
        out = torch.empty((3, self.num_elems), dtype=F64, device=self.device)
        check(_lib.lib().pyto_thermal_gradient(self.handle, ptr(self.as_dof(T), F64, self.num_dofs),
                                               ptr(out), stream_ptr()), "pyto_thermal_gradient")
        return out
