"""Many independent designs and load cases solved together on one GPU.

``BatchSolver(operators, loads)`` takes one ``GridOperator`` per design
(each with its own grid, BCs and densities) and, per design, a list of load
vectors.  Every (design, load case) pair is one system; all systems iterate
in lockstep through batched kernels (one launch per kernel per iteration for
the whole batch, captured as a hipGraph), each with its own scalars and
stopping test.  Per system: PCG in FP32 inside FP64 iterative refinement,
deflated with the operator's coarse space when its handle has deflation
groups (attach them, with ``set_coarse_shift(0, 0)``, before
``set_density``), Jacobi PCG otherwise.  ``autotune()`` picks the operator's tile shape (element
rows of the 16-column tiles, node planes per workgroup) for this batch by
timing the batched FP32 operator.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import SolveStats, check, stream_ptr

F64 = torch.float64


class BatchSolver:
    def __init__(self, operators, loads):
        if len(operators) != len(loads):
            raise ValueError("one list of loads per operator")
        self.ops = list(operators)
        self.systems = []          # (design index, load index)
        self.forces = []
        for d, (op, fl) in enumerate(zip(self.ops, loads)):
            if len(fl) == 0:
                raise ValueError("design %d has no load case" % d)
            for k, f in enumerate(fl):
                ft = op.as_dof(torch.as_tensor(f, dtype=F64))
                if ft.numel() != op.num_dofs:
                    raise ValueError("load %d of design %d has %d DOFs, expected %d"
                                     % (k, d, ft.numel(), op.num_dofs))
                self.systems.append((d, k))
                self.forces.append(ft)
        self.nsys = len(self.systems)
        self.device = self.ops[0].device
        hs = (ctypes.c_void_p * self.nsys)(*[self.ops[d].handle for d, _ in self.systems])
        self._b = ctypes.c_void_p()
        check(_lib.lib().pyto_batch_create(ctypes.byref(self._b), self.nsys, hs, stream_ptr()),
              "pyto_batch_create")
        self.tuning = None

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            _lib.lib().pyto_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_densities(self, xs):
        """New densities for every design at once (one per operator, in
        operator order): element moduli, Jacobi diagonals and Galerkin
        matrices per design, then one batched FP64 Cholesky factorisation and
        inverse of all the designs' coarse matrices."""
        uniq, seen = [], {}
        for op, x in zip(self.ops, xs):
            if id(op) in seen:
                if seen[id(op)] is not x:
                    raise ValueError("a design listed twice needs the same density array")
                continue
            seen[id(op)] = x
            uniq.append((op, x))
        set_density_batch([op for op, _ in uniq], [x for _, x in uniq])

    def info(self):
        z, m, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().pyto_batch_info(self._b, ctypes.byref(z), ctypes.byref(m), ctypes.byref(v)))
        return {"zchunk": z.value, "matvec_blocks": m.value, "vector_blocks": v.value}

    def set_zchunk(self, z):
        check(_lib.lib().pyto_batch_set_zchunk(self._b, int(z), stream_ptr()), "pyto_batch_set_zchunk")

    def time_matvec(self, reps=10):
        ms = ctypes.c_float()
        check(_lib.lib().pyto_batch_time_matvec(self._b, int(reps), ctypes.byref(ms), stream_ptr()))
        return ms.value

    @property
    def tile_rows(self):
        return int(_lib.lib().pyto_batch_get_tile_rows(self._b))

    def set_tile_rows(self, rows):
        """Operator tile rows (8, 16 or 32) of the batch and of its designs'
        operators."""
        check(_lib.lib().pyto_batch_set_tile_rows(self._b, int(rows), stream_ptr()),
              "pyto_batch_set_tile_rows")

    def autotune(self, candidates=(4, 6, 8, 12, 16, 24, 32, 48, 62), rows=(8, 16, 32),
                 reps=10):
        """Time the batched FP32 operator for each tile shape (element rows of
        the 16-column tiles x node planes per workgroup) and keep the fastest.
        Densities must be set (the operator reads E).  The operator's result
        does not depend on the shape."""
        times = {}
        for r in rows:
            self.set_tile_rows(r)
            times[int(r)] = {}
            for z in candidates:
                self.set_zchunk(z)
                self.time_matvec(2)
                times[int(r)][int(z)] = self.time_matvec(reps)
        best = min(((r, z) for r in times for z in times[r]), key=lambda rz: times[rz[0]][rz[1]])
        self.set_tile_rows(best[0])
        self.set_zchunk(best[1])
        self.tuning = {"rows": best[0], "zchunk": best[1], "ms_per_matvec": times}
        return self.tuning

    def solve(self, rtol=1e-8, inner_rtol=1e-2, max_iters=200000, max_outer=60, out=None):
        """Returns (list of solutions per system in system order, list of
        stats dicts).  Systems are ordered design by design, load case by
        load case (``self.systems``)."""
        us = out if out is not None else [torch.empty(self.ops[d].num_dofs, dtype=F64,
                                                      device=self.device)
                                          for d, _ in self.systems]
        fp = (ctypes.c_void_p * self.nsys)(*[f.data_ptr() for f in self.forces])
        up = (ctypes.c_void_p * self.nsys)(*[u.data_ptr() for u in us])
        st = (SolveStats * self.nsys)()
        rc = _lib.lib().pyto_batch_solve(self._b, fp, up, float(rtol), float(inner_rtol),
                                         int(max_iters), int(max_outer), st, stream_ptr())
        check(rc, "batch solve")
        return us, [st[i].as_dict() for i in range(self.nsys)]


def set_density_batch(operators, xs):
    """pyto_set_density for several distinct GridOperators (one density
    array per operator): one batched coarse factorisation for all of them."""
    ops = list(operators)
    if len(ops) != len(xs):
        raise ValueError("one density array per operator")
    if len(set(id(op) for op in ops)) != len(ops):
        raise ValueError("an operator is listed twice")
    xt = [op.as_elem(x) for op, x in zip(ops, xs)]
    for op, x in zip(ops, xt):
        if x.numel() != op.num_elems:
            raise ValueError("density of %d entries for a grid of %d elements"
                             % (x.numel(), op.num_elems))
    hs = (ctypes.c_void_p * len(ops))(*[op.handle for op in ops])
    xp = (ctypes.c_void_p * len(ops))(*[x.data_ptr() for x in xt])
    check(_lib.lib().pyto_set_density_batch(len(ops), hs, xp, stream_ptr()),
          "pyto_set_density_batch")
    for op, x in zip(ops, xt):
        op.x = x


def measure_grid_barrier(blocks_per_cu=1, rounds=2000):
    """Microseconds per grid-wide barrier of a persistent kernel on this GPU
    (pyto_measure_grid_barrier): one 256-thread workgroup on each of
    blocks_per_cu x CU count slots, cooperative launch, the launch itself
    subtracted.  A persistent form of the batched iteration would pay one
    such barrier per dependent phase (per kernel of the captured graph)."""
    _lib.require_gpu()
    us = ctypes.c_float()
    check(_lib.lib().pyto_measure_grid_barrier(int(blocks_per_cu), int(rounds), ctypes.byref(us),
                                               stream_ptr()), "pyto_measure_grid_barrier")
    return float(us.value)


def block_dpcg(op, forces, rtol=1e-8, max_iters=5000):
    """Block deflated PCG (O'Leary's block CG with the deflation projection
    of src/deflation.py:504-584 applied per column) over the load cases of
    one design, in FP64 on the GPU: the operator, Jacobi diagonal and
    coarse space are the handle's (GridOperator.matvec, diagonal, restrict,
    coarse_solve, prolong), the s x s algebra is torch's.  One shared Krylov
    space for all s right-hand sides: Q = K P, alpha = (P^T Q)^-1 (Z^T R),
    X += P alpha, R -= Q alpha, Z = M^-1 R, beta = (Z^T R)_old^-1 (Z^T R),
    P = (I - W E^-1 W^T K)(Z + P beta).  Every column stops on the
    reference's test sqrt(r^T M r / r0^T M r0) <= rtol; the iteration ends
    when all have.  With s = 1 it is the per-system deflated PCG in the same
    form, the baseline of the block measurement (the projection costs a
    second operator per column here, in both forms).  No reference
    counterpart: the reference solves one load case at a time.
    Returns (U as an s x n tensor, {"iterations", "converged", "rel"})."""
    F = torch.stack([torch.as_tensor(f, dtype=F64).to(op.device) for f in forces])
    s = F.shape[0]
    d = op.diagonal()
    free = d > 0
    invd = torch.where(free, 1.0 / torch.where(free, d, torch.ones_like(d)), torch.zeros_like(d))
    F = F * free
    defl = op.ngroups > 0

    def kmul(V):
        return torch.stack([op.matvec(v) for v in V])

    def project(V):   # V - W E^-1 W^T K V, row by row
        if not defl:
            return V
        return torch.stack([v - op.prolong(op.coarse_solve(op.restrict(op.matvec(v)))) for v in V])

    if defl:
        X = torch.stack([op.prolong(op.coarse_solve(op.restrict(f))) for f in F])
    else:
        X = torch.zeros_like(F)
    R = F - kmul(X)
    Z = invd * R
    P = project(Z)
    D = Z @ R.T                       # (Z^T R), s x s
    rz0 = torch.diagonal(D).clone()
    rz0 = torch.where(rz0 > 0, rz0, torch.ones_like(rz0))
    it, rel = 0, float("inf")
    converged = bool((torch.diagonal(D) <= 0).all())
    def small_solve(A, B):   # s x s systems on the host (FP64)
        a = A.cpu().numpy()
        return torch.from_numpy(np.linalg.solve(0.5 * (a + a.T), B.cpu().numpy())).to(op.device)

    while not converged and it < max_iters:
        Q = kmul(P)
        alpha = small_solve(P @ Q.T, D)
        X = X + alpha.T @ P
        R = R - alpha.T @ Q
        Z = invd * R
        Dn = Z @ R.T
        it += 1
        rel = float(torch.sqrt(torch.clamp(torch.diagonal(Dn), min=0.0) / rz0).max())
        if rel <= rtol:
            converged = True
            break
        beta = small_solve(D, Dn)
        P = project(Z + beta.T @ P)
        D = Dn
    return X, {"iterations": it, "converged": converged, "rel": rel, "columns": s}
