"""Linear solver dispatch (src/linear_solvers.py:126-202).

``solve(A0, b0, solver, bc, **kwargs)`` keeps the reference signature.  A0 is
the ``GridOperator`` that ``HexStructuralFEA`` exposes as ``stiff_mtrx`` (or
the ``ThermalOperator`` of ``HexThermalFEA``).
DPCG runs the deflated PCG on the GPU (kwargs: dsolver, rtol, maxIters and
``precision`` = "fp64" (default, the reference algorithm) or "mixed"); PCG
runs Jacobi PCG (no deflation).  The direct and external solvers of the
reference (SPSOLVE, SPLU, PYAMG, PARDISO, PETSC) are not part of this path.
"""
import enum

import numpy as np
import torch

from . import bound_cond

DEFAULT_TOL = 1.e-6


class Solvers(enum.Enum):
    """src/linear_solvers.py:126-133."""
    SPSOLVE = enum.auto()
    PCG = enum.auto()
    PYAMG = enum.auto()
    DPCG = enum.auto()
    PARDISO = enum.auto()
    SPLU = enum.auto()
    PETSC = enum.auto()


def _to_device(A0, b0):
    b = torch.as_tensor(b0, dtype=torch.float64)
    return b.to(A0.device).contiguous()


def solve(A0, b0, solver, bc, **kwargs):
    """src/linear_solvers.py:136-202.  Returns u like b0 (numpy in, numpy
    out; a CUDA tensor in, a CUDA tensor out)."""
    if bc.constraint_matrix is not None:
        raise ValueError('Only Pardiso and DPCG solvers support Lagrange multipliers currently.')
    bound_cond.check_supported(bc)
    rtol = kwargs.get('rtol', DEFAULT_TOL)
    maxIters = kwargs.get('maxIters', 2000)
    precision = kwargs.get('precision', 'fp64')
    b = _to_device(A0, b0)
    # non-zero Dirichlet values: b' = b - A[:, fixed] g, u[fixed] = g
    # (src/bound_cond.py:57-66, src/linear_solvers.py:199-201)
    g = bound_cond.dirichlet_vector(bc, A0.num_dofs)
    if g is not None:
        g = torch.from_numpy(g).to(A0.device)
        b = b - A0.lift(g)
    if solver == Solvers.DPCG:
        dsolver = kwargs['dsolver']
        u = dsolver.deflatedPCG(A0, b, rtol=rtol, maxIters=maxIters, precision=precision)
    elif solver == Solvers.PCG:
        if A0.ngroups:
            A0.set_groups(np.zeros(0, np.int32), np.zeros((0, 3)))
            A0._groups_from = None
        u, A0.last_stats = A0.solve(b, rtol=rtol, max_iters=maxIters, precision=precision)
    elif solver in (Solvers.SPSOLVE, Solvers.SPLU, Solvers.PYAMG, Solvers.PARDISO, Solvers.PETSC):
        raise NotImplementedError("%s is not part of the GPU path; use Solvers.DPCG" % solver.name)
    else:
        raise ValueError('Unknown solver type')
    if g is not None:
        u = u + g
    if isinstance(b0, torch.Tensor):
        return u
    return u.cpu().numpy()
