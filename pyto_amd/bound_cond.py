"""Boundary conditions (src/bound_cond.py:8-71).

``BC`` keeps the reference's fields.  On the GPU the Dirichlet elimination of
``impose_dirichlet_bc`` is a per-node fixed-DOF mask applied inside the
operator (fixed entries of inputs ignored, of outputs zeroed), which equals
``A[free][:, free]``.  Non-zero Dirichlet values g are lifted: the right-hand
side becomes b - A[:, fixed] g (one operator pass that keeps g's fixed
entries) and g is written into the fixed entries of the solution.
"""
import dataclasses

import numpy as np


@dataclasses.dataclass
class BC:
    """src/bound_cond.py:8-38."""
    force: np.ndarray
    fixed_dofs: np.ndarray
    dirichlet_values: np.ndarray
    constraint_matrix: np.ndarray = None
    constraint_rhs: np.ndarray = None

    @property
    def num_dofs(self) -> int:
        return self.force.shape[0]

    @property
    def free_dofs(self) -> np.ndarray:
        return np.setdiff1d(np.arange(self.num_dofs), self.fixed_dofs)

    def set_force(self, force):
        self.force = force


def check_supported(bc):
    """Lagrange-multiplier constraints are not part of the GPU path."""
    if bc.constraint_matrix is not None:
        raise NotImplementedError("constraint_matrix BCs are not supported "
                                  "(the reference solves them with PARDISO only, src/linear_solvers.py:204-225)")


def dirichlet_vector(bc, num_dofs):
    """g at the fixed DOFs and 0 elsewhere, or None when every value is 0."""
    dv = np.asarray(bc.dirichlet_values, dtype=np.float64).ravel()
    if dv.size == 0 or not np.any(dv != 0.0):
        return None
    fixed = np.asarray(bc.fixed_dofs, dtype=np.int64).ravel()
    g = np.zeros(num_dofs, dtype=np.float64)
    g[fixed] = np.broadcast_to(dv, fixed.shape) if dv.size in (1, fixed.size) else dv
    return g


def homogeneous(bc):
    """The same fixed DOFs with zero Dirichlet values: the boundary conditions
    of every adjoint solve (the adjoint variable vanishes where the state is
    prescribed, whatever the prescribed value).  Returns bc itself when its
    values are already zero."""
    if dirichlet_vector(bc, bc.num_dofs) is None:
        return bc
    fixed = np.asarray(bc.fixed_dofs)
    return BC(force=bc.force, fixed_dofs=fixed, dirichlet_values=np.zeros(fixed.size))


def impose_dirichlet_bc(A, b, bc):
    """src/bound_cond.py:40-71 for the matrix-free operator: returns the
    operator (which applies the fixed-DOF mask itself) and
    b - A[:, fixed] g restricted to the free DOFs."""
    if bc.constraint_matrix is not None:
        raise NotImplementedError("impose_dirichlet_bc not implemented for constraint_matrix BCs")
    check_supported(bc)
    b = np.asarray(b, dtype=np.float64)
    g = dirichlet_vector(bc, b.shape[0])
    if g is not None:
        b = b - A.lift(g).cpu().numpy()
    return A, b[bc.free_dofs]
