"""Boundary conditions (src/bound_cond.py:8-71).

``BC`` keeps the reference's fields.  On the GPU the Dirichlet elimination of
``impose_dirichlet_bc`` is a per-node fixed-DOF mask applied inside the
operator (fixed entries of inputs ignored, of outputs zeroed), which equals
``A[free][:, free]`` for zero Dirichlet values.
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
    """The GPU path eliminates zero Dirichlet values only (all the reference's
    structural recipes use zeros, e.g. src/hex_structural_examples.py:356)."""
    if bc.constraint_matrix is not None:
        raise NotImplementedError("constraint_matrix BCs are not supported "
                                  "(the reference solves them with PARDISO only, src/linear_solvers.py:204-225)")
    dv = np.asarray(bc.dirichlet_values, dtype=np.float64)
    if dv.size and np.any(dv != 0.0):
        raise NotImplementedError("non-zero Dirichlet values are not supported on the GPU path")


def impose_dirichlet_bc(A, b, bc):
    """src/bound_cond.py:40-71 for the matrix-free operator: returns the
    operator (which applies the fixed-DOF mask itself) and b restricted to
    the free DOFs."""
    if bc.constraint_matrix is not None:
        raise NotImplementedError("impose_dirichlet_bc not implemented for constraint_matrix BCs")
    check_supported(bc)
    return A, np.asarray(b)[bc.free_dofs]
