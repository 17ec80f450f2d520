"""Thermal FEA on the voxel grid (src/hex_thermal_fea.py:18-270).

``HexThermalFEA`` keeps the reference constructor and the attributes the
optimisers read (``sol``, ``stiff_mtrx``, ``elem_stiff``, ``strain``,
``thermoElasticReferenceTemperature``).  The assembled COO matrix is replaced
by the matrix-free ``ThermalOperator`` (one temperature per node, the 8x8
KE_T scaled by the conductivity law) and the solve runs on the GPU, with
non-zero Dirichlet temperatures eliminated by lifting.

Solvers: the thermal operator carries no deflation space, so ``Solvers.DPCG``
and ``Solvers.PCG`` both run Jacobi PCG on the GPU (the converged temperature
does not depend on the deflation space; the reference's DPCG takes its
thermal W as the rows ``W[free_dofs]`` of the structural rigid-body matrix,
src/linear_solvers.py:184).  Arrays passed as numpy come back as numpy;
CUDA tensors stay on the device.
"""
import numpy as np
import torch

from . import bound_cond, hex_element_stiffness, linear_solvers
from . import topopt_material_model as tmm
from .operator import ThermalOperator


class HexThermalFEA:
    """src/hex_thermal_fea.py:18-52."""

    def __init__(self, mesh, mat_prop, bc, solver, dsolver=None, elem_body_force=None,
                 thermoElasticReferenceTemperature=23.0, **kwargs):
        if isinstance(mat_prop, list):
            raise NotImplementedError("multi-material KE stacks are not part of the GPU path")
        if elem_body_force is not None:
            raise NotImplementedError("element heat sources are not part of the GPU path")
        bound_cond.check_supported(bc)
        self.mesh, self.mat_prop, self.bc = mesh, mat_prop, bc
        self.thermoElasticReferenceTemperature = thermoElasticReferenceTemperature
        self.solver, self.kwargs, self.dsolver = solver, kwargs, dsolver
        self.elem_body_force = None
        self.sol = None
        self.x = None
        self.last_stats = None
        self.elem_stiff = np.expand_dims(hex_element_stiffness.hex8_stiffness_matrix_thermal(
            mat_prop.thermal_conductivity, mesh.elem_size), axis=0)
        self._law = None
        self.stiff_mtrx = None
        self._force = None
        self._force_src = None
        self._operator(tmm.MaterialModel.SIMP)

    def set_material(self, mat_prop):
        """src/hex_thermal_fea.py:55-71 (single material)."""
        if isinstance(mat_prop, list):
            raise NotImplementedError("multi-material KE stacks are not part of the GPU path")
        self.mat_prop = mat_prop
        self.elem_stiff = np.expand_dims(hex_element_stiffness.hex8_stiffness_matrix_thermal(
            mat_prop.thermal_conductivity, self.mesh.elem_size), axis=0)
        self._law = None

    def _operator(self, material_model):
        # the conductivity law is baked into the handle: k = x (no model,
        # src/topopt_material_model.py:106-107) or the thermal SIMP law
        if material_model is None:
            law = (1.0, 0.0)
        else:
            tmm._require_simp(material_model)
            law = (tmm.get_SIMP_THERMAL_PENALTY(), tmm.get_KVOID_RELATIVE())
        law = law + (float(self.mat_prop.thermal_conductivity),)
        if self.stiff_mtrx is None or self._law != law:
            # (a previous operator stays alive for whoever still holds it,
            # e.g. a filter built on it)
            self.stiff_mtrx = ThermalOperator(self.mesh.grid, self.mesh.elem_size, law[2],
                                              self.bc.fixed_dofs, penal=law[0], k_void=law[1])
            self._law = law
        return self.stiff_mtrx

    def solve(self, x=None, material_model=None, elem_mat_id=None):
        """src/hex_thermal_fea.py:74-117."""
        if elem_mat_id is not None:
            raise NotImplementedError("multi-material element ids are not part of the GPU path")
        op = self._operator(material_model)
        as_numpy = x is None or not isinstance(x, torch.Tensor)
        if x is None:
            x = np.ones(self.mesh.num_elems)
        self.x = x
        op.set_density(x)
        if self._force is None or self._force_src is not self.bc.force:
            self._force = op.as_dof(self.bc.force)
            self._force_src = self.bc.force
        solver = self.solver
        if solver == linear_solvers.Solvers.DPCG:
            solver = linear_solvers.Solvers.PCG
        kwargs = dict(self.kwargs)
        kwargs.setdefault("maxIters", 50000)
        T = linear_solvers.solve(op, self._force, solver, self.bc, **kwargs)
        self.last_stats = getattr(op, "last_stats", None)
        self._T = T
        self.sol = T.cpu().numpy() if as_numpy else T
        return self.sol

    def postprocess(self):
        """Centre temperature gradient in reference coordinates, 3 x num_elems
        (src/hex_thermal_fea.py:242-270)."""
        grad = self.stiff_mtrx.gradient(self._T)
        self.strain = grad.cpu().numpy() if not isinstance(self.sol, torch.Tensor) else grad
        return
