"""Structural FEA on the voxel grid (src/hex_structural_fea.py:18-234).

``HexStructuralFEA`` keeps the reference constructor and the attributes the
optimisers read (``sol``, ``stiff_mtrx``, ``elem_stiff``, ``deformation``,
``max_deformation``, ``stressComponents``, ``vonMisesStress``,
``pNormStress``, ``strainComponents``, ``elemStrainEnergy``, ``x``).  The
assembled COO matrix is replaced by the matrix-free ``GridOperator``
(``stiff_mtrx``), and the solve runs on the GPU.  Arrays passed as numpy
come back as numpy; CUDA tensors stay on the device.
"""
import time

import numpy as np
import torch

from . import bound_cond, hex_element_stiffness, linear_solvers
from . import topopt_material_model as tmm
from .operator import GridOperator


class HexStructuralFEA:
    """src/hex_structural_fea.py:18-52."""

    def __init__(self, mesh, mat_prop, bc, solver, dsolver=None, elem_body_force=None,
                 thermo_elastic_force=None, **kwargs):
        if isinstance(mat_prop, list):
            raise NotImplementedError("multi-material KE stacks are not part of the GPU path")
        if elem_body_force is not None:
            raise NotImplementedError("element body forces are not part of the GPU path")
        bound_cond.check_supported(bc)
        self.mesh, self.mat_prop, self.bc = mesh, mat_prop, bc
        self.solver, self.kwargs, self.dsolver = solver, kwargs, dsolver
        self.elem_body_force = None
        self.thermo_elastic_force = thermo_elastic_force
        self._material = None
        self.stiff_mtrx = None
        self.set_material(mat_prop)
        self._force = None
        self.x = None
        self.last_stats = None

    def set_material(self, mat_prop):
        """src/hex_structural_fea.py:56-65 (one material): KE of the material;
        the operator is rebuilt when E or nu change."""
        if isinstance(mat_prop, list):
            raise NotImplementedError("multi-material KE stacks are not part of the GPU path")
        self.elem_stiff = np.expand_dims(hex_element_stiffness.hex8_stiffness_matrix_structural(
            mat_prop.youngs_modulus, mat_prop.poissons_ratio, self.mesh.elem_size), axis=0)
        key = (mat_prop.youngs_modulus, mat_prop.poissons_ratio)
        if key != self._material:
            groups = getattr(self.stiff_mtrx, "_groups_from", None)
            self.stiff_mtrx = GridOperator(self.mesh.grid, self.mesh.elem_size, key[0], key[1],
                                           self.bc.fixed_dofs,
                                           penal=tmm.get_SIMP_STRUCTURAL_PENALTY(),
                                           eps_void=tmm.get_EVOID_RELATIVE())
            if groups is not None:
                groups.attach(self.stiff_mtrx)
            self._material = key

    def set_thermal_forces(self, thermo_elastic_force):
        """src/hex_structural_fea.py:68-69."""
        self.thermo_elastic_force = thermo_elastic_force

    def _sync_material_law(self):
        op = self.stiff_mtrx
        p, e = tmm.get_SIMP_STRUCTURAL_PENALTY(), tmm.get_EVOID_RELATIVE()
        if op.penal != p or op.eps_void != e:
            # the SIMP constants are baked into the handle: rebuild it
            groups = getattr(op, "_groups_from", None)
            self.stiff_mtrx = GridOperator(self.mesh.grid, self.mesh.elem_size,
                                           self._material[0], self._material[1],
                                           self.bc.fixed_dofs, penal=p, eps_void=e)
            if groups is not None:
                groups.attach(self.stiff_mtrx)

    def solve(self, x=None, material_model=tmm.MaterialModel.SIMP):
        """src/hex_structural_fea.py:72-149."""
        if material_model != tmm.MaterialModel.SIMP:
            raise ValueError("Invalid material model specified (only SIMP is supported).")
        self._sync_material_law()
        op = self.stiff_mtrx
        as_numpy = x is None or not isinstance(x, torch.Tensor)
        if x is None:
            x = np.ones(self.mesh.num_elems)
        self.x = x
        if self.solver == linear_solvers.Solvers.DPCG:
            if self.dsolver is None:
                raise ValueError("Solvers.DPCG needs a DeflationSolver (dsolver)")
            self.dsolver.attach(op)
        op.set_density(x)
        if self._force is None or self._force_src is not self.bc.force:
            self._force = op.as_dof(self.bc.force)
            self._force_src = self.bc.force
        self.external_force = self.bc.force
        u = linear_solvers.solve(op, self._force, self.solver, self.bc, dsolver=self.dsolver,
                                 **self.kwargs)
        self.last_stats = getattr(self.dsolver, "last_stats", None) or getattr(op, "last_stats", None)
        self.total_force = self.bc.force
        u_elastic = None
        if self.thermo_elastic_force is not None:
            # elastic part first, then the solve with the thermal forces added
            # (src/hex_structural_fea.py:134-147)
            u_elastic = u
            f_total = self._force + op.as_dof(self.thermo_elastic_force)
            self.total_force = f_total
            u = linear_solvers.solve(op, f_total, self.solver, self.bc, dsolver=self.dsolver,
                                     **self.kwargs)
        self._u = u
        d = torch.sqrt(u[0::3] ** 2 + u[1::3] ** 2 + u[2::3] ** 2)
        if as_numpy:
            self.sol = u.cpu().numpy()
            self.deformation = d.cpu().numpy()
            self.max_deformation = float(self.deformation.max())
        else:
            self.sol = u
            self.deformation = d
            self.max_deformation = float(d.max())
        self.solElastic = self.sol
        if u_elastic is not None:
            self.solElastic = u_elastic.cpu().numpy() if as_numpy else u_elastic
        return self.sol

    def postprocess(self):
        """Centre-point strains, relaxed stresses, von Mises, p-norm and
        element strain energy (src/hex_structural_fea.py:152-234)."""
        op = self.stiff_mtrx
        q = tmm.get_stress_relaxation_factor()
        p = tmm.get_pNorm_exponent()
        r = op.stress(self._u, op.x, q=q, pexp=p)
        pn = float(r["sum_vm_p"][0]) ** (1.0 / p)
        self._post = r
        as_numpy = not isinstance(self.sol, torch.Tensor)
        conv = (lambda t: t.cpu().numpy()) if as_numpy else (lambda t: t)
        self.strainComponents = conv(r["strain"])
        self.stressComponents = conv(r["stress"])
        self.vonMisesStress = conv(r["vm"])
        self.elemStrainEnergy = conv(r["energy"])
        self.pNormStress = pn
        return
