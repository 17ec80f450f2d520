"""Thermal FEA on the voxel grid (src/hex_thermal_fea.py:18-270).

``HexThermalFEA`` keeps the reference constructor and the attributes the
optimisers read (``sol``, ``stiff_mtrx``, ``elem_stiff``, ``strain``,
``thermoElasticReferenceTemperature``).  The assembled COO matrix is replaced
by the matrix-free ``ThermalOperator`` (one temperature per node, the 8x8
KE_T scaled by the conductivity law) and the solve runs on the GPU, with
non-zero Dirichlet temperatures eliminated by lifting.

Solvers: ``Solvers.DPCG`` with a DeflationSolver runs the deflated PCG on
the GPU with one constant per deflation group (the thermal coarse space: the
temperature analogue of the rigid-body modes); ``Solvers.PCG``, or DPCG
without a DeflationSolver, runs Jacobi PCG.  The reference's DPCG takes its
thermal W as the rows ``W[free_dofs]`` of the structural rigid-body matrix
(src/linear_solvers.py:184), which pairs temperatures with unrelated
rigid-mode entries; the converged temperature does not depend on W, only the
iteration count does.  Arrays passed as numpy come back as numpy; CUDA
tensors stay on the device.
"""
import ctypes

import numpy as np
import torch

from . import _lib, bound_cond, hex_element_stiffness, linear_solvers
from . import topopt_material_model as tmm
from ._lib import check, ptr, stream_ptr
from .operator import ThermalOperator

_CORNERS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def thermoelastic_coupling_matrix(elem_size, nu, layout="reference"):
    """24 x 8 H with f_e = E alpha E(x) H (T_e - T_ref), exact:
    H[(c, j), k] = int dN_j/dx_c N_k dV / (1 - 2 nu), a product of 1D
    integrals (+-1/2 along c; 1/3 or 1/6 along the other two axes).

    layout "reference": row 8 c + j, as src/hex_thermal_fea.py:119-196
    returns it (the reference then reads row 3 j + c for node j, component c,
    :224-237, so its coupling is this matrix with permuted rows).
    layout "physical": row 3 j + c, the node-major order of the DOFs.
    """
    h = np.asarray(elem_size, dtype=np.float64)
    scale = float(np.prod(h)) / (1.0 - 2.0 * nu)
    H = np.zeros((24, 8))
    for j, cj in enumerate(_CORNERS):
        for c in range(3):
            row = 8 * c + j if layout == "reference" else 3 * j + c
            for k, ck in enumerate(_CORNERS):
                v = scale / h[c] * (0.5 if cj[c] else -0.5)
                for a in range(3):
                    if a != c:
                        v *= (1.0 / 3.0) if cj[a] == ck[a] else (1.0 / 6.0)
                H[row, k] = v
    if layout not in ("reference", "physical"):
        raise ValueError("layout must be 'reference' or 'physical'")
    return H


def structural_law(material_model):
    """(penal, eps) of get_structural_material_model_scaling: x itself without
    a model (src/topopt_material_model.py:84-85), else the SIMP law."""
    if material_model is None:
        return 1.0, 0.0
    tmm._require_simp(material_model)
    return tmm.get_SIMP_STRUCTURAL_PENALTY(), tmm.get_EVOID_RELATIVE()


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
        self.h_layout = "reference"
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
        kwargs = dict(self.kwargs)
        kwargs.setdefault("maxIters", 50000)
        if solver == linear_solvers.Solvers.DPCG:
            if self.dsolver is None or self.dsolver.ws_nGroups == 0:
                solver = linear_solvers.Solvers.PCG
            else:
                kwargs["dsolver"] = self.dsolver
        T = linear_solvers.solve(op, self._force, solver, self.bc, **kwargs)
        self.last_stats = (getattr(self.dsolver, "last_stats", None)
                           if solver == linear_solvers.Solvers.DPCG
                           else getattr(op, "last_stats", None))
        self._T = T
        self.sol = T.cpu().numpy() if as_numpy else T
        return self.sol

    def postprocess(self):
        """Centre temperature gradient in reference coordinates, 3 x num_elems
        (src/hex_thermal_fea.py:242-270)."""
        grad = self.stiff_mtrx.gradient(self._T)
        self.strain = grad.cpu().numpy() if not isinstance(self.sol, torch.Tensor) else grad
        return

    def getHMatrix(self, dx, dy, dz, nu):
        """24 x 8 thermo-elastic coupling matrix (src/hex_thermal_fea.py:119-196);
        ``self.h_layout`` = "reference" (default) or "physical"."""
        return thermoelastic_coupling_matrix((dx, dy, dz), nu, layout=self.h_layout)

    def get_thermoelastic_force(self, x=None, material_model=None):
        """Structural nodal forces E alpha E(x) H (T_e - T_ref) summed over the
        elements (src/hex_thermal_fea.py:198-237), single material."""
        if self.sol is None:
            raise ValueError("Solution not computed yet. Call solve() before get_thermal_force().")
        op = self.stiff_mtrx
        as_numpy = not isinstance(self.sol, torch.Tensor)
        if x is None:
            x = np.ones(self.mesh.num_elems)
        self.x = x
        xd = op.as_elem(x)
        penal, eps = structural_law(material_model)
        mp = self.mat_prop
        H = np.ascontiguousarray(self.getHMatrix(*self.mesh.elem_size, mp.poissons_ratio))
        f = torch.empty(3 * op.num_nodes, dtype=torch.float64, device=op.device)
        check(_lib.lib().pyto_thermoelastic_force(
            op.handle, ptr(op.as_dof(self._T), torch.float64, op.num_nodes), ptr(xd),
            float(penal), float(eps), float(mp.youngs_modulus * mp.thermal_expansion_coefficient),
            float(self.thermoElasticReferenceTemperature), H.ctypes.data_as(ctypes.c_void_p),
            ptr(f), stream_ptr()), "pyto_thermoelastic_force")
        return f.cpu().numpy() if as_numpy else f
