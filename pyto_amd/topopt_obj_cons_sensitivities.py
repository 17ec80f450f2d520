"""Objectives, constraints and their sensitivities on the GPU
(src/topopt_obj_cons_sensitivities.py:11-287).

Signatures follow the reference.  ``sol`` and ``x`` may be numpy arrays or
CUDA tensors; results come back in the same kind.  The p-norm sensitivity
implements the reference's intended math (its adjoint right-hand side
``g_e`` is a (1,24) array at :265 and ``g[edof] += g_e`` raises; here g_e is
the 24-vector), with one adjoint DPCG solve on the same operator and BCs.
"""
import numpy as np
import torch

from . import linear_solvers
from . import topopt_material_model as tmm
from .topopt_common import FEA_MODE, TO_QOI


def _like(ref, t):
    return t if isinstance(ref, torch.Tensor) else t.cpu().numpy()


def compute_volume_constraint_and_gradient(x, volfracUpper):
    """src/topopt_obj_cons_sensitivities.py:100-116."""
    if isinstance(x, torch.Tensor):
        grad = torch.ones_like(x) / volfracUpper / x.numel()
    else:
        grad = np.ones_like(x) / volfracUpper / x.size
    return float(x.mean()) / volfracUpper - 1.0, grad


def compute_compliance(sol, x, fe_solver, KE, material_model):
    """src/topopt_obj_cons_sensitivities.py:118-145."""
    C, _ = compute_compliance_and_gradient(FEA_MODE.STRUCTURAL, sol, x, fe_solver, KE, material_model)
    return C


def compute_compliance_and_gradient(feaMode, sol, x, fe_solver, KE, material_model):
    """src/topopt_obj_cons_sensitivities.py:148-179: C = sum E(x) u_e^T KE u_e,
    dC/dx = -E'(x) u_e^T KE u_e (one fused element kernel); in the thermal
    mode C = sum k(x) T_e^T KE_T T_e with the conductivity law."""
    if feaMode not in (FEA_MODE.STRUCTURAL, FEA_MODE.THERMAL):
        raise ValueError("Invalid number of rows in element stiffness matrix.")
    tmm._require_simp(material_model)
    op = fe_solver.stiff_mtrx
    C, grad, _ = op.compliance(sol, x)
    return float(C), _like(sol, grad)


def compute_pnorm_stress_and_sensitivity(sol, x, fe_solver, KE, material_model):
    """src/topopt_obj_cons_sensitivities.py:181-287 -> (pn, grad, max_vm)."""
    tmm._require_simp(material_model)
    op = fe_solver.stiff_mtrx
    q = tmm.get_stress_relaxation_factor()
    p = tmm.get_pNorm_exponent()
    u = op.as_dof(sol)
    xd = op.as_elem(x)
    st = op.stress(u, xd, q=q, pexp=p, want_strain=False, want_stress=False, want_energy=False)
    pn = float(st["sum_vm_p"][0]) ** (1.0 / p)
    g, T1 = op.pnorm_adjoint_rhs(u, xd, st["vm"], st["sum_vm_p"], q=q, pexp=p)
    lam = linear_solvers.solve(op, g, fe_solver.solver, fe_solver.bc, dsolver=fe_solver.dsolver,
                               **fe_solver.kwargs)
    xc = torch.clamp(xd, min=1e-12)
    ce = op.bilinear(lam, u)
    T2 = -tmm.get_structural_material_model_sensitivity(xc, material_model) * ce
    grad = T1 + T2
    return pn, _like(sol, grad), float(st["vm"].max())


def compute_objective_and_gradient(feaMode, to_params, sol, x, fe_solver, KE):
    """src/topopt_obj_cons_sensitivities.py:11-40."""
    kind = to_params.Objective[0]
    mm = to_params.materialModel
    if kind == TO_QOI.COMPLIANCE:
        return compute_compliance_and_gradient(feaMode, sol, x, fe_solver, KE, mm)
    if kind == TO_QOI.VOLUME_FRACTION:
        n = x.numel() if isinstance(x, torch.Tensor) else x.size
        return float(x.mean()), (torch.ones_like(x) if isinstance(x, torch.Tensor) else np.ones_like(x)) / n
    if kind == TO_QOI.PNORM_STRESS:
        pn, g, _ = compute_pnorm_stress_and_sensitivity(sol, x, fe_solver, KE, mm)
        return pn, g
    if kind == TO_QOI.MAX_VONMISES_STRESS:
        pn, g, mx = compute_pnorm_stress_and_sensitivity(sol, x, fe_solver, KE, mm)
        return mx, g
    raise NotImplementedError(f"Objective {kind} is not implemented yet.")
