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

from . import bound_cond, linear_solvers
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
    # adjoint: same fixed DOFs, zero values (lambda vanishes where u is prescribed)
    lam = linear_solvers.solve(op, g, fe_solver.solver, bound_cond.homogeneous(fe_solver.bc),
                               dsolver=fe_solver.dsolver, **fe_solver.kwargs)
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
    if kind == TO_QOI.MASS:
        w = _elem_mass(fe_solver)
        ones = torch.ones_like(x) if isinstance(x, torch.Tensor) else np.ones_like(x)
        return float(x.sum()) * w, ones * w
    if kind == TO_QOI.PNORM_STRESS:
        pn, g, _ = compute_pnorm_stress_and_sensitivity(sol, x, fe_solver, KE, mm)
        return pn, g
    if kind == TO_QOI.MAX_VONMISES_STRESS:
        pn, g, mx = compute_pnorm_stress_and_sensitivity(sol, x, fe_solver, KE, mm)
        return mx, g
    raise NotImplementedError(f"Objective {kind} is not implemented yet.")


def _elem_mass(fe_solver):
    h = fe_solver.mesh.elem_size
    return h[0] * h[1] * h[2] * fe_solver.mat_prop.mass_density


def compute_constraint_and_gradient(feaMode, to_params, sol, x, fe_solver, KE):
    """src/topopt_obj_cons_sensitivities.py:42-96 -> (c, dc): c is an (m, 1)
    numpy array, dc (m, num_elems) in the kind of x (CUDA tensor or numpy).
    The stress-scaled constraints keep the reference's running scale
    ``compute_constraint_and_gradient.stress_scaling`` (:78-93)."""
    tensor = isinstance(x, torch.Tensor)
    n = x.numel() if tensor else x.size
    nC = len(to_params.Constraints)
    c = np.zeros((nC, 1))
    dc = torch.zeros((nC, n), dtype=torch.float64, device=x.device) if tensor else np.zeros((nC, n))
    mm = to_params.materialModel
    for m in range(nC):
        kind, limit = to_params.Constraints[m][0], to_params.Constraints[m][2]
        if kind == TO_QOI.COMPLIANCE:
            C, g = compute_compliance_and_gradient(feaMode, sol, x, fe_solver, KE, mm)
            c[m, 0], dc[m, :] = C / limit - 1.0, g / limit
        elif kind == TO_QOI.MASS:
            w = _elem_mass(fe_solver)
            c[m, 0] = float(x.sum()) * w / limit - 1.0
            dc[m, :] = w / limit
        elif kind == TO_QOI.VOLUME_FRACTION:
            v, g = compute_volume_constraint_and_gradient(x, limit)
            c[m, 0], dc[m, :] = v, g
        elif kind == TO_QOI.PNORM_STRESS:
            pn, g, _ = compute_pnorm_stress_and_sensitivity(sol, x, fe_solver, KE, mm)
            c[m, 0], dc[m, :] = pn / limit - 1.0, g / limit
        elif kind in (TO_QOI.MAX_VONMISES_STRESS, TO_QOI.STRESS_FAILURE_FACTOR):
            pn, g, mx = compute_pnorm_stress_and_sensitivity(sol, x, fe_solver, KE, mm)
            sc = compute_constraint_and_gradient.stress_scaling
            div = limit if kind == TO_QOI.MAX_VONMISES_STRESS else \
                fe_solver.mat_prop.yield_strength * limit
            c[m, 0], dc[m, :] = sc * pn / div - 1.0, sc * g / div
            compute_constraint_and_gradient.stress_scaling = 0.25 * mx / pn + 0.75 * sc
        else:
            raise NotImplementedError(f"Constraint {kind} is not implemented yet.")
    return c, dc


compute_constraint_and_gradient.stress_scaling = 1.0


def solve_thermal_adjoint(x, d, fe_thermal_solver, fe_structural_solver, material_model):
    """src/topopt_obj_cons_sensitivities.py:322-385: K_T lambda =
    -sum_e 2 E alpha E(x_e) H^T d_e with zero temperatures on the thermal
    fixed DOFs, K_T as last assembled by the thermal solve."""
    import ctypes
    from . import _lib, bound_cond
    from .hex_thermal_fea import structural_law
    ts, ss = fe_thermal_solver, fe_structural_solver
    opT, opS = ts.stiff_mtrx, ss.stiff_mtrx
    mp = ss.mat_prop
    H = np.ascontiguousarray(ts.getHMatrix(*ss.mesh.elem_size, mp.poissons_ratio))
    penal, eps = structural_law(material_model)
    rhs = torch.empty(opT.num_nodes, dtype=torch.float64, device=opT.device)
    _lib.check(_lib.lib().pyto_thermoelastic_adjoint_rhs(
        opS.handle, _lib.ptr(opS.as_dof(d), torch.float64, opS.num_dofs),
        _lib.ptr(opS.as_elem(x), torch.float64, opS.num_elems), float(penal), float(eps),
        float(-2.0 * mp.youngs_modulus * mp.thermal_expansion_coefficient),
        H.ctypes.data_as(ctypes.c_void_p), _lib.ptr(rhs), _lib.stream_ptr()),
        "pyto_thermoelastic_adjoint_rhs")
    force = np.zeros(opT.num_nodes)
    bcAdjoint = bound_cond.BC(force=force, fixed_dofs=ts.bc.fixed_dofs,
                              dirichlet_values=np.zeros(np.asarray(ts.bc.fixed_dofs).size))
    solver = ts.solver
    kwargs = dict(ts.kwargs)
    kwargs.setdefault("maxIters", 50000)
    if solver == linear_solvers.Solvers.DPCG:
        # the thermal coarse space of the forward solve, when it has one
        if ts.dsolver is None or ts.dsolver.ws_nGroups == 0:
            solver = linear_solvers.Solvers.PCG
        else:
            kwargs["dsolver"] = ts.dsolver
    lam = linear_solvers.solve(opT, rhs, solver, bcAdjoint, **kwargs)
    return _like(d, lam)


def compute_thermoelastic_compliance_and_gradient(x, temperature, displacement, to_params,
                                                  fe_thermal_solver, fe_structural_solver,
                                                  exact=False):
    """src/topopt_obj_cons_sensitivities.py:387-470: J = d^T K d and
    dJ/dx_e = -E'(x) d_e^T KE d_e + 2 k'(x) E alpha d_e^T H (T_e - T_ref)
              + k'(x) lambda_e^T KE_T T_e,
    with lambda from ``solve_thermal_adjoint``.  The force carries the
    structural law E(x), so its direct derivative is E'(x); the reference
    writes k'(x) (:455-458).  exact=True uses E'(x) there: that is the
    derivative of J (checked by finite differences); the default keeps the
    reference's expression."""
    import ctypes
    from . import _lib
    material_model = to_params.materialModel
    tmm._require_simp(material_model)
    ts, ss = fe_thermal_solver, fe_structural_solver
    opT, opS = ts.stiff_mtrx, ss.stiff_mtrx
    mp = ss.mat_prop
    d = opS.as_dof(displacement)
    T = opT.as_dof(temperature)
    xd = opS.as_elem(x)
    J = float(torch.dot(d, opS.matvec(d)))
    lam = opT.as_dof(solve_thermal_adjoint(xd, d, ts, ss, material_model))
    _, term1, _ = opS.compliance(d, xd)
    H = np.ascontiguousarray(ts.getHMatrix(*ss.mesh.elem_size, mp.poissons_ratio))
    bil = torch.empty(opS.num_elems, dtype=torch.float64, device=opS.device)
    _lib.check(_lib.lib().pyto_thermoelastic_bilinear(
        opS.handle, _lib.ptr(d, torch.float64, opS.num_dofs),
        _lib.ptr(T, torch.float64, opT.num_nodes), float(ts.thermoElasticReferenceTemperature),
        H.ctypes.data_as(ctypes.c_void_p), _lib.ptr(bil), _lib.stream_ptr()),
        "pyto_thermoelastic_bilinear")
    kprime = tmm.get_thermal_material_model_sensitivity(xd, material_model)
    dlaw = tmm.get_structural_material_model_sensitivity(xd, material_model) if exact else kprime
    term2 = 2.0 * dlaw * (mp.youngs_modulus * mp.thermal_expansion_coefficient) * bil
    term3 = kprime * opT.bilinear(lam, T)
    return J, _like(x, term1 + term2 + term3)
