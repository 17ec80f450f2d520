"""Compliance and p-norm stress sensitivities (oracle; test infrastructure only).

Restates ``src/topopt_obj_cons_sensitivities.py``:
  * ``compute_compliance_and_gradient`` ``:148-179``:
    ce = rowsum((U_e KE) o U_e), C = sum E_e ce, dC/dx = -E'(x) ce;
  * ``compute_volume_constraint_and_gradient`` ``:100-116``;
  * ``compute_pnorm_stress_and_sensitivity`` ``:181-287`` with the intended
    math of the adjoint right-hand side: the reference builds g_e with shape
    (1,24) at ``:265`` (``get_stress_relaxation_correction`` reshapes to
    (-1,1), src/topopt_material_model.py:201) and ``g[edof] += g_e`` raises;
    here g_e is the flattened (24,) vector.  x is clipped to >= 1e-12 (``:252``)
    for T1, g and T2 exactly as in the reference;
  * the thermo-elastic force, thermal adjoint right-hand side and coupled
    compliance gradient, src/hex_thermal_fea.py:198-237 and
    ``:322-470``.
"""
import numpy as np
import scipy.sparse.linalg as spla

from . import fea


def element_energy(sol, edof, KE):
    """ce_e = u_e^T KE u_e; src/topopt_obj_cons_sensitivities.py:166."""
    Ue = sol[edof]
    return (Ue @ KE * Ue).sum(1)


def compliance_and_gradient(sol, x, edof, KE, penal=3.0, eps=1e-8):
    """src/topopt_obj_cons_sensitivities.py:148-179 (structural branch)."""
    ce = element_energy(sol, edof, KE)
    C = np.sum(fea.simp_scaling(x, penal, eps) * ce)
    grad = -fea.simp_sensitivity(x, penal, eps) * ce
    return C, grad


def volume_constraint(x, volfrac):
    """src/topopt_obj_cons_sensitivities.py:100-116."""
    return np.mean(x) / volfrac - 1.0, np.ones_like(x) / volfrac / x.size


def B_centre(elem_size):
    """src/topopt_obj_cons_sensitivities.py:204-225 (rows xx,yy,zz,xy,xz,yz)."""
    g = fea.centre_gradN(elem_size)
    Bi = np.zeros((6, 3, 8))
    Bi[0, 0] = g[0]
    Bi[1, 1] = g[1]
    Bi[2, 2] = g[2]
    Bi[3, 0] = g[1]
    Bi[3, 1] = g[0]
    Bi[4, 0] = g[2]
    Bi[4, 2] = g[0]
    Bi[5, 1] = g[2]
    Bi[5, 2] = g[1]
    B = np.zeros((6, 24))
    idx = np.arange(8)
    B[:, (3 * idx)[:, None] + np.arange(3)] = Bi.transpose(0, 2, 1)
    return B


def dvm_dsigma(stress, vm):
    """src/topopt_obj_cons_sensitivities.py:237-248 (vectorised)."""
    s = stress
    d = np.empty_like(s)
    d[:, 0] = (2 * s[:, 0] - s[:, 1] - s[:, 2]) / (2 * vm)
    d[:, 1] = (2 * s[:, 1] - s[:, 0] - s[:, 2]) / (2 * vm)
    d[:, 2] = (2 * s[:, 2] - s[:, 0] - s[:, 1]) / (2 * vm)
    d[:, 3] = 3 * s[:, 3] / vm
    d[:, 4] = 3 * s[:, 4] / vm
    d[:, 5] = 3 * s[:, 5] / vm
    return d


def pnorm_rhs_and_direct(sol, x, mesh, E, nu, q=0.5, pexp=6.0):
    """Adjoint right-hand side g and direct term T1 (intended math).

    src/topopt_obj_cons_sensitivities.py:229-266.
    Returns (pn, vm, g, T1).
    """
    from .ke import isotropic_D
    post = fea.postprocess(mesh, sol, x, E, nu, q, pexp)
    vm = post["vonMisesStress"]
    stress = post["stressComponents"]
    pn = post["pNormStress"]
    D = isotropic_D(E, nu)
    B = B_centre(mesh.elem_size)
    edof = mesh.edofMatStructural
    dpn = np.sum(vm ** pexp) ** (1.0 / pexp - 1.0)
    dv = dvm_dsigma(stress, vm)
    xc = np.maximum(x, 1e-12)
    Ue = sol[edof]
    DBu = (Ue @ B.T) @ D.T                          # (N_el, 6) = D B u_e
    relax_sens = q * xc ** (q - 1.0)                 # src/topopt_material_model.py:203-205
    beta = relax_sens * vm ** (pexp - 1.0) * np.sum(dv * DBu, axis=1)
    T1 = dpn * beta
    ge = (xc ** q * dpn * vm ** (pexp - 1.0))[:, None] * (dv @ D @ B)   # (N_el, 24)
    g = np.zeros(3 * mesh.num_nodes)
    np.add.at(g, edof.ravel(), ge.ravel())
    return pn, vm, g, T1


def pnorm_and_sensitivity(sol, x, mesh, KE, E, nu, solve_fn, penal=3.0, eps=1e-8,
                          q=0.5, pexp=6.0):
    """src/topopt_obj_cons_sensitivities.py:181-287; solve_fn(rhs) -> lambda."""
    pn, vm, g, T1 = pnorm_rhs_and_direct(sol, x, mesh, E, nu, q, pexp)
    lam = solve_fn(g)
    xc = np.maximum(x, 1e-12)
    edof = mesh.edofMatStructural
    ce = (lam[edof] @ KE * sol[edof]).sum(1)
    T2 = -fea.simp_sensitivity(xc, penal, eps) * ce
    return pn, T1 + T2, np.max(vm), lam


def direct_solver(mesh, KE, x, fixed, gvals, penal=3.0, eps=1e-8):
    """Sparse-direct reference solve (scipy splu) used to pin the DPCG."""
    E_e = fea.simp_scaling(x, penal, eps)
    K = fea.assemble_K(mesh, KE, E_e)
    ndof = K.shape[0]
    free = fea.free_dofs(ndof, fixed)
    Kff = K[free, :][:, free].tocsc()
    lu = spla.splu(Kff)

    def solve(b):
        bm = b.copy()
        if len(fixed):
            bm -= K[:, fixed] @ gvals
        return fea.scatter_solution(ndof, fixed, gvals, free, lu.solve(bm[free]))
    return solve, K


# ---------------------------------------------------------------------------
# Thermo-elastic coupling (src/hex_thermal_fea.py:198-237,
# src/topopt_obj_cons_sensitivities.py:322-470).  H is indexed by row 3 j + c
# for element node j and component c, as the reference indexes it.
# ---------------------------------------------------------------------------
def thermoelastic_force(mesh, T, x, H, e_alpha, t_ref, penal=3.0, eps=1e-8):
    """f = sum_e E(x_e) e_alpha H (T_e - t_ref) assembled by nodes;
    src/hex_thermal_fea.py:209-237 (single material)."""
    s = fea.simp_scaling(x, penal, eps) * e_alpha
    fe = s[:, None] * ((T[mesh.elemArray] - t_ref) @ H.T)          # (N_el, 24)
    f = np.zeros(3 * mesh.num_nodes)
    for j in range(8):
        for c in range(3):
            np.add.at(f, 3 * mesh.elemArray[:, j] + c, fe[:, 3 * j + c])
    return f


def thermal_adjoint_rhs(mesh, d, x, H, coef, penal=3.0, eps=1e-8):
    """rhs = sum_e coef E(x_e) H^T d_e; src/topopt_obj_cons_sensitivities.py:352-366
    (coef = -2 E alpha)."""
    s = fea.simp_scaling(x, penal, eps) * coef
    edof = mesh.edofMatStructural
    re = s[:, None] * (d[edof] @ H)                                # (N_el, 8)
    rhs = np.zeros(mesh.num_nodes)
    for k in range(8):
        np.add.at(rhs, mesh.elemArray[:, k], re[:, k])
    return rhs


def thermoelastic_gradient(mesh, x, T, d, lam, KE, KT, H, e_alpha, t_ref, penal_s=3.0,
                           eps_s=1e-8, penal_t=1.0, kvoid=1e-8, exact=False):
    """dJ/dx_e of J = d^T K d, src/topopt_obj_cons_sensitivities.py:428-468:
    -E'(x) d_e^T KE d_e + 2 k'(x) E alpha d_e^T H (T_e - T_ref)
    + k'(x) lambda_e^T KE_T T_e (exact: E'(x) in the middle term)."""
    edof = mesh.edofMatStructural
    de = d[edof]
    Te = T[mesh.elemArray]
    term1 = -fea.simp_sensitivity(x, penal_s, eps_s) * np.einsum("ei,ij,ej->e", de, KE, de)
    kprime = fea.thermal_sensitivity(x, penal_t, kvoid)
    dlaw = fea.simp_sensitivity(x, penal_s, eps_s) if exact else kprime
    term2 = 2.0 * dlaw * e_alpha * np.einsum("ei,ik,ek->e", de, H, Te - t_ref)
    term3 = kprime * np.einsum("ei,ij,ej->e", lam[mesh.elemArray], KT, Te)
    return term1 + term2 + term3
