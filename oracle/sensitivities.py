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
    for T1, g and T2 exactly as in the reference.
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
