"""GPU element kernels (compliance, stresses, p-norm adjoint pieces) and the
full sensitivities against the oracle, plus central finite differences."""
import numpy as np
import pytest
import torch

from oracle import fea as OF
from oracle import sensitivities as OS
from tests.helpers import Case, rel_err

pytestmark = pytest.mark.gpu


def _u_direct(c):
    solve, _ = OS.direct_solver(c.mesh, c.KE, c.x, c.fixed, c.gv, c.penal, c.eps)
    return solve


@pytest.mark.parametrize("dims", [(60, 20, 10), (17, 9, 5)])
def test_compliance_kernel(gpu, dims):
    print("compliance kernel", dims, flush=True)
    c = Case(dims, seed=1, lo=0.05)
    op = c.operator(gpu, groups=False)
    u = _u_direct(c)(c.f)
    C, g, ce = op.compliance(torch.from_numpy(u).to(gpu), want_ce=True)
    C_ref, g_ref = OS.compliance_and_gradient(u, c.x, c.mesh.edofMatStructural, c.KE, c.penal, c.eps)
    # element energies of rigid-dominated elements cancel (u_e^T KE u_e with a
    # large rigid part), so both sides carry ~1e-12 relative rounding in C
    assert abs(float(C) - C_ref) <= 1e-10 * abs(C_ref)
    # same cancellation (the oracle's dense KE.u cancels the rigid part less
    # exactly than the butterfly's differences)
    assert rel_err(g.cpu().numpy(), g_ref) <= 1e-10
    assert rel_err(ce.cpu().numpy(), OS.element_energy(u, c.mesh.edofMatStructural, c.KE)) <= 1e-12
    # identity C = f^T u
    assert abs(float(C) - c.f @ u) <= 1e-10 * abs(C_ref)


@pytest.mark.parametrize("q,p", [(0.5, 6.0), (0.5, 8.0)])
def test_stress_and_pnorm_pieces(gpu, q, p):
    print("stress/pnorm pieces", q, p, flush=True)
    c = Case((20, 8, 6), seed=2, lo=0.2)
    op = c.operator(gpu, groups=False)
    u = _u_direct(c)(c.f)
    ut = torch.from_numpy(u).to(gpu)
    r = op.stress(ut, q=q, pexp=p)
    post = OF.postprocess(c.mesh, u, c.x, c.E0, c.nu, q=q, pexp=p)
    assert rel_err(r["strain"].cpu().numpy(), post["strainComponents"]) <= 1e-12
    assert rel_err(r["stress"].cpu().numpy(), post["stressComponents"]) <= 1e-12
    assert rel_err(r["vm"].cpu().numpy(), post["vonMisesStress"]) <= 1e-12
    assert rel_err(r["energy"].cpu().numpy(), post["elemStrainEnergy"]) <= 1e-12
    pn = float(r["sum_vm_p"][0]) ** (1 / p)
    assert abs(pn - post["pNormStress"]) <= 1e-12 * post["pNormStress"]
    g, T1 = op.pnorm_adjoint_rhs(ut, c.x, r["vm"], r["sum_vm_p"], q=q, pexp=p)
    pn_ref, vm_ref, g_ref, T1_ref = OS.pnorm_rhs_and_direct(u, c.x, c.mesh, c.E0, c.nu, q, p)
    assert rel_err(g.cpu().numpy(), g_ref) <= 1e-11
    assert rel_err(T1.cpu().numpy(), T1_ref) <= 1e-11


def test_pnorm_sensitivity_end_to_end(gpu):
    """GPU solve + adjoint DPCG vs oracle with direct solves; then central FD
    of the GPU p-norm on 20 seeded elements (exact coarse matrix, tight
    solves so that solver noise / h stays below the 1e-5 target)."""
    from pyto_amd import hex_mesher, hex_structural_fea, linear_solvers, deflation, bound_cond, mat_lib
    from pyto_amd import topopt_material_model as tmm
    from pyto_amd import topopt_obj_cons_sensitivities as sens
    print("pnorm end to end", flush=True)
    dims = (16, 8, 6)
    c = Case(dims, seed=7, lo=0.2)
    solve = _u_direct(c)
    u_ref = solve(c.f)
    pn_ref, grad_ref, vmax_ref, _ = OS.pnorm_and_sensitivity(u_ref, c.x, c.mesh, c.KE, c.E0, c.nu,
                                                            solve, c.penal, c.eps, q=0.5, pexp=8.0)
    mesh = hex_mesher.HexMesher().grid_mesh(dims, (1, 1, 1))
    bc = bound_cond.BC(force=c.f, fixed_dofs=c.fixed, dirichlet_values=np.zeros(len(c.fixed)))
    ds = deflation.DeflationSolver()
    ds.create_deflation_groups(mesh, 10)
    tmm.set_pNorm_exponent(8.0)
    try:
        # FD at h = 1e-6 of a quantity first order in the solve error: solve
        # to 1e-13 (exact coarse matrix; FP64 DPCG breaks down below ~1e-14)
        # so that solver noise / h stays ~1e-8
        fe = hex_structural_fea.HexStructuralFEA(mesh, mat_lib.Material("Unit", 1.0, 0.3), bc,
                                                 linear_solvers.Solvers.DPCG, dsolver=ds, rtol=1e-13,
                                                 maxIters=5000)
        fe.stiff_mtrx.set_coarse_shift(0.0, 0.0)
        u = fe.solve(c.x)
        fe.postprocess()
        pn, grad, vmax = sens.compute_pnorm_stress_and_sensitivity(u, c.x, fe, fe.elem_stiff[0],
                                                                   tmm.MaterialModel.SIMP)
        assert abs(pn - pn_ref) <= 1e-9 * pn_ref
        assert abs(vmax - vmax_ref) <= 1e-9 * vmax_ref
        assert rel_err(grad, grad_ref) <= 1e-7
        # central FD of the GPU p-norm on 20 seeded elements.  At h = 1e-6
        # FP64 differences of PN carry ~1e-7 absolute noise on the gradient
        # for this grid even with sparse-direct solves (the oracle gives the
        # same FD values, e.g. element 567: 2.0638e-4 vs analytic 2.0647e-4,
        # but 2.06472e-4 at h = 1e-5), hence the absolute floor; h = 1e-5
        # is checked with a 10x smaller floor.
        gmax = np.abs(grad).max()
        idx = np.random.default_rng(1).choice(mesh.num_elems, 20, replace=False)
        for h, floor in ((1e-6, 2e-7), (1e-5, 2e-8)):
            for e in idx:
                vals = []
                for s in (+1, -1):
                    xx = c.x.copy()
                    xx[e] += s * h
                    fe.solve(xx)
                    fe.postprocess()
                    vals.append(fe.pNormStress)
                fd = (vals[0] - vals[1]) / (2 * h)
                assert abs(grad[e] - fd) <= 1e-5 * abs(grad[e]) + floor * gmax, (h, e, grad[e], fd)
    finally:
        tmm.set_pNorm_exponent(6)


def test_compliance_fd_on_gpu(gpu):
    from pyto_amd import hex_mesher, hex_structural_fea, linear_solvers, deflation, bound_cond, mat_lib
    from pyto_amd import topopt_material_model as tmm
    from pyto_amd import topopt_obj_cons_sensitivities as sens
    from pyto_amd.topopt_common import FEA_MODE
    print("compliance fd", flush=True)
    dims = (20, 10, 6)
    c = Case(dims, seed=8, lo=0.1)
    mesh = hex_mesher.HexMesher().grid_mesh(dims, (1, 1, 1))
    bc = bound_cond.BC(force=c.f, fixed_dofs=c.fixed, dirichlet_values=np.zeros(len(c.fixed)))
    ds = deflation.DeflationSolver()
    ds.create_deflation_groups(mesh, 20)
    fe = hex_structural_fea.HexStructuralFEA(mesh, mat_lib.Material("Unit", 1.0, 0.3), bc,
                                             linear_solvers.Solvers.DPCG, dsolver=ds, rtol=1e-12)
    fe.stiff_mtrx.set_coarse_shift(0.0, 0.0)
    u = fe.solve(c.x)
    C, grad = sens.compute_compliance_and_gradient(FEA_MODE.STRUCTURAL, u, c.x, fe, fe.elem_stiff[0],
                                                   tmm.MaterialModel.SIMP)
    # C(x) estimated as 2 f^T u - u^T K(x) u, which is second order in the
    # solve error (f^T u alone is first order and swamps the FD at small h)
    h = 1e-5
    op = fe.stiff_mtrx
    for e in np.random.default_rng(2).choice(mesh.num_elems, 10, replace=False):
        vals = []
        for s in (+1, -1):
            xx = c.x.copy()
            xx[e] += s * h
            uu = fe.solve(xx)
            uKu, _, _ = op.compliance(uu, xx, want_grad=False)
            vals.append(2.0 * float(c.f @ uu) - float(uKu))
        assert abs(grad[e] - (vals[0] - vals[1]) / (2 * h)) <= 1e-5 * abs(grad[e])
