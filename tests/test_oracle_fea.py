"""Pins of the oracle's assembly, elimination, DPCG, stresses, sensitivities,
filter and OC update (closed forms, brute force, direct solves, FD)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import deflation as DF
from oracle import fea, filters, ke
from oracle import mesh as M
from oracle import sensitivities as S


def _setup(dims, seed=0, load="traction", lo=0.1, hi=1.0):
    g = M.GridMesh(dims, (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    KE = ke.hex8_stiffness(1.0, 0.3, g.elem_size)
    x = np.random.default_rng(seed).uniform(lo, hi, g.num_elems)
    fixed, gv = fea.clamp_x0(g)
    f = fea.face_traction_load(g) if load == "traction" else fea.point_load(g, dims[0], dims[1] // 2, dims[2] // 2)
    return g, KE, x, fixed, gv, f


def test_assembly_matches_dense_brute_force():
    g, KE, x, *_ = _setup((3, 2, 2))
    E = fea.simp_scaling(x)
    K = fea.assemble_K(g, KE, E).toarray()
    D = np.zeros_like(K)
    for e in range(g.num_elems):
        idx = g.edofMatStructural[e]
        for a in range(24):
            for b in range(24):
                D[idx[a], idx[b]] += E[e] * KE[a, b]
    assert np.abs(K - D).max() <= 1e-14 * np.abs(D).max()
    assert np.allclose(K, K.T)


def test_free_floating_rigid_modes():
    g, KE, x, *_ = _setup((3, 3, 2))
    K = fea.assemble_K(g, KE, fea.simp_scaling(x))
    ngn = np.zeros(g.num_nodes, dtype=np.int32)
    W = DF.deflation_matrix(g.node_xyz, ngn, g.node_xyz.mean(axis=0, keepdims=True))
    assert np.abs(K @ W).max() <= 1e-12 * abs(K).max()


def test_dpcg_matches_direct_solve():
    g, KE, x, fixed, gv, f = _setup((12, 6, 4), seed=1)
    E = fea.simp_scaling(x)
    K = fea.assemble_K(g, KE, E)
    A, b, free = fea.impose_dirichlet(K, f, fixed, gv)
    ngn, cnt, cen = DF.create_groups(g.node_xyz, DF.recommended_groups_ocm(g.num_nodes))
    W = DF.deflation_matrix(g.node_xyz, ngn, cen)[free, :]
    Minv = fea.jacobi_inverse_diag(A)
    x_s = spla.spsolve(A.tocsc(), b)
    # the reference algorithm at its own tolerance
    x_d, info = DF.deflated_pcg(A, b, W, Minv, rtol=1e-8)
    assert info["converged"]
    assert np.linalg.norm(x_d - x_s) <= 1e-6 * np.linalg.norm(x_s)
    # without the 1e-10 coarse shift the same iteration reaches 1e-12
    x_d, info = DF.deflated_pcg(A, b, W, Minv, rtol=1e-12, shift=0.0)
    assert info["converged"]
    assert np.linalg.norm(x_d - x_s) <= 1e-9 * np.linalg.norm(x_s)
    assert np.linalg.norm(b - A @ x_d) <= 1e-10 * np.linalg.norm(b)


def test_reference_coarse_shift_stalls_below_1e9():
    """The reference's E + 1e-10 I (src/deflation.py:521) makes the deflation
    inexact: at rtol 1e-10 this case runs to maxIters and diverges."""
    g, KE, x, fixed, gv, f = _setup((12, 6, 4), seed=1)
    K = fea.assemble_K(g, KE, fea.simp_scaling(x))
    A, b, free = fea.impose_dirichlet(K, f, fixed, gv)
    ngn, cnt, cen = DF.create_groups(g.node_xyz, DF.recommended_groups_ocm(g.num_nodes))
    W = DF.deflation_matrix(g.node_xyz, ngn, cen)[free, :]
    _, info = DF.deflated_pcg(A, b, W, fea.jacobi_inverse_diag(A), rtol=1e-10, maxIters=400)
    assert not info["converged"]


def test_uniaxial_bar_closed_form():
    """Bar 6x1x1 with sliding supports (u=0 at x=0, one corner pinned in y/z
    against rigid motion with roller faces) under uniform end traction: the
    exact solution is uniform strain sigma = F/A, delta = F L / (E A)."""
    nelx = 6
    g = M.GridMesh((nelx, 1, 1), (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    KE = ke.hex8_stiffness(2.0, 0.0, g.elem_size)   # nu = 0: no lateral coupling
    K = fea.assemble_K(g, KE, np.ones(g.num_elems))
    x0 = g.getNodesOnBoundingBoxPlane(0, True)
    y0 = g.getNodesOnBoundingBoxPlane(1, True)
    z0 = g.getNodesOnBoundingBoxPlane(2, True)
    fixed = np.unique(np.concatenate([3 * x0, 3 * y0 + 1, 3 * z0 + 2]))
    f = np.zeros(3 * g.num_nodes)
    xe = g.getNodesOnBoundingBoxPlane(0, False)
    f[3 * xe] = 1.0 / len(xe)                      # total force 1 on area 1
    A, b, free = fea.impose_dirichlet(K, f, fixed, np.zeros(fixed.size))
    u = fea.scatter_solution(3 * g.num_nodes, fixed, np.zeros(fixed.size), free,
                             spla.spsolve(A.tocsc(), b))
    assert np.allclose(u[3 * xe], nelx * 1.0 / 2.0, rtol=1e-12)   # F L / (E A)
    post = fea.postprocess(g, u, np.ones(g.num_elems), 2.0, 0.0, q=0.5, pexp=6)
    assert np.allclose(post["stressComponents"][:, 0], 1.0, rtol=1e-12)
    assert np.allclose(post["vonMisesStress"], 1.0, rtol=1e-12)


def test_von_mises_pure_shear():
    """Pure shear gamma_xy: vm = sqrt(3) * mu * gamma."""
    g = M.GridMesh((1, 1, 1), (1.0, 1.0, 1.0))
    g.createEdofMatStructural()
    gam = 0.01
    u = np.zeros(3 * g.num_nodes)
    u[0::3] = gam * g.node_xyz[:, 1]      # u_x = gamma * y
    post = fea.postprocess(g, u, np.ones(1), 1.0, 0.3, q=0.5, pexp=6)
    mu = 1.0 / (2 * 1.3)
    assert np.isclose(post["stressComponents"][0, 3], mu * gam)
    assert np.isclose(post["vonMisesStress"][0], np.sqrt(3) * mu * gam)


def _direct(g, KE, x, fixed, gv):
    solve, K = S.direct_solver(g, KE, x, fixed, gv)
    return solve, K


def test_compliance_identity_and_fd():
    g, KE, x, fixed, gv, f = _setup((6, 3, 3), seed=2)
    solve, K = _direct(g, KE, x, fixed, gv)
    u = solve(f)
    C, grad = S.compliance_and_gradient(u, x, g.edofMatStructural, KE)
    assert np.isclose(C, f @ u, rtol=1e-12)
    assert np.isclose(C, u @ (K @ u), rtol=1e-12)
    h = 1e-4   # solver rounding (~1e-12 relative in C) limits how small h can be
    for e in [0, 7, 25, 53]:
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        sp_, _ = _direct(g, KE, xp, fixed, gv)
        sm_, _ = _direct(g, KE, xm, fixed, gv)
        Cp = f @ sp_(f)
        Cm = f @ sm_(f)
        assert np.isclose(grad[e], (Cp - Cm) / (2 * h), rtol=1e-6)


def test_pnorm_sensitivity_fd():
    g, KE, x, fixed, gv, f = _setup((6, 3, 3), seed=4, lo=0.3)
    solve, K = _direct(g, KE, x, fixed, gv)
    u = solve(f)
    pn, grad, _, _ = S.pnorm_and_sensitivity(u, x, g, KE, 1.0, 0.3, solve, q=0.5, pexp=8.0)
    h = 1e-5
    for e in [1, 10, 30, 50]:
        vals = []
        for s in (+1, -1):
            xx = x.copy()
            xx[e] += s * h
            so, _ = _direct(g, KE, xx, fixed, gv)
            vals.append(fea.postprocess(g, so(f), xx, 1.0, 0.3, q=0.5, pexp=8.0)["pNormStress"])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert np.isclose(grad[e], fd, rtol=1e-5, atol=1e-9 * abs(pn))


def test_filter_equals_all_pairs_and_closed_form():
    g = M.GridMesh((5, 4, 3), (1.0, 1.0, 1.0))
    H, Hs = filters.cone_filter(g, 1.5)
    H2, Hs2 = filters.cone_filter_all_pairs(g, 1.5)
    assert (H != H2).nnz == 0
    assert np.allclose(Hs, Hs2)
    vals = np.unique(np.round(H.data, 12))
    assert np.allclose(vals, sorted({1.0, 1 / 3, round((1.5 - np.sqrt(2)) / 1.5, 12)}))
    assert (H - H.T).nnz == 0
    interior = 1 + 1 + 2 + 3 * 0  # element (2,1,1) is interior in a 5x4x3 grid
    e = 2 + 5 * 1 + 20 * 1
    assert H[e].nnz == 19
    assert np.isclose(Hs[e], 1 + 6 / 3 + 12 * (1.5 - np.sqrt(2)) / 1.5)


def test_filter_anisotropic_and_radius():
    g = M.GridMesh((6, 5, 4), (1.0, 0.6, 1.4))
    for r in (1.5, 2.2):
        H, _ = filters.cone_filter(g, r)
        H2, _ = filters.cone_filter_all_pairs(g, r)
        assert (H != H2).nnz == 0


def test_oc_update_meets_volume():
    rng = np.random.default_rng(0)
    x = rng.uniform(0.2, 0.8, 500)
    g = -rng.uniform(0.1, 2.0, 500)
    xn = filters.oc_update(x, g, 0.4)
    assert abs(xn.mean() - 0.4) < 1e-5
    assert np.all(np.abs(xn - x) <= 0.2 + 1e-15)
    assert xn.min() >= 0.001 and xn.max() <= 1.0


def test_oc_update_closed_form_multiplier():
    """Uniform gradient -c and move limits inactive: x_new = x sqrt(c/(dv l)),
    so the volume constraint fixes l* = c mean(x)^2 / (dv volfrac^2); the
    bisection must land within its 1e-7 relative stopping width."""
    rng = np.random.default_rng(3)
    n, volfrac, c = 4000, 0.45, 0.7
    x = rng.uniform(0.4, 0.5, n)
    xn, lmid, steps = filters.oc_update(x, np.full(n, -c), volfrac, move=0.5, return_info=True)
    dv = 1.0 / volfrac / n
    lstar = c * x.mean() ** 2 / (dv * volfrac ** 2)
    assert abs(lmid - lstar) <= 2e-7 * lstar
    assert abs(xn.mean() - volfrac) <= 1e-6
    # every step halves the bracket [0, 1e12] until its width relative to the
    # midpoint is <= 1e-7: the count follows from l* alone
    width, expect = 1e12, 0
    lo, hi = 0.0, 1e12
    while (hi - lo) / ((lo + hi) / 2 + 1e-10) > 1e-7:
        expect += 1
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if mid < lstar else (lo, mid)
    assert steps == expect
