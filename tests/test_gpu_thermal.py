"""GPU parity of the thermal path (one temperature per node) against the
oracle's assembled K_T and sparse direct solves, plus the Dirichlet lifting
on the structural operator."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from oracle import fea as OF
from oracle import filters as OFL
from oracle import ke as OK
from oracle import mesh as OM
from tests.helpers import Case, rel_err

pytestmark = pytest.mark.gpu


class TCase:
    """Seeded thermal problem: T fixed to T0 on the x = 0 face and to T1 on
    the lower half of the x = max face, unit heat load on the top face."""

    def __init__(self, dims, seed=0, h=(1.0, 1.0, 1.0), k0=1.0, penal=1.0, kvoid=1e-8, lo=0.01,
                 hi=1.0, T0=20.0, T1=100.0, x=None):
        self.dims, self.h, self.k0, self.penal, self.kvoid = dims, h, k0, penal, kvoid
        m = OM.GridMesh(dims, h)
        self.mesh = m
        self.KE = OK.hex8_stiffness_thermal(k0, h)
        rng = np.random.default_rng(seed)
        self.x = rng.uniform(lo, hi, m.num_elems) if x is None else x
        self.K = OF.assemble_KT(m, self.KE, OF.thermal_scaling(self.x, penal, kvoid))
        ix, iy, iz = m.node_indices[:, 0], m.node_indices[:, 1], m.node_indices[:, 2]
        a = np.nonzero(ix == 0)[0]
        b = np.nonzero((ix == dims[0]) & (iz <= dims[2] // 2))[0]
        self.fixed = np.concatenate([a, b])
        self.gv = np.concatenate([np.full(a.size, T0), np.full(b.size, T1)])
        self.f = np.zeros(m.num_nodes)
        self.f[iz == dims[2]] = 1.0
        self.f[self.fixed] = 0.0
        self.mask = np.ones(m.num_nodes)
        self.mask[self.fixed] = 0.0

    def operator(self, device):
        from pyto_amd.operator import ThermalOperator
        op = ThermalOperator(self.dims, self.h, self.k0, self.fixed, penal=self.penal,
                             k_void=self.kvoid, device=device)
        op.set_density(self.x)
        return op

    def solve_oracle(self):
        A, b, free = OF.impose_dirichlet(self.K, self.f, self.fixed, self.gv)
        return OF.scatter_solution(self.mesh.num_nodes, self.fixed, self.gv, free,
                                   spla.spsolve(A.tocsc(), b))


GRIDS = [(1, 1, 1), (3, 2, 2), (16, 15, 3), (31, 17, 5), (40, 20, 10)]


@pytest.mark.parametrize("dims", GRIDS)
def test_thermal_matvec_f64_equals_csr(gpu, dims):
    print("thermal matvec f64", dims, flush=True)
    c = TCase(dims, seed=1)
    op = c.operator(gpu)
    u = np.random.default_rng(5).normal(size=c.mesh.num_nodes)
    y = op.matvec(torch.from_numpy(u).to(gpu)).cpu().numpy()
    assert rel_err(y, c.mask * (c.K @ (c.mask * u))) <= 1e-13


@pytest.mark.parametrize("dims", [(3, 2, 2), (31, 17, 5)])
def test_thermal_matvec_f32_close(gpu, dims):
    print("thermal matvec f32", dims, flush=True)
    c = TCase(dims, seed=2, h=(0.5, 1.25, 0.8), k0=3.0)
    op = c.operator(gpu)
    u = np.random.default_rng(6).normal(size=c.mesh.num_nodes)
    y = op.matvec(torch.from_numpy(u).to(gpu).float()).double().cpu().numpy()
    uf = u.astype(np.float32).astype(np.float64)
    assert rel_err(y, c.mask * (c.K @ (c.mask * uf))) <= 2e-6


def test_thermal_diagonal_and_lift(gpu):
    print("thermal diagonal and lift", flush=True)
    c = TCase((9, 7, 6), seed=3, penal=3.0, kvoid=1e-3)
    op = c.operator(gpu)
    d = op.diagonal().cpu().numpy()
    assert rel_err(d, c.mask * c.K.diagonal()) <= 1e-14
    g = np.zeros(c.mesh.num_nodes)
    g[c.fixed] = c.gv
    y = op.lift(torch.from_numpy(g).to(gpu)).cpu().numpy()
    assert rel_err(y, c.mask * (c.K @ g)) <= 1e-13


@pytest.mark.parametrize("dims,penal", [((12, 6, 6), 1.0), ((24, 12, 8), 3.0)])
def test_thermal_fea_solve_with_dirichlet(gpu, dims, penal):
    """HexThermalFEA with T0 = 20 and T1 = 100 on two faces and a heat load:
    temperatures vs the sparse direct solve of the oracle."""
    from pyto_amd import bound_cond, hex_mesher, linear_solvers, mat_lib
    from pyto_amd import topopt_material_model as tmm
    from pyto_amd.hex_thermal_fea import HexThermalFEA
    print("thermal fea", dims, penal, flush=True)
    tmm.initialize_SIMP_THERMAL_PENALTY(penal)
    tmm.set_KVOID_RELATIVE(1e-4)
    try:
        c = TCase(dims, seed=4, k0=50.0, penal=penal, kvoid=1e-4, lo=0.2)
        mesh = hex_mesher.HexMesher()
        mesh.grid_mesh(dims, (1.0, 1.0, 1.0))
        bc = bound_cond.BC(force=c.f, fixed_dofs=c.fixed, dirichlet_values=c.gv)
        mat = mat_lib.Material("T", 1.0, 0.3, 1.0, 50.0, 1.0, 1e-5)
        fe = HexThermalFEA(mesh, mat, bc, linear_solvers.Solvers.DPCG, rtol=1e-12)
        T = fe.solve(c.x, tmm.MaterialModel.SIMP)
        assert fe.last_stats["converged"]
        T_ref = c.solve_oracle()
        assert rel_err(T, T_ref) <= 1e-9
        assert np.array_equal(T[c.fixed], c.gv)
        fe.postprocess()
        assert rel_err(fe.strain, OF.thermal_gradient(c.mesh, T_ref)) <= 1e-9
    finally:
        tmm.initialize_SIMP_THERMAL_PENALTY(1.0)
        tmm.set_KVOID_RELATIVE(1e-8)


def test_thermal_bar_closed_form(gpu):
    """T = 0 on x = 0, T = 1 on x = L, no source: T = x / L."""
    from pyto_amd import bound_cond, hex_mesher, linear_solvers, mat_lib
    from pyto_amd.hex_thermal_fea import HexThermalFEA
    print("thermal bar", flush=True)
    mesh = hex_mesher.HexMesher()
    mesh.grid_mesh((32, 4, 4), (1.0, 1.0, 1.0))
    ix = OM.GridMesh((32, 4, 4), (1, 1, 1)).node_indices[:, 0]
    fixed = np.concatenate([np.nonzero(ix == 0)[0], np.nonzero(ix == 32)[0]])
    gv = np.concatenate([np.zeros((ix == 0).sum()), np.ones((ix == 32).sum())])
    bc = bound_cond.BC(force=np.zeros(ix.size), fixed_dofs=fixed, dirichlet_values=gv)
    fe = HexThermalFEA(mesh, mat_lib.get_material("Steel"), bc, linear_solvers.Solvers.PCG,
                       rtol=1e-12)
    T = fe.solve(np.full(32 * 16, 0.6), None)
    assert np.abs(T - ix / 32.0).max() <= 1e-10


def test_thermal_compliance_and_gradient(gpu):
    from pyto_amd import topopt_obj_cons_sensitivities as S
    from pyto_amd import topopt_material_model as tmm
    from pyto_amd.topopt_common import FEA_MODE
    print("thermal compliance", flush=True)
    c = TCase((16, 8, 6), seed=5, penal=1.0, kvoid=1e-8)
    op = c.operator(gpu)
    T = c.solve_oracle()

    class FE:
        stiff_mtrx = op

    C, g = S.compute_compliance_and_gradient(FEA_MODE.THERMAL, T, c.x, FE, c.KE,
                                             tmm.MaterialModel.SIMP)
    Te = T[c.mesh.elemArray]
    ce = np.einsum("ei,ij,ej->e", Te, c.KE, Te)
    C_ref = float((OF.thermal_scaling(c.x, 1.0, 1e-8) * ce).sum())
    g_ref = -OF.thermal_sensitivity(c.x, 1.0, 1e-8) * ce
    assert abs(C - C_ref) <= 1e-12 * abs(C_ref)
    assert rel_err(g, g_ref) <= 1e-12


def test_structural_nonzero_dirichlet(gpu):
    """Prescribed displacement of the clamped face (u_z = 0.01 y): the lifted
    GPU solve equals the oracle's b - A[:, fixed] g elimination."""
    from pyto_amd import bound_cond, linear_solvers
    print("structural dirichlet", flush=True)
    c = Case((12, 6, 4), seed=6)
    op = c.operator(gpu, groups=False)
    yy = c.mesh.node_xyz[c.fixed // 3, 1]
    gv = np.where(c.fixed % 3 == 2, 0.01 * yy, 0.0)
    bc = bound_cond.BC(force=c.f, fixed_dofs=c.fixed, dirichlet_values=gv)
    u = linear_solvers.solve(op, c.f, linear_solvers.Solvers.PCG, bc, rtol=1e-12, maxIters=20000)
    A, b, free = OF.impose_dirichlet(c.K, c.f, c.fixed, gv)
    u_ref = OF.scatter_solution(c.K.shape[0], c.fixed, gv, free, spla.spsolve(A.tocsc(), b))
    assert rel_err(u, u_ref) <= 1e-8


def _oracle_thermal_oc(c, H, Hs, volfrac, iters, force_elems):
    x = np.full(c.mesh.num_elems, volfrac)
    hist, scale = [], None
    for _ in range(iters):
        K = OF.assemble_KT(c.mesh, c.KE, OF.thermal_scaling(x, c.penal, c.kvoid))
        A, b, free = OF.impose_dirichlet(K, c.f, c.fixed, c.gv)
        T = OF.scatter_solution(c.mesh.num_nodes, c.fixed, c.gv, free,
                                spla.spsolve(A.tocsc(), b))
        Te = T[c.mesh.elemArray]
        ce = np.einsum("ei,ij,ej->e", Te, c.KE, Te)
        C = float((OF.thermal_scaling(x, c.penal, c.kvoid) * ce).sum())
        g = -OF.thermal_sensitivity(x, c.penal, c.kvoid) * ce
        scale = abs(C) if scale is None else scale
        g = OFL.sensitivity_filter(H, Hs, x, g / scale)
        g[force_elems] = g.min()
        x = OFL.oc_update(x, g, volfrac)
        hist.append(C)
    return x, hist


def test_thermal_oc_loop_matches_oracle(gpu):
    """Five OC iterations in the thermal mode (heat sink at T = 0 on x = 0,
    heat load on the far face) vs the oracle loop with sparse direct solves."""
    from pyto_amd import bound_cond, hex_mesher, linear_solvers, mat_lib
    from pyto_amd import topopt_common, topopt_ocm
    from pyto_amd import topopt_material_model as tmm
    from pyto_amd.hex_thermal_fea import HexThermalFEA
    print("thermal oc loop", flush=True)
    dims = (20, 10, 6)
    c = TCase(dims, x=np.full(np.prod(dims), 0.4), T0=0.0, T1=0.0)
    ix = c.mesh.node_indices[:, 0]
    c.fixed = np.nonzero(ix == 0)[0]
    c.gv = np.zeros(c.fixed.size)
    c.f = np.zeros(c.mesh.num_nodes)
    c.f[ix == dims[0]] = 0.01        # heat enters at the far face
    mesh = hex_mesher.HexMesher()
    mesh.grid_mesh(dims, (1.0, 1.0, 1.0))
    bc = bound_cond.BC(force=c.f, fixed_dofs=c.fixed, dirichlet_values=c.gv)
    mat = mat_lib.Material("T", 1.0, 0.3, 1.0, 1.0, 1.0, 0.0)
    tmm.initialize_SIMP_THERMAL_PENALTY(1.0)
    fe = HexThermalFEA(mesh, mat, bc, linear_solvers.Solvers.DPCG, rtol=1e-12)
    H, Hs = OFL.cone_filter(c.mesh, 1.5)
    fel = topopt_common.find_elements_with_forces(mesh, c.f, 1)
    x_ref, hist_ref = _oracle_thermal_oc(c, H, Hs, 0.4, 5, fel)
    params = topopt_common.TOParams()
    params.Constraints = [(topopt_common.TO_QOI.VOLUME_FRACTION, None, 0.4)]
    sol, hist, ok, msg, n = topopt_ocm.topopt_optimality_criteria(
        topopt_common.FEA_MODE.THERMAL, fe, params, maxIterations=5, print_progress=False,
        binarize_topology=False)
    assert n == 5
    assert np.allclose(hist["objective"][:5], hist_ref, rtol=1e-8, atol=0)
    assert np.abs(fe.stiff_mtrx.x.cpu().numpy() - x_ref).max() <= 1e-6


@pytest.mark.parametrize("dims", [(24, 12, 8), (64, 32, 32)])
def test_thermal_mixed_matches_fp64(gpu, dims):
    """The FP32-inside-FP64-refinement solve on the thermal operator (one
    temperature per node, Jacobi PCG) against the FP64 solve: thermal
    compliance within 1e-6, temperatures within 1e-5, both to rel res 1e-8."""
    from pyto_amd.operator import ThermalOperator
    print("thermal mixed", dims, flush=True)
    nelx, nely, nelz = dims
    sx, sy, sz = nelx + 1, nely + 1, nelz + 1
    idx = np.arange(sx * sy * sz)
    ix = idx % sx
    x = np.random.default_rng(4).uniform(0.01, 1.0, nelx * nely * nelz)
    op = ThermalOperator(dims, (1.0, 1.0, 1.0), 1.0, idx[ix == 0], penal=3.0, k_void=1e-6)
    op.set_density(x)
    q = torch.zeros(idx.size, dtype=torch.float64, device=gpu)
    q[torch.from_numpy(idx[ix == nelx]).to(gpu)] = 1.0 / (sy * sz)
    T64, s64 = op.solve(q, rtol=1e-8, max_iters=100000)
    Tm, sm = op.solve(q, rtol=1e-8, max_iters=100000, precision="mixed")
    print("fp64", s64["iterations"], "mixed", sm["iterations"], sm["outer_iterations"], flush=True)
    assert s64["converged"] and sm["converged"] and sm["rel"] <= 1e-8
    C64, Cm = float(torch.dot(q, T64)), float(torch.dot(q, Tm))
    assert abs(C64 - Cm) <= 1e-6 * abs(C64)
    assert rel_err(Tm.cpu().numpy(), T64.cpu().numpy()) <= 1e-5
    op.close()


def thermal_groups(c, boxes):
    """Box groups of the case's nodes (the reference's box rule on a unit
    grid, no merge) and the oracle W: one constant per group on the free
    nodes."""
    import scipy.sparse as sp
    m = c.mesh
    n = m.node_indices[:, :3].astype(np.int64)
    size = np.array(c.dims, dtype=np.int64) + 1
    b = np.array(boxes, dtype=np.int64)
    box = np.minimum(n * b // size, b - 1)
    ngn = (box[:, 0] + b[0] * (box[:, 1] + b[1] * box[:, 2])).astype(np.int32)
    G = int(b.prod())
    cen = np.zeros((G, 3))
    for i in range(3):
        cen[:, i] = np.bincount(ngn, weights=m.node_xyz[:, i], minlength=G) / np.bincount(ngn, minlength=G)
    free = np.setdiff1d(np.arange(m.num_nodes), c.fixed)
    W = sp.csr_matrix((np.ones(free.size), (np.arange(free.size), ngn[free])), shape=(free.size, G))
    return ngn, cen, W, free


@pytest.mark.parametrize("dims,boxes", [((12, 6, 4), (3, 2, 2)), ((17, 9, 6), (4, 2, 3))])
def test_thermal_coarse_space_matches_oracle(gpu, dims, boxes):
    """Thermal deflation groups (one constant per group): E = W^T K_T,ff W,
    W^T v and W mu against the oracle's assembled K_T and explicit W."""
    print("thermal coarse space", dims, boxes, flush=True)
    c = TCase(dims, seed=7, penal=3.0, kvoid=1e-6)
    ngn, cen, W, free = thermal_groups(c, boxes)
    op = c.operator(gpu)
    op.set_coarse_shift(0.0, 0.0)
    op.set_groups(ngn, cen)
    op.set_density(c.x)
    Kff = c.K.tocsr()[free][:, free]
    E_ref = (W.T @ (Kff @ W)).toarray()
    E = op.coarse_matrix().cpu().numpy()
    assert E.shape == E_ref.shape
    assert rel_err(E, E_ref) <= 1e-12
    v = np.random.default_rng(3).normal(size=c.mesh.num_nodes)
    t = op.restrict(torch.from_numpy(v).to(gpu)).cpu().numpy()
    assert rel_err(t, W.T @ v[free]) <= 1e-13
    mu = np.random.default_rng(4).normal(size=E.shape[0])
    pv = op.prolong(torch.from_numpy(mu).to(gpu)).cpu().numpy()
    ref = np.zeros(c.mesh.num_nodes)
    ref[free] = W @ mu
    assert np.array_equal(pv, ref)
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_thermal_deflated_solve(gpu, precision):
    """Deflated PCG on the thermal operator: the temperatures of the sparse
    direct solve, far fewer iterations than Jacobi PCG, and (FP64) the
    iteration count of the oracle's deflated PCG with the same W within 2."""
    import scipy.sparse.linalg as spla
    print("thermal deflated solve", precision, flush=True)
    dims = (32, 16, 12)
    c = TCase(dims, seed=8, penal=3.0, kvoid=1e-6, lo=0.01)
    c.gv[:] = 0.0    # homogeneous temperatures: the operator solves K_ff T_f = f_f
    ngn, cen, W, free = thermal_groups(c, (4, 2, 2))
    op = c.operator(gpu)
    op.set_coarse_shift(0.0, 0.0)
    f = torch.from_numpy(c.f).to(gpu)
    Tj, sj = op.solve(f, rtol=1e-10, max_iters=100000, precision=precision)
    op.set_groups(ngn, cen)
    op.set_density(c.x)
    T, st = op.solve(f, rtol=1e-10, max_iters=100000, precision=precision)
    print("jacobi", sj["iterations"], "deflated", st["iterations"], flush=True)
    assert st["converged"] and sj["converged"]
    A, b, _ = OF.impose_dirichlet(c.K, c.f, c.fixed, c.gv)
    T_ref = np.zeros(c.mesh.num_nodes)
    T_ref[free] = spla.spsolve(A.tocsc(), b)
    assert rel_err(T.cpu().numpy(), T_ref) <= 1e-7
    assert st["iterations"] < 0.6 * sj["iterations"]
    if precision == "fp64":
        from oracle import deflation as OD
        _, info = OD.deflated_pcg(A.tocsr(), b, W, OF.jacobi_inverse_diag(A), rtol=1e-10,
                                  maxIters=5000, shift=0.0)
        assert abs(st["iterations"] - info["iterations"]) <= 2
    op.close()


def test_thermal_fea_dpcg_uses_groups(gpu):
    """HexThermalFEA with Solvers.DPCG and a DeflationSolver runs the
    deflated solve (fewer iterations than PCG, same temperatures)."""
    from pyto_amd import bound_cond, deflation, hex_mesher, linear_solvers, mat_lib
    from pyto_amd.hex_thermal_fea import HexThermalFEA
    print("thermal fea dpcg", flush=True)
    dims = (24, 12, 8)
    c = TCase(dims, seed=9, k0=50.0, lo=0.2)
    mesh = hex_mesher.HexMesher()
    mesh.grid_mesh(dims, (1.0, 1.0, 1.0))
    bc = bound_cond.BC(force=c.f, fixed_dofs=c.fixed, dirichlet_values=c.gv)
    mat = mat_lib.Material("T", 1.0, 0.3, 1.0, 50.0, 1.0, 1e-5)
    ds = deflation.DeflationSolver()
    ds.create_box_groups(mesh, (4, 2, 2))
    fe_d = HexThermalFEA(mesh, mat, bc, linear_solvers.Solvers.DPCG, dsolver=ds, rtol=1e-12)
    Td = fe_d.solve(c.x, None)
    fe_p = HexThermalFEA(mesh, mat, bc, linear_solvers.Solvers.PCG, rtol=1e-12)
    Tp = fe_p.solve(c.x, None)
    print("dpcg", fe_d.last_stats["iterations"], "pcg", fe_p.last_stats["iterations"], flush=True)
    assert fe_d.last_stats["converged"]
    assert fe_d.last_stats["iterations"] < fe_p.last_stats["iterations"]
    assert rel_err(Td, Tp) <= 1e-9
    assert rel_err(Td, c.solve_oracle()) <= 1e-9
