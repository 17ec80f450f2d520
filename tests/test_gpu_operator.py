"""GPU parity of the matrix-free operator, Jacobi diagonal and coarse space
against the oracle's assembled CSR K and explicit W (FP64 to rounding)."""
import numpy as np
import pytest
import torch

from tests.helpers import Case, rel_err

pytestmark = pytest.mark.gpu

GRIDS = [(1, 1, 1), (3, 2, 2), (16, 15, 3), (31, 17, 5), (15, 30, 18), (60, 20, 10)]


@pytest.mark.parametrize("dims", GRIDS)
def test_matvec_f64_equals_csr(gpu, dims):
    print("matvec f64", dims, flush=True)
    c = Case(dims, seed=1)
    op = c.operator(gpu, groups=False)
    rng = np.random.default_rng(5)
    for _ in range(2):
        u = rng.normal(size=3 * c.mesh.num_nodes)
        y = op.matvec(torch.from_numpy(u).to(gpu)).cpu().numpy()
        assert rel_err(y, c.Kff_full(u)) <= 1e-13


@pytest.mark.parametrize("dims", [(3, 2, 2), (31, 17, 5), (60, 20, 10)])
def test_matvec_f32_close(gpu, dims):
    print("matvec f32", dims, flush=True)
    c = Case(dims, seed=2)
    op = c.operator(gpu, groups=False)
    u = np.random.default_rng(6).normal(size=3 * c.mesh.num_nodes)
    y = op.matvec(torch.from_numpy(u).to(gpu).float()).double().cpu().numpy()
    assert rel_err(y, c.Kff_full(u.astype(np.float32).astype(np.float64))) <= 2e-6


@pytest.mark.parametrize("dims,h", [((3, 2, 2), (1, 1, 1)), ((31, 17, 5), (1, 1, 1)),
                                    ((60, 20, 10), (1, 1, 1)), ((9, 7, 6), (0.5, 1.25, 0.8))])
def test_matvec_f32_mfma_close(gpu, dims, h):
    """The dense-KE matrix-core variant of the FP32 operator against the
    assembled CSR K (FP32 tolerance) and against the modal FP32 operator."""
    print("matvec f32 mfma", dims, h, flush=True)
    c = Case(dims, seed=2, h=h)
    op = c.operator(gpu, groups=False)
    u = np.random.default_rng(8).normal(size=3 * c.mesh.num_nodes)
    u32 = torch.from_numpy(u).to(gpu).float()
    y = op.matvec_mfma(u32).double().cpu().numpy()
    assert rel_err(y, c.Kff_full(u.astype(np.float32).astype(np.float64))) <= 2e-6
    assert rel_err(y, op.matvec(u32).double().cpu().numpy()) <= 2e-6


def test_anisotropic_brick_and_material(gpu):
    print("matvec anisotropic", flush=True)
    c = Case((9, 7, 6), seed=3, h=(0.5, 1.25, 0.8), E0=7.0, nu=0.41)
    op = c.operator(gpu, groups=False)
    u = np.random.default_rng(7).normal(size=3 * c.mesh.num_nodes)
    y = op.matvec(torch.from_numpy(u).to(gpu)).cpu().numpy()
    assert rel_err(y, c.Kff_full(u)) <= 1e-13


@pytest.mark.parametrize("dims", [(3, 2, 2), (60, 20, 10)])
def test_diagonal(gpu, dims):
    print("diagonal", dims, flush=True)
    c = Case(dims, seed=4)
    op = c.operator(gpu, groups=False)
    d = op.diagonal().cpu().numpy()
    assert rel_err(d, c.mask * c.K.diagonal()) <= 1e-14


def test_symmetry_at_full_size(gpu):
    """u^T K v = v^T K u on the 192x64x64 grid (2.4M DOF)."""
    from pyto_amd.operator import GridOperator
    from pyto_amd import problems
    print("symmetry 192x64x64", flush=True)
    mesh, mat, bc = problems.cantilever((192, 64, 64))
    op = GridOperator(mesh.grid, mesh.elem_size, 1.0, 0.3, bc.fixed_dofs)
    op.set_density(problems.density_uniform(mesh.num_elems, 0.01, 1.0, 0))
    g = torch.Generator(device="cuda").manual_seed(0)
    u = torch.randn(op.num_dofs, dtype=torch.float64, device=gpu, generator=g)
    v = torch.randn(op.num_dofs, dtype=torch.float64, device=gpu, generator=g)
    a = float(torch.dot(u, op.matvec(v)))
    b = float(torch.dot(v, op.matvec(u)))
    assert abs(a - b) <= 1e-12 * abs(a)
    # positive on a random vector and zero on fixed DOFs
    Ku = op.matvec(u)
    assert float(torch.dot(u * (op.diagonal() != 0), Ku)) > 0
    fixed = torch.from_numpy(np.asarray(bc.fixed_dofs)).to(gpu)
    assert float(Ku[fixed].abs().max()) == 0.0


@pytest.mark.parametrize("dims,h", [((12, 6, 4), (1.0, 1.0, 1.0)), ((60, 20, 10), (1.0, 1.0, 1.0)),
                                    ((14, 7, 5), (0.5, 1.0, 1.5))])
def test_coarse_space(gpu, dims, h):
    print("coarse", dims, h, flush=True)
    c = Case(dims, seed=5, h=h)
    op = c.operator(gpu, groups=True)
    W = c.W_free()
    rng = np.random.default_rng(8)
    v = rng.normal(size=3 * c.mesh.num_nodes)
    t = op.restrict(torch.from_numpy(v).to(gpu)).cpu().numpy()
    assert rel_err(t, W.T @ v) <= 1e-13
    mu = rng.normal(size=W.shape[1])
    pv = op.prolong(torch.from_numpy(mu).to(gpu)).cpu().numpy()
    assert rel_err(pv, W @ mu) <= 1e-14
    Kf = c.K.multiply(1.0)
    E_ref = (W.T @ (Kf @ W)).toarray()
    E_ref = 0.5 * (E_ref + E_ref.T) + 1e-10 * np.eye(E_ref.shape[0])
    E = op.coarse_matrix().cpu().numpy()
    assert rel_err(E, E_ref) <= 1e-12
    rhs = rng.normal(size=W.shape[1])
    sol = op.coarse_solve(torch.from_numpy(rhs).to(gpu)).cpu().numpy()
    ref = np.linalg.solve(E_ref, rhs)
    assert np.linalg.norm(E_ref @ sol - rhs) <= 1e-10 * np.linalg.norm(rhs)
    assert rel_err(sol, ref) <= 1e-8


def test_coarse_matrix_is_reproducible(gpu):
    """The Galerkin matrix is summed in a fixed order: the same densities give
    the same bits on every assembly."""
    print("coarse reproducible", flush=True)
    c = Case((40, 20, 12), seed=6)
    op = c.operator(gpu, groups=True)
    E1 = op.coarse_matrix().clone()
    for _ in range(3):
        op.set_density(c.x)
        assert torch.equal(op.coarse_matrix(), E1)


def test_bad_calls_raise(gpu):
    from pyto_amd.operator import GridOperator
    print("bad calls", flush=True)
    op = GridOperator((2, 2, 2), (1, 1, 1), 1.0, 0.3, [0, 1, 2])
    u = torch.zeros(op.num_dofs, dtype=torch.float64, device=gpu)
    with pytest.raises(RuntimeError):
        op.matvec(u)                      # density not set
    op.set_density(np.ones(8))
    with pytest.raises(ValueError):
        op.matvec(torch.zeros(5, dtype=torch.float64, device=gpu))
    with pytest.raises(ValueError):
        op.set_groups(np.full(op.num_nodes, 3, np.int32), np.zeros((2, 3)))
    with pytest.raises(ValueError):
        GridOperator((0, 2, 2), (1, 1, 1), 1.0, 0.3, [])


@pytest.mark.parametrize("dims,h", [((31, 17, 5), (1, 1, 1)), ((16, 45, 7), (1, 1, 1)),
                                    ((9, 23, 6), (0.5, 1.25, 0.8))])
def test_tile_rows_same_result(gpu, dims, h):
    """Every tile shape (16 x 8, 16 x 16, 16 x 32 element columns) gives the
    same bits: each node is summed by one tile in a fixed order.  Grids whose
    y extent is not a multiple of any tile's owned rows, with fixed nodes on
    tile seams; FP64 also against the assembled CSR K."""
    print("tile rows", dims, h, flush=True)
    c = Case(dims, seed=9, h=h)
    op = c.operator(gpu, groups=False)
    assert op.tile_rows == 16
    rng = np.random.default_rng(10)
    u = torch.from_numpy(rng.normal(size=3 * c.mesh.num_nodes)).to(gpu)
    ref64, ref32, reflift = op.matvec(u), op.matvec(u.float()), op.lift(u)
    refm = op.matvec_mfma(u.float()) if h == (1, 1, 1) else None
    for rows in (8, 32, 16):
        op.set_tile_rows(rows)
        assert op.tile_rows == rows
        y64 = op.matvec(u)
        assert torch.equal(y64, ref64)
        assert rel_err(y64.cpu().numpy(), c.Kff_full(u.cpu().numpy())) <= 1e-13
        assert torch.equal(op.matvec(u.float()), ref32)
        assert torch.equal(op.lift(u), reflift)
        if refm is not None:
            assert torch.equal(op.matvec_mfma(u.float()), refm)


def test_tile_rows_thermal_and_solve(gpu):
    """Tile shapes on the thermal operator, and a deflated mixed solve after a
    shape change (the solver launches with the new shape)."""
    from pyto_amd.operator import ThermalOperator
    print("tile rows thermal + solve", flush=True)
    dims = (20, 33, 6)
    idx = np.arange((dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1))
    op = ThermalOperator(dims, (1.0, 1.0, 1.0), 1.0, idx[idx % (dims[0] + 1) == 0], penal=3.0,
                         k_void=1e-6)
    op.set_density(np.random.default_rng(11).uniform(0.05, 1.0, op.num_elems))
    T = torch.from_numpy(np.random.default_rng(12).normal(size=op.num_nodes)).to(gpu)
    ref = op.matvec(T)
    for rows in (8, 32):
        op.set_tile_rows(rows)
        assert torch.equal(op.matvec(T), ref)
        assert torch.equal(op.matvec(T.float()), op.matvec(T.float()))

    c = Case((24, 40, 8), seed=13)
    s = c.operator(gpu, groups=True)
    f = torch.from_numpy(c.f).to(gpu)
    u16, info16 = s.solve(f, rtol=1e-8, precision="mixed")
    for rows in (32, 8):
        s.set_tile_rows(rows)
        u, info = s.solve(f, rtol=1e-8, precision="mixed")
        assert info["converged"]
        assert abs(info["iterations"] - info16["iterations"]) <= 3
        assert rel_err(u.cpu().numpy(), u16.cpu().numpy()) <= 1e-7
    with pytest.raises(ValueError):
        s.set_tile_rows(12)
