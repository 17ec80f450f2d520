"""Pins of the oracle's thermal restatement (oracle/ke.py thermal parts,
oracle/fea.py conductivity law, K_T assembly and gradient)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import fea as OF
from oracle import ke as OK
from oracle import mesh as OM

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


@pytest.mark.parametrize("h", [(1, 1, 1), (0.5, 0.7, 1.3), (2.0, 1.0, 0.25)])
def test_thermal_gauss_equals_exact(h):
    A = OK.hex8_stiffness_thermal(2.5, h)
    B = OK.hex8_stiffness_thermal_exact(2.5, h)
    assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max()


def test_thermal_unit_cube_closed_form():
    """Trilinear brick, unit cube, k = 1: 1/3 on the diagonal, 0 between
    nodes that share an edge, -1/12 across a face or the body diagonal."""
    A = OK.hex8_stiffness_thermal(1.0, (1, 1, 1))
    for i in range(8):
        for j in range(8):
            nd = int(np.abs(CORNERS[i] - CORNERS[j]).sum())
            want = 1.0 / 3.0 if nd == 0 else (0.0 if nd == 1 else -1.0 / 12.0)
            assert abs(A[i, j] - want) < 1e-15


@pytest.mark.parametrize("h", [(1, 1, 1), (0.5, 0.7, 1.3)])
def test_thermal_invariants(h):
    A = OK.hex8_stiffness_thermal(1.7, h)
    assert np.allclose(A, A.T, atol=1e-15)
    assert np.abs(A.sum(axis=1)).max() < 1e-14          # constants are in the null space
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-13 and int((ev > 1e-10 * ev.max()).sum()) == 7


def test_thermal_energy_linear_field():
    """T = g.x: T^T KE_T T = k V |g|^2 for any constant gradient g."""
    rng = np.random.default_rng(4)
    h = (0.9, 1.1, 1.7)
    A = OK.hex8_stiffness_thermal(3.0, h)
    xyz = CORNERS * np.array(h)
    for _ in range(5):
        g = rng.normal(size=3)
        T = xyz @ g + rng.normal()
        assert abs(T @ A @ T - 3.0 * np.prod(h) * g @ g) < 1e-12 * abs(g @ g)


def test_patch_test_interior_nodes():
    """A linear temperature field on a grid has zero net flux at every
    interior node of the assembled K_T (uniform and graded conductivity
    along an axis orthogonal to the gradient)."""
    m = OM.GridMesh((4, 3, 3), (1.0, 0.5, 2.0))
    KE = OK.hex8_stiffness_thermal(1.0, m.elem_size)
    K = OF.assemble_KT(m, KE, np.ones(m.num_elems))
    T = m.node_xyz @ np.array([0.3, -1.2, 0.7]) + 5.0
    r = K @ T
    ix, iy, iz = m.node_indices[:, 0], m.node_indices[:, 1], m.node_indices[:, 2]
    inner = (ix > 0) & (ix < 4) & (iy > 0) & (iy < 3) & (iz > 0) & (iz < 3)
    assert np.abs(r[inner]).max() < 1e-13


def test_bar_conduction_closed_form():
    """Bar along x, T = 0 on x = 0 and T = 1 on x = L, no source: T = x / L
    exactly (the linear field is in the trilinear space)."""
    m = OM.GridMesh((8, 2, 2), (1.0, 1.0, 1.0))
    KE = OK.hex8_stiffness_thermal(2.0, m.elem_size)
    K = OF.assemble_KT(m, KE, np.full(m.num_elems, 0.7))
    ix = m.node_indices[:, 0]
    fixed = np.concatenate([np.where(ix == 0)[0], np.where(ix == 8)[0]])
    g = np.concatenate([np.zeros((ix == 0).sum()), np.ones((ix == 8).sum())])
    A, b, free = OF.impose_dirichlet(K, np.zeros(m.num_nodes), fixed, g)
    T = OF.scatter_solution(m.num_nodes, fixed, g, free, spla.spsolve(A.tocsc(), b))
    assert np.abs(T - ix / 8.0).max() < 1e-12


def test_thermal_compliance_gradient_fd():
    """dC/dx = -k'(x) T_e^T KE_T T_e (zero Dirichlet values) against
    central differences of C = f^T T."""
    m = OM.GridMesh((3, 2, 2), (1.0, 1.0, 1.0))
    KE = OK.hex8_stiffness_thermal(1.0, m.elem_size)
    rng = np.random.default_rng(7)
    x = rng.uniform(0.2, 1.0, m.num_elems)
    ix = m.node_indices[:, 0]
    fixed = np.where(ix == 0)[0]
    f = np.zeros(m.num_nodes)
    f[np.where(ix == 3)[0]] = 1.0
    penal, kv = 2.0, 1e-3

    def solve(xx):
        K = OF.assemble_KT(m, KE, OF.thermal_scaling(xx, penal, kv))
        A, b, free = OF.impose_dirichlet(K, f, fixed, np.zeros(fixed.size))
        return OF.scatter_solution(m.num_nodes, fixed, np.zeros(fixed.size), free,
                                   spla.spsolve(A.tocsc(), b))

    T = solve(x)
    Te = T[m.elemArray]
    ce = np.einsum("ei,ij,ej->e", Te, KE, Te)
    grad = -OF.thermal_sensitivity(x, penal, kv) * ce
    for e in (0, 5, 11):
        hh = 1e-6
        xp, xm = x.copy(), x.copy()
        xp[e] += hh
        xm[e] -= hh
        fd = (f @ solve(xp) - f @ solve(xm)) / (2 * hh)
        assert abs(fd - grad[e]) <= 1e-6 * abs(grad[e])


def test_thermal_gradient_of_linear_field():
    m = OM.GridMesh((3, 2, 2), (1.0, 1.0, 1.0))
    T = m.node_xyz @ np.array([0.5, -2.0, 1.5])
    G = OF.thermal_gradient(m, T)
    # reference coordinates span 2 per element: (1/8) * 4 * (difference) = g/2
    assert np.allclose(G, np.array([[0.25], [-1.0], [0.75]]) * np.ones((1, m.num_elems)))


@pytest.mark.parametrize("h", [(1, 1, 1), (0.6, 1.3, 0.8)])
@pytest.mark.parametrize("nu", [0.0, 0.3])
def test_thermoelastic_H_free_expansion(h, nu):
    """Physical layout: a free element heated by dT expands as u = alpha dT x,
    a stress-free state, so KE u equals the thermal load E alpha dT H 1."""
    E, alpha, dT = 2.0, 1.5e-2, 3.0
    KE = OK.hex8_stiffness_exact(E, nu, h)
    H = OK.thermoelastic_H(h, nu, layout="physical")
    xyz = CORNERS * np.array(h)
    u = (alpha * dT * xyz).ravel()
    assert np.abs(KE @ u - E * alpha * dT * H @ np.ones(8)).max() < 1e-13


def test_thermoelastic_H_reference_layout():
    """The reference's getHMatrix (src/hex_thermal_fea.py:119-196) puts the
    x, y, z rows of the 8 nodes in blocks of 8; its first entry is
    dy dz / (18 (2 nu - 1))."""
    h, nu = (0.6, 1.3, 0.8), 0.3
    Hp = OK.thermoelastic_H(h, nu, layout="physical")
    Hr = OK.thermoelastic_H(h, nu, layout="reference")
    for j in range(8):
        for c in range(3):
            assert np.array_equal(Hr[8 * c + j], Hp[3 * j + c])
    assert abs(Hr[0, 0] - h[1] * h[2] / (18.0 * (2 * nu - 1))) < 1e-15
