"""Pins of the oracle's element stiffness (oracle/ke.py)."""
import numpy as np
import pytest

from oracle import ke


@pytest.mark.parametrize("h", [(1, 1, 1), (0.5, 0.7, 1.3), (2.0, 1.0, 0.25)])
@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_gauss_equals_exact_integration(h, nu):
    A = ke.hex8_stiffness(2.5, nu, h)
    B = ke.hex8_stiffness_exact(2.5, nu, h)
    assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max()


@pytest.mark.parametrize("h", [(1, 1, 1), (0.5, 0.7, 1.3)])
def test_invariants(h):
    A = ke.hex8_stiffness(1.0, 0.3, h)
    assert np.allclose(A, A.T, atol=1e-15 * np.abs(A).max())
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-12 * ev.max()
    assert int((ev > 1e-10 * ev.max()).sum()) == 18
    R = ke.rigid_body_modes(h)
    assert np.abs(A @ R).max() <= 1e-13 * np.abs(A).max()


def test_patch_energy_linear_field():
    """u = G x (any constant gradient): u^T KE u = V eps^T D eps."""
    rng = np.random.default_rng(3)
    h = (0.9, 1.1, 1.7)
    E, nu = 3.0, 0.27
    A = ke.hex8_stiffness(E, nu, h)
    D = ke.isotropic_D(E, nu)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float) * np.array(h)
    for _ in range(5):
        G = rng.normal(size=(3, 3))
        u = (xyz @ G.T).ravel()
        eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1],
                        G[0, 2] + G[2, 0]])
        V = np.prod(h)
        assert np.isclose(u @ A @ u, V * eps @ D @ eps, rtol=1e-12)


def test_unit_cube_known_diagonal():
    """Diagonal entry of the unit cube: (lam+2mu)/9 + 2 mu/9 (closed form of
    int (dN/dx)^2 = 1/9 per axis)."""
    E, nu = 1.0, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    A = ke.hex8_stiffness(E, nu, (1, 1, 1))
    assert np.allclose(np.diag(A), (lam + 2 * mu) / 9 + 2 * mu / 9, rtol=1e-14)
