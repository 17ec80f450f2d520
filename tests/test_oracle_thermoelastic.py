"""Pins of the oracle's thermo-elastic coupling (oracle/sensitivities.py,
oracle/ke.py thermoelastic_H)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import fea as OF
from oracle import ke as OK
from oracle import mesh as OM
from oracle import sensitivities as OS


def test_force_of_free_uniform_heating_is_K_times_expansion():
    """A free body heated uniformly by dT expands as u = alpha dT x without
    stress, so the assembled thermal force equals K u (physical layout)."""
    m = OM.GridMesh((3, 2, 2), (0.7, 1.0, 1.3))
    m.createEdofMatStructural()
    E, nu, alpha, dT, t_ref = 2.0, 0.3, 1e-2, 5.0, 23.0
    KE = OK.hex8_stiffness(E, nu, m.elem_size)
    K = OF.assemble_K(m, KE, np.ones(m.num_elems))
    H = OK.thermoelastic_H(m.elem_size, nu, layout="physical")
    T = np.full(m.num_nodes, t_ref + dT)
    f = OS.thermoelastic_force(m, T, np.ones(m.num_elems), H, E * alpha, t_ref, 3.0, 0.0)
    u = (alpha * dT * m.node_xyz).ravel()
    assert np.abs(K @ u - f).max() < 1e-13 * np.abs(f).max()


class Coupled:
    """Clamped at x = 0 (all components), heat sink T = t_ref on x = 0,
    heat load on the x = max face; J = d^T K d with K d = f + f_te(T)."""

    def __init__(self, dims, layout, seed=3):
        m = OM.GridMesh(dims, (1.0, 1.0, 1.0))
        m.createEdofMatStructural()
        self.m = m
        self.E, self.nu, self.alpha, self.k0, self.t_ref = 1.0, 0.3, 1e-2, 1.0, 23.0
        self.KE = OK.hex8_stiffness(self.E, self.nu, m.elem_size)
        self.KT = OK.hex8_stiffness_thermal(self.k0, m.elem_size)
        self.H = OK.thermoelastic_H(m.elem_size, self.nu, layout=layout)
        ix = m.node_indices[:, 0]
        self.fixT = np.nonzero(ix == 0)[0]
        self.gT = np.full(self.fixT.size, self.t_ref)
        self.q = np.zeros(m.num_nodes)
        self.q[ix == dims[0]] = 0.5
        self.fixS = np.concatenate([3 * self.fixT + c for c in range(3)])
        self.fS = np.zeros(3 * m.num_nodes)
        self.fS[3 * np.nonzero(ix == dims[0])[0] + 2] = -0.01
        self.x = np.random.default_rng(seed).uniform(0.3, 1.0, m.num_elems)
        self.ps, self.es, self.pt, self.kv = 3.0, 1e-3, 1.0, 1e-3

    def temperature(self, x):
        K = OF.assemble_KT(self.m, self.KT, OF.thermal_scaling(x, self.pt, self.kv))
        A, b, free = OF.impose_dirichlet(K, self.q, self.fixT, self.gT)
        return OF.scatter_solution(self.m.num_nodes, self.fixT, self.gT, free,
                                   spla.spsolve(A.tocsc(), b)), K

    def displacement(self, x, T):
        K = OF.assemble_K(self.m, self.KE, OF.simp_scaling(x, self.ps, self.es))
        fte = OS.thermoelastic_force(self.m, T, x, self.H, self.E * self.alpha, self.t_ref,
                                     self.ps, self.es)
        z = np.zeros(self.fixS.size)
        A, b, free = OF.impose_dirichlet(K, self.fS + fte, self.fixS, z)
        return OF.scatter_solution(K.shape[0], self.fixS, z, free, spla.spsolve(A.tocsc(), b)), K

    def J(self, x):
        T, _ = self.temperature(x)
        d, K = self.displacement(x, T)
        return float(d @ (K @ d))

    def gradient(self, x, exact):
        T, KTa = self.temperature(x)
        d, _ = self.displacement(x, T)
        rhs = OS.thermal_adjoint_rhs(self.m, d, x, self.H, -2 * self.E * self.alpha, self.ps,
                                     self.es)
        z = np.zeros(self.fixT.size)
        A, b, free = OF.impose_dirichlet(KTa, rhs, self.fixT, z)
        lam = OF.scatter_solution(self.m.num_nodes, self.fixT, z, free,
                                  spla.spsolve(A.tocsc(), b))
        return OS.thermoelastic_gradient(self.m, x, T, d, lam, self.KE, self.KT, self.H,
                                         self.E * self.alpha, self.t_ref, self.ps, self.es,
                                         self.pt, self.kv, exact=exact), T, d, lam


@pytest.mark.parametrize("layout", ["physical", "reference"])
def test_exact_coupled_gradient_matches_fd(layout):
    c = Coupled((6, 3, 2), layout)
    g, _, _, _ = c.gradient(c.x, exact=True)
    for e in (0, 7, 20, 35):
        h = 1e-6
        xp, xm = c.x.copy(), c.x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (c.J(xp) - c.J(xm)) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * np.abs(g).max(), (e, fd, g[e])


def test_reference_gradient_differs_only_in_the_direct_law():
    c = Coupled((6, 3, 2), "reference")
    g_ref, T, d, _ = c.gradient(c.x, exact=False)
    g_ex, _, _, _ = c.gradient(c.x, exact=True)
    bil = np.einsum("ei,ik,ek->e", d[c.m.edofMatStructural], c.H, T[c.m.elemArray] - c.t_ref)
    want = 2 * c.E * c.alpha * bil * (OF.thermal_sensitivity(c.x, c.pt, c.kv)
                                      - OF.simp_sensitivity(c.x, c.ps, c.es))
    assert np.abs((g_ref - g_ex) - want).max() <= 1e-12 * np.abs(g_ref).max()
