"""Pins of the oracle's MMA restatement (oracle/mma.py) by what the
mathematics fixes: the approximation's value and gradient at the expansion
point, an independent dual solve of the subproblem, the subproblem's KKT
conditions, a published optimum and an SLSQP optimum."""
import numpy as np
import pytest
import scipy.optimize as so

from oracle import mma as OM


def _random_problem(rng, n, m, it=1):
    xval = rng.uniform(0.1, 0.9, n)
    xmin, xmax = np.full(n, 1e-6), np.full(n, 1 - 1e-6)
    xold1 = np.clip(xval + rng.normal(0, 0.05, n), 0.01, 0.99)
    xold2 = np.clip(xold1 + rng.normal(0, 0.05, n), 0.01, 0.99)
    low, upp = xmin.copy(), xmax.copy()
    if it > 2:
        low, upp = xval - 0.3, xval + 0.4
    df0dx = rng.normal(size=n)
    dfdx = rng.normal(size=(m, n)) / n
    fval = rng.normal(0, 0.1, m)
    return dict(xval=xval, xmin=xmin, xmax=xmax, xold1=xold1, xold2=xold2, low=low, upp=upp,
                df0dx=df0dx, dfdx=dfdx, fval=fval)


@pytest.mark.parametrize("it", [1, 5])
def test_approximation_matches_value_and_gradient(it):
    rng = np.random.default_rng(11)
    n, m = 40, 2
    P_ = _random_problem(rng, n, m, it)
    low, upp = OM.asymptotes(it, P_["xval"], P_["xmin"], P_["xmax"], P_["xold1"], P_["xold2"],
                             P_["low"], P_["upp"])
    alfa, beta, p0, q0, P, Q, b = OM.approximation(P_["xval"], P_["xmin"], P_["xmax"], low, upp,
                                                   P_["df0dx"], P_["fval"], P_["dfdx"], 0.2)
    x = P_["xval"]
    _, fi = OM.approx_value(x, low, upp, p0, q0, P, Q, b)
    assert np.abs(fi - P_["fval"]).max() < 1e-14
    g0 = p0 / (upp - x) ** 2 - q0 / (x - low) ** 2
    gi = P / (upp - x) ** 2 - Q / (x - low) ** 2
    assert np.abs(g0 - P_["df0dx"]).max() < 1e-12
    assert np.abs(gi - P_["dfdx"]).max() < 1e-14
    assert np.all(alfa > low) and np.all(beta < upp) and np.all(alfa <= x) and np.all(x <= beta)
    assert np.all(beta - alfa <= 0.4 * (P_["xmax"] - P_["xmin"]) + 1e-15)


def _dual_solution(low, upp, alfa, beta, p0, q0, P, Q, b, c, d):
    """Maximise the subproblem's dual (a0 = 1, a = 0) over lambda >= 0 with the
    closed-form minimisers x(lambda), y(lambda), z = 0."""
    def xy(lam):
        pl = np.sqrt(p0 + P.T @ lam)
        ql = np.sqrt(q0 + Q.T @ lam)
        x = np.clip((pl * low + ql * upp) / (pl + ql), alfa, beta)
        y = np.maximum(0.0, (lam - c) / d)
        return x, y

    def negW(lam):
        x, y = xy(lam)
        g = P @ (1 / (upp - x)) + Q @ (1 / (x - low))
        W = np.sum(p0 / (upp - x) + q0 / (x - low)) + np.sum(c * y + 0.5 * d * y * y) \
            + lam @ (g - y - b)
        return -W, -(g - y - b)

    m = b.size
    r = so.minimize(negW, np.ones(m), jac=True, method="L-BFGS-B", bounds=[(0, None)] * m,
                    options=dict(ftol=1e-15, gtol=1e-13, maxiter=5000))
    x, y = xy(r.x)
    return x, y, r.x


@pytest.mark.parametrize("m", [1, 2])
def test_subsolv_matches_dual_solve_and_kkt(m):
    rng = np.random.default_rng(5 + m)
    n = 60
    P_ = _random_problem(rng, n, m)
    # make every constraint active: the approximation at xval is violated
    P_["fval"] = np.full(m, 0.05)
    low, upp = OM.asymptotes(1, P_["xval"], P_["xmin"], P_["xmax"], P_["xold1"], P_["xold2"],
                             P_["low"], P_["upp"])
    alfa, beta, p0, q0, P, Q, b = OM.approximation(P_["xval"], P_["xmin"], P_["xmax"], low, upp,
                                                   P_["df0dx"], P_["fval"], P_["dfdx"], 0.2)
    c, d = 100.0 * np.ones(m), np.ones(m)
    v = OM.subsolv(m, n, OM.EPSIMIN, low, upp, alfa, beta, p0, q0, P, Q, 1.0, np.zeros(m), b,
                   c, d)
    xd, yd, lamd = _dual_solution(low, upp, alfa, beta, p0, q0, P, Q, b, c, d)
    # a bound whose multiplier is nearly zero leaves the barrier solution
    # epsi / eta inside it, up to ~2e-5 here
    assert np.abs(v.x - xd).max() < 1e-4
    assert np.abs(v.lam - lamd).max() < 1e-5 * max(1.0, np.abs(lamd).max())
    assert np.abs(v.y - yd).max() < 1e-5
    # KKT of the subproblem at epsi = 0: stationarity and complementarity
    r0, _ = OM._residual(v, 0.0, low, upp, alfa, beta, p0, q0, P, Q, 1.0, np.zeros(m), b, c, d)
    assert r0 < 1e-5
    # the interior point respects its bounds strictly
    assert np.all(v.x > alfa) and np.all(v.x < beta) and np.all(v.lam > 0)


def test_cantilever_reaches_published_optimum():
    """Svanberg (1987), five-segment cantilever: min 0.0624 sum x subject to
    61/x1^3 + 37/x2^3 + 19/x3^3 + 7/x4^3 + 1/x5^3 <= 1; optimum f = 1.340
    at x = (6.016, 5.309, 4.494, 3.502, 2.153)."""
    cc = np.array([61.0, 37.0, 19.0, 7.0, 1.0])

    def func(x):
        return (0.0624 * x.sum(), np.full(5, 0.0624), np.array([cc @ x ** -3 - 1.0]),
                (-3 * cc * x ** -4)[None, :])

    x, f, _, g, _, it, _ = OM.run_mma(5, 1, func, np.full(5, 5.0), np.full(5, 1.0),
                                      np.full(5, 10.0), fTolerance=1e-9, gTolerance=1e-9,
                                      maxIterations=200, move_limit=0.5, kktTol=1e-9)
    f_abs = 0.0624 * x.sum()
    assert abs(f_abs - 1.340) < 5e-4
    assert np.abs(x - np.array([6.016, 5.309, 4.494, 3.502, 2.153])).max() < 2e-3
    assert g[0] < 1e-6


def test_toy_problem_matches_slsqp():
    """min x1^2+x2^2+x3^2 subject to two spheres of radius 3, 0 <= x <= 5."""
    c1, c2 = np.array([5.0, 2.0, 1.0]), np.array([3.0, 4.0, 3.0])

    def func(x):
        g = np.array([np.sum((x - c1) ** 2) - 9.0, np.sum((x - c2) ** 2) - 9.0])
        return x @ x, 2 * x, g, np.stack([2 * (x - c1), 2 * (x - c2)])

    x, *_ = OM.run_mma(3, 2, func, np.array([4.0, 3.0, 2.0]), np.zeros(3), np.full(3, 5.0),
                       fTolerance=1e-12, gTolerance=1e-9, maxIterations=300, move_limit=0.5,
                       kktTol=1e-10)
    cons = [{"type": "ineq", "fun": lambda x, cc=cc: 9.0 - np.sum((x - cc) ** 2)} for cc in (c1, c2)]
    r = so.minimize(lambda x: x @ x, np.array([4.0, 3.0, 2.0]), jac=lambda x: 2 * x,
                    method="SLSQP", bounds=[(0, 5)] * 3, constraints=cons,
                    options=dict(ftol=1e-11, maxiter=500))
    assert r.success
    assert np.abs(x - r.x).max() < 1e-4


def test_kktcheck_is_zero_at_an_exact_kkt_point():
    """One variable, min x subject to nothing active, lower bound 0.2: the
    multiplier of the bound is 1, everything else vanishes."""
    r, nrm, mx = OM.kktcheck(1, 1, np.array([0.2]), np.zeros(1), 0.0, np.zeros(1),
                             np.array([1.0]), np.zeros(1), np.array([100.0]), 1.0, np.ones(1),
                             np.array([0.2]), np.array([1.0]), np.array([1.0]),
                             np.array([-1.0]), np.zeros((1, 1)), 1.0, np.zeros(1),
                             100.0 * np.ones(1), np.ones(1))
    assert nrm < 1e-15 and mx < 1e-15
