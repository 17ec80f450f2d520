"""Method of moving asymptotes (oracle; test infrastructure only).

The reference drives MMA through ``runMMA`` (src/mmaWrapper.py:18-134), which
calls ``mmasub`` and ``kktcheck`` of the third-party package mmapy
(GCMMA-MMA-Python, requirements.txt:25, unpinned; not installed here and not
under the reference tree).  mmapy is a port of K. Svanberg's MATLAB
mmasub.m / subsolv.m / kktcheck.m (version September 2007, described in
"MMA and GCMMA - two methods for nonlinear optimization", Svanberg 2007).
This module restates that published algorithm with its published constants
(asyinit 0.5, asyincr 1.2, asydecr 0.7, asymptote bounds 0.01 and 10 times
xmax - xmin, raa0 1e-5, albefa 0.1, epsimin 1e-7, at most 200 Newton steps per
barrier level and 50 step halvings).

Vectors are 1-D: n-vectors (x, bounds, gradients), m-vectors (constraints);
``dfdx`` is (m, n).

Pinning (tests/test_oracle_mma.py): the convex approximation reproduces f and
its gradient at xval; the subproblem solution matches an independent dual
solve (closed-form x(lambda), maximised dual) and satisfies the subproblem's
KKT conditions; the outer loop reaches Svanberg's published optimum of the
five-segment cantilever (f = 1.340) and the SLSQP optimum of a two-constraint
toy problem.  No golden vector of mmapy exists in the reference, so identity
with mmapy's own iterates, beyond these pins of the published algorithm, is
not checked.
"""
import numpy as np

EPSIMIN = 1e-7
RAA0 = 1e-5
ALBEFA = 0.1
ASYINIT = 0.5
ASYINCR = 1.2
ASYDECR = 0.7


def asymptotes(it, xval, xmin, xmax, xold1, xold2, low, upp):
    """mmasub's asymptote update (Svanberg 2007, eqs. for L and U)."""
    span = xmax - xmin
    if it <= 2:
        return xval - ASYINIT * span, xval + ASYINIT * span
    zzz = (xval - xold1) * (xold1 - xold2)
    factor = np.ones_like(xval)
    factor[zzz > 0] = ASYINCR
    factor[zzz < 0] = ASYDECR
    low = xval - factor * (xold1 - low)
    upp = xval + factor * (upp - xold1)
    low = np.minimum(np.maximum(low, xval - 10 * span), xval - 0.01 * span)
    upp = np.maximum(np.minimum(upp, xval + 10 * span), xval + 0.01 * span)
    return low, upp


def approximation(xval, xmin, xmax, low, upp, df0dx, fval, dfdx, move):
    """Move limits alfa, beta and the coefficients p0, q0, P, Q, b of the
    convex separable approximation (mmasub after the asymptotes)."""
    alfa = np.maximum(np.maximum(low + ALBEFA * (xval - low), xval - move * (xmax - xmin)), xmin)
    beta = np.minimum(np.minimum(upp - ALBEFA * (upp - xval), xval + move * (xmax - xmin)), xmax)
    xmamiinv = 1.0 / np.maximum(xmax - xmin, 1e-5)
    ux1 = upp - xval
    xl1 = xval - low
    ux2 = ux1 * ux1
    xl2 = xl1 * xl1
    p0 = np.maximum(df0dx, 0.0)
    q0 = np.maximum(-df0dx, 0.0)
    pq0 = 0.001 * (p0 + q0) + RAA0 * xmamiinv
    p0 = (p0 + pq0) * ux2
    q0 = (q0 + pq0) * xl2
    P = np.maximum(dfdx, 0.0)
    Q = np.maximum(-dfdx, 0.0)
    PQ = 0.001 * (P + Q) + RAA0 * xmamiinv[None, :]
    P = (P + PQ) * ux2[None, :]
    Q = (Q + PQ) * xl2[None, :]
    b = P @ (1.0 / ux1) + Q @ (1.0 / xl1) - fval
    return alfa, beta, p0, q0, P, Q, b


def approx_value(x, low, upp, p0, q0, P, Q, b, f0val=0.0, xval=None):
    """Values of the approximating functions at x: objective part (up to the
    constant r0) and the m constraint approximations sum P/(U-x)+Q/(x-L)-b."""
    f0 = np.sum(p0 / (upp - x) + q0 / (x - low))
    fi = P @ (1.0 / (upp - x)) + Q @ (1.0 / (x - low)) - b
    return f0, fi


class _State:
    __slots__ = ("x", "y", "z", "lam", "xsi", "eta", "mu", "zet", "s")

    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)

    def axpy(self, steg, d):
        return _State(**{k: getattr(self, k) + steg * getattr(d, k) for k in self.__slots__})


def _residual(v, epsi, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d):
    ux1 = upp - v.x
    xl1 = v.x - low
    plam = p0 + P.T @ v.lam
    qlam = q0 + Q.T @ v.lam
    gvec = P @ (1.0 / ux1) + Q @ (1.0 / xl1)
    dpsidx = plam / (ux1 * ux1) - qlam / (xl1 * xl1)
    parts = [dpsidx - v.xsi + v.eta,
             c + d * v.y - v.mu - v.lam,
             np.atleast_1d(a0 - v.zet - a @ v.lam),
             gvec - a * v.z - v.y + v.s - b,
             v.xsi * (v.x - alfa) - epsi,
             v.eta * (beta - v.x) - epsi,
             v.mu * v.y - epsi,
             np.atleast_1d(v.zet * v.z - epsi),
             v.lam * v.s - epsi]
    r = np.concatenate(parts)
    return np.sqrt(r @ r), np.max(np.abs(r))


def subsolv(m, n, epsimin, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d):
    """Primal-dual Newton method for the MMA subproblem (Svanberg's subsolv):
    barrier levels epsi = 1, 0.1, ... down to epsimin, at most 200 Newton
    steps per level while max|residual| > 0.9 epsi, each step shortened to
    stay 1% inside the bounds and then halved (at most 50 times) until the
    residual norm does not grow.  Returns the final _State; the counts of
    Newton steps, residual evaluations and barrier levels are kept in
    ``subsolv.last_counts``."""
    newton = evals = levels = 0
    epsi = 1.0
    x = 0.5 * (alfa + beta)
    v = _State(x=x, y=np.ones(m), z=1.0, lam=np.ones(m),
               xsi=np.maximum(1.0 / (x - alfa), 1.0), eta=np.maximum(1.0 / (beta - x), 1.0),
               mu=np.maximum(np.ones(m), 0.5 * c), zet=1.0, s=np.ones(m))
    args = (low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
    while epsi > epsimin:
        residunorm, residumax = _residual(v, epsi, *args)
        evals += 1
        ittt = 0
        while residumax > 0.9 * epsi and ittt < 200:
            ittt += 1
            newton += 1
            ux1 = upp - v.x
            xl1 = v.x - low
            ux2 = ux1 * ux1
            xl2 = xl1 * xl1
            plam = p0 + P.T @ v.lam
            qlam = q0 + Q.T @ v.lam
            gvec = P @ (1.0 / ux1) + Q @ (1.0 / xl1)
            GG = P / ux2[None, :] - Q / xl2[None, :]
            dpsidx = plam / ux2 - qlam / xl2
            delx = dpsidx - epsi / (v.x - alfa) + epsi / (beta - v.x)
            dely = c + d * v.y - v.lam - epsi / v.y
            delz = a0 - a @ v.lam - epsi / v.z
            dellam = gvec - a * v.z - v.y - b + epsi / v.lam
            diagx = 2.0 * (plam / (ux2 * ux1) + qlam / (xl2 * xl1)) \
                + v.xsi / (v.x - alfa) + v.eta / (beta - v.x)
            diagy = d + v.mu / v.y
            diaglamyi = v.s / v.lam + 1.0 / diagy
            if m < n:
                blam = dellam + dely / diagy - GG @ (delx / diagx)
                Alam = np.diag(diaglamyi) + (GG / diagx[None, :]) @ GG.T
                AA = np.zeros((m + 1, m + 1))
                AA[:m, :m] = Alam
                AA[:m, m] = a
                AA[m, :m] = a
                AA[m, m] = -v.zet / v.z
                sol = np.linalg.solve(AA, np.concatenate([blam, [delz]]))
                dlam, dz = sol[:m], sol[m]
                dx = -delx / diagx - (GG.T @ dlam) / diagx
            else:
                diaglamyiinv = 1.0 / diaglamyi
                dellamyi = dellam + dely / diagy
                Axx = np.diag(diagx) + (GG.T * diaglamyiinv[None, :]) @ GG
                azz = v.zet / v.z + a @ (a / diaglamyi)
                axz = -GG.T @ (a / diaglamyi)
                bx = delx + GG.T @ (dellamyi / diaglamyi)
                bz = delz - a @ (dellamyi / diaglamyi)
                AA = np.zeros((n + 1, n + 1))
                AA[:n, :n] = Axx
                AA[:n, n] = axz
                AA[n, :n] = axz
                AA[n, n] = azz
                sol = np.linalg.solve(AA, -np.concatenate([bx, [bz]]))
                dx, dz = sol[:n], sol[n]
                dlam = (GG @ dx) / diaglamyi - dz * (a / diaglamyi) + dellamyi / diaglamyi
            dy = -dely / diagy + dlam / diagy
            dxsi = -v.xsi + epsi / (v.x - alfa) - (v.xsi * dx) / (v.x - alfa)
            deta = -v.eta + epsi / (beta - v.x) + (v.eta * dx) / (beta - v.x)
            dmu = -v.mu + epsi / v.y - (v.mu * dy) / v.y
            dzet = -v.zet + epsi / v.z - v.zet * dz / v.z
            ds = -v.s + epsi / v.lam - (v.s * dlam) / v.lam
            dv = _State(x=dx, y=dy, z=dz, lam=dlam, xsi=dxsi, eta=deta, mu=dmu, zet=dzet, s=ds)
            xx = np.concatenate([v.y, [v.z], v.lam, v.xsi, v.eta, v.mu, [v.zet], v.s])
            dxx = np.concatenate([dy, [dz], dlam, dxsi, deta, dmu, [dzet], ds])
            stmxx = np.max(-1.01 * dxx / xx)
            stmalfa = np.max(-1.01 * dx / (v.x - alfa))
            stmbeta = np.max(1.01 * dx / (beta - v.x))
            steg = 1.0 / max(stmalfa, stmbeta, stmxx, 1.0)
            old = v
            itto = 0
            resinew = 2.0 * residunorm
            while resinew > residunorm and itto < 50:
                itto += 1
                v = old.axpy(steg, dv)
                resinew, residumax = _residual(v, epsi, *args)
                evals += 1
                steg = steg / 2.0
            residunorm = resinew
        epsi = 0.1 * epsi
        levels += 1
    subsolv.last_counts = (newton, evals, levels)
    return v


subsolv.last_counts = None


def mmasub(m, n, it, xval, xmin, xmax, xold1, xold2, f0val, df0dx, fval, dfdx, low, upp,
           a0, a, c, d, move):
    """One MMA step: asymptotes, approximation, subproblem solve.
    Returns (xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, low, upp)."""
    low, upp = asymptotes(it, xval, xmin, xmax, xold1, xold2, low, upp)
    alfa, beta, p0, q0, P, Q, b = approximation(xval, xmin, xmax, low, upp, df0dx, fval, dfdx,
                                                move)
    v = subsolv(m, n, EPSIMIN, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
    return v.x, v.y, v.z, v.lam, v.xsi, v.eta, v.mu, v.zet, v.s, low, upp


def kktcheck(m, n, x, y, z, lam, xsi, eta, mu, zet, s, xmin, xmax, df0dx, fval, dfdx, a0, a, c,
             d):
    """KKT residual of the original problem at an MMA iterate:
    (residu, residunorm, residumax)."""
    parts = [df0dx + dfdx.T @ lam - xsi + eta,
             c + d * y - mu - lam,
             np.atleast_1d(a0 - zet - a @ lam),
             fval - a * z - y + s,
             xsi * (x - xmin),
             eta * (xmax - x),
             mu * y,
             np.atleast_1d(zet * z),
             lam * s]
    r = np.concatenate(parts)
    return r, np.sqrt(r @ r), np.max(np.abs(r))


def run_mma(n, m, func, x0, xmin, xmax, fTolerance=1e-4, gTolerance=1e-4, maxIterations=250,
            move_limit=0.2, kktTol=1e-6):
    """The outer loop of src/mmaWrapper.py:18-134 without its timing and
    logging: c = 100, d = 1, a0 = 1, a = 0; the first gradient is passed to
    mmasub unscaled (:73,90) and later ones divided by the first objective
    value (:80,101-102); stop on kktnorm <= kktTol, on fErr < fTolerance with
    max(g) < gTolerance (:117), or after maxIterations.
    Returns (xval, f0val, df0dx, gval, dgdx, outit, history)."""
    xval = np.array(x0, dtype=float)
    low, upp = xmin.copy(), xmax.copy()
    xold1, xold2 = xval.copy(), xval.copy()
    c, d, a0, a = 100.0 * np.ones(m), np.ones(m), 1.0, np.zeros(m)
    f0val, df0dx, gval, dgdx = func(xval)
    kktnorm = kktTol + 10
    outit = 0
    f0Scaling = f0val if abs(f0val) > 1e-6 else 1
    f0valPrev = f0val / f0Scaling
    history = []
    while kktnorm > kktTol and outit < maxIterations:
        outit += 1
        xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, low, upp = mmasub(
            m, n, outit, xval, xmin, xmax, xold1, xold2, f0val, df0dx, gval, dgdx, low, upp,
            a0, a, c, d, move_limit)
        xold2, xold1, xval = xold1.copy(), xval.copy(), xmma.copy()
        f0val, df0dx, gval, dgdx = func(xval)
        f0val = f0val / f0Scaling
        df0dx = df0dx / f0Scaling
        _, kktnorm, _ = kktcheck(m, n, xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, xmin, xmax,
                                 df0dx, gval, dgdx, a0, a, c, d)
        fErr = abs(f0val - f0valPrev) / (1e-10 + abs(f0val))
        gErr = np.max(gval)
        history.append((f0val, gErr, kktnorm))
        if fErr < fTolerance and gErr < gTolerance:
            break
        f0valPrev = f0val
        if np.isnan(kktnorm):
            break
    return xval, f0val, df0dx, gval, dgdx, outit, history
