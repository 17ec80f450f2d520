"""The method of moving asymptotes on the GPU (src/mmaWrapper.py).

``runMMA`` keeps the reference's signature, loop, stopping tests and return
value ``[xval, f0val, df0dx, gval, dgdx, outit]``.  ``mmasub`` and
``kktcheck`` keep mmapy's signatures (the package the reference imports at
src/mmaWrapper.py:14) and run on the device through libpytohip: the design
variables, bounds, asymptotes, gradients and the interior-point state stay in
HBM, and only the m constraint values and the m-sized multipliers cross to
the host.

Arrays may be CUDA tensors or numpy arrays, flat (n,) or columns (n, 1) as
mmapy uses; results are CUDA tensors of shape (n,) (the m-sized multipliers
are numpy arrays).  ``optimizationFunction`` receives the design variables as
a CUDA tensor of shape (n,) and may return tensors or numpy arrays.

Differences from mmapy: ``kktcheck`` returns ``residu=None`` (the residual
vector over the variables is reduced on the device, not materialised); the
subproblem solver requires 1 <= m <= 4 constraints and m < n.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib

MAX_CONSTRAINTS = 4


class MMAWorkspace:
    """Device workspace of the MMA subproblem for n variables and m
    constraints (8 (7 + 2m) n bytes)."""

    def __init__(self, n, m, device=None):
        self.n, self.m = int(n), int(m)
        self.device = torch.device("cuda") if device is None else torch.device(device)
        lib = _lib.lib()
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.pyto_mma_create(ctypes.byref(h), self.n, self.m))
        self._h = h
        self._lib = lib

    def close(self):
        if self._h:
            self._lib.pyto_mma_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_workspaces = {}


def _workspace(n, m, device):
    key = (n, m, str(device))
    w = _workspaces.get(key)
    if w is None:
        w = MMAWorkspace(n, m, device)
        _workspaces.clear()   # one live workspace: the last shape used
        _workspaces[key] = w
    return w


def _dev(a, device, shape=None):
    t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a, dtype=np.float64))
    t = t.to(device=device, dtype=torch.float64)
    if shape is not None:
        t = t.reshape(shape)
    return t.contiguous()


def _host(a, m):
    """m values on the host; a scalar is broadcast."""
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    if a.size == 1:
        a = np.full(m, a[0])
    if a.size != m:
        raise ValueError(f"expected {m} values, got {a.size}")
    return np.ascontiguousarray(a)


def _scalar(a):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return float(np.asarray(a, dtype=np.float64).reshape(-1)[0])


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _hptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pick_device(*arrays):
    for a in arrays:
        if isinstance(a, torch.Tensor) and a.is_cuda:
            return a.device
    return torch.device("cuda")


def mmasub(m, n, iter, xval, xmin, xmax, xold1, xold2, f0val, df0dx, fval, dfdx, low, upp, a0,
           a, c, d, move=0.5):
    """One MMA step (mmapy's mmasub; called at src/mmaWrapper.py:90-91).

    Returns (xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, low, upp): xmma,
    xsi, eta, low, upp are CUDA tensors (n,), the others numpy arrays (m,)
    (zmma and zet floats)."""
    m, n = int(m), int(n)
    if not 1 <= m <= MAX_CONSTRAINTS:
        raise ValueError(f"the GPU MMA subproblem supports 1..{MAX_CONSTRAINTS} constraints, got {m}")
    dev = _pick_device(xval, df0dx, dfdx)
    w = _workspace(n, m, dev)
    xv, lo, hi = _dev(xval, dev, (n,)), _dev(xmin, dev, (n,)), _dev(xmax, dev, (n,))
    o1, o2 = _dev(xold1, dev, (n,)), _dev(xold2, dev, (n,))
    g0, G = _dev(df0dx, dev, (n,)), _dev(dfdx, dev, (m, n))
    low_t, upp_t = _dev(low, dev, (n,)).clone(), _dev(upp, dev, (n,)).clone()
    xmma = torch.empty(n, dtype=torch.float64, device=dev)
    xsi, eta = torch.empty_like(xmma), torch.empty_like(xmma)
    fv, av, cv, dv = (_host(v, m) for v in (fval, a, c, d))
    scal = np.zeros(4 * m + 2)
    st = _lib.MmaStats()
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(w._lib.pyto_mma_sub(
            w._h, int(iter), _ptr(xv), _ptr(lo), _ptr(hi), _ptr(o1), _ptr(o2), _ptr(g0), _ptr(G),
            _hptr(fv), _scalar(a0), _hptr(av), _hptr(cv), _hptr(dv),
            float(move), _ptr(low_t), _ptr(upp_t), _ptr(xmma), _ptr(xsi), _ptr(eta), _hptr(scal),
            ctypes.byref(st), ctypes.c_void_p(stream)))
    mmasub.last_stats = st.as_dict()
    y, z = scal[:m].copy(), float(scal[m])
    lam, mu = scal[m + 1:2 * m + 1].copy(), scal[2 * m + 1:3 * m + 1].copy()
    zet, s = float(scal[3 * m + 1]), scal[3 * m + 2:4 * m + 2].copy()
    return xmma, y, z, lam, xsi, eta, mu, zet, s, low_t, upp_t


mmasub.last_stats = None


def kktcheck(m, n, x, y, z, lam, xsi, eta, mu, zet, s, xmin, xmax, df0dx, fval, dfdx, a0, a, c,
             d):
    """KKT residual of the original problem (mmapy's kktcheck; called at
    src/mmaWrapper.py:108-109) -> (None, residunorm, residumax)."""
    m, n = int(m), int(n)
    dev = _pick_device(x, df0dx, dfdx)
    w = _workspace(n, m, dev)
    xt, xs, et = _dev(x, dev, (n,)), _dev(xsi, dev, (n,)), _dev(eta, dev, (n,))
    lo, hi = _dev(xmin, dev, (n,)), _dev(xmax, dev, (n,))
    g0, G = _dev(df0dx, dev, (n,)), _dev(dfdx, dev, (m, n))
    lamh = _host(lam, m)
    out = np.zeros(2)
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        _lib.check(w._lib.pyto_mma_kkt(w._h, _ptr(xt), _ptr(xs), _ptr(et), _ptr(lo), _ptr(hi),
                                       _ptr(g0), _ptr(G), _hptr(lamh), _hptr(out),
                                       ctypes.c_void_p(stream)))
    yv, muv, sv, fv = _host(y, m), _host(mu, m), _host(s, m), _host(fval, m)
    av, cv, dv = _host(a, m), _host(c, m), _host(d, m)
    z, zet, a0 = _scalar(z), _scalar(zet), _scalar(a0)
    rows = np.concatenate([cv + dv * yv - muv - lamh, [a0 - zet - av @ lamh],
                           fv - av * z - yv + sv, muv * yv, [zet * z], lamh * sv])
    sq = out[0] + rows @ rows
    return None, float(np.sqrt(sq)), float(max(out[1], np.abs(rows).max()))


def runMMA(nVariables, nConstraints, optimizationFunction, X0, lowerBound, upperBound,
           fTolerance=1e-4, gTolerance=1e-4, maxIterations=250, timeLimitSecs=3600,
           move_limit=0.2, kktTol=1e-6, verbose=False, progress_callback=None):
    """src/mmaWrapper.py:18-134.  The loop, its scaling of the objective by
    its first value (the first gradient goes to mmasub unscaled, :73,90), its
    stopping tests (kktnorm, fErr < fTolerance with max(g) < gTolerance, the
    time limit, a NaN kktnorm) and the return value are the reference's."""
    def log_message(msg):
        if progress_callback:
            progress_callback(str(msg))
        else:
            print(msg)

    n, m = int(nVariables), int(nConstraints)
    dev = _pick_device(X0, lowerBound, upperBound)
    xval = _dev(X0, dev, (n,)).clone()
    xmin, xmax = _dev(lowerBound, dev, (n,)), _dev(upperBound, dev, (n,))
    low, upp = xmin.clone(), xmax.clone()
    xold1, xold2 = xval.clone(), xval.clone()
    c, d, a0, a = 100.0 * np.ones(m), np.ones(m), 1.0, np.zeros(m)

    def evaluate(x):
        f0, g0, gv, G = optimizationFunction(x)
        return _scalar(f0), _dev(g0, dev, (n,)), _host(gv, m), _dev(G, dev, (m, n))

    f0val, df0dx, gval, dgdx = evaluate(xval)
    kktnorm = kktTol + 10
    outit = 0
    timeMMA = timeFuncEval = 0.0
    f0Scaling = f0val if abs(f0val) > 1e-6 else 1
    f0valPrev = f0val / f0Scaling
    tStart = time.time()
    while kktnorm > kktTol and outit < maxIterations:
        outit += 1
        t0 = time.time()
        xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, low, upp = mmasub(
            m, n, outit, xval, xmin, xmax, xold1, xold2, f0val, df0dx, gval, dgdx, low, upp, a0,
            a, c, d, move_limit)
        timeMMA += time.time() - t0
        xold2, xold1, xval = xold1, xval, xmma.clone()
        t0 = time.time()
        f0val, df0dx, gval, dgdx = evaluate(xval)
        f0val = f0val / f0Scaling
        df0dx = df0dx / f0Scaling
        timeFuncEval += time.time() - t0
        t0 = time.time()
        _, kktnorm, _ = kktcheck(m, n, xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, xmin, xmax,
                                 df0dx, gval, dgdx, a0, a, c, d)
        timeMMA += time.time() - t0
        fErr = abs(f0val - f0valPrev) / (1e-10 + abs(f0val))
        gErr = float(np.max(gval))
        if verbose:
            log_message(f"iter: {outit}, f: {f0val:.3e}, max(g): {gErr:.3e}, fErr: {fErr:.3e}")
        if fErr < fTolerance and gErr < gTolerance:
            if verbose:
                log_message(f"Convergence reached with fEerr: {fErr:.3e} and max(gval): {gErr:.3e}")
            break
        f0valPrev = f0val
        if time.time() - tStart > timeLimitSecs:
            log_message(f"Time limit of {timeLimitSecs:0.2f} reached, exiting...")
            break
        if np.isnan(kktnorm):
            log_message("kktnorm is nan, something wrong in the optimization function, exiting...")
            break
    if verbose:
        log_message(f"MMA (secs):  {timeMMA:0.2f}")
        log_message(f"FuncEval (secs): {timeFuncEval:0.2f}")
    return [xval, f0val, df0dx, gval, dgdx, outit]
