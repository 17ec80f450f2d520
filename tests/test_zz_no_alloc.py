"""The allocation-free OC update (pyto_oc_update_noalloc) against the
handle-based pyto_oc_update, called through the C ABI: the same xnew bit for
bit, the same multiplier and step count, on two calls in a row, the second
with a smaller input (the static scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _noalloc(x, dc, volfrac, move=0.2, xmin=0.001, xmax=1.0, max_steps=160):
    from pyto_amd import _lib
    n = x.numel()
    xnew = torch.empty_like(x)
    lm = ctypes.c_double()
    st = ctypes.c_int()
    _lib.check(_lib.lib().pyto_oc_update_noalloc(
        _lib.ptr(x, torch.float64, n), _lib.ptr(dc, torch.float64, n), n, 1.0 / volfrac / n,
        float(volfrac), float(move), float(xmin), float(xmax), int(max_steps), _lib.ptr(xnew),
        ctypes.byref(lm), ctypes.byref(st), _lib.stream_ptr()), "pyto_oc_update_noalloc")
    return xnew, lm.value, st.value


def test_noalloc_oc_update_matches_handle_entry_point(gpu):
    from pyto_amd.operator import GridOperator
    print("no-alloc OC update", flush=True)
    op = GridOperator((4, 3, 2), (1, 1, 1), 1.0, 0.3, [])
    for n, volfrac, seed in ((200003, 0.45, 11), (777, 0.3, 12)):
        rng = np.random.default_rng(seed)
        x = torch.from_numpy(rng.uniform(0.001, 1.0, n)).to(gpu)
        g = -rng.uniform(0.0, 3.0, n)
        g[rng.choice(n, n // 20, replace=False)] *= -0.1
        dc = torch.from_numpy(g).to(gpu)
        ref, lref, sref = op.oc_update(x, dc, volfrac, read_back=True)
        out, lm, st = _noalloc(x, dc, volfrac)
        print("n", n, "steps", st, sref, "lmid", lm, lref, flush=True)
        assert st == sref and st > 0
        assert lm == lref
        assert torch.equal(out, ref)


def test_noalloc_rejects_bad_arguments(gpu):
    from pyto_amd import _lib
    x = torch.ones(10, dtype=torch.float64, device=gpu)
    with pytest.raises(ValueError, match="max_steps"):
        _lib.check(_lib.lib().pyto_oc_update_noalloc(
            _lib.ptr(x), _lib.ptr(x), 10, 0.2, 0.5, 0.2, 0.001, 1.0, 5000, _lib.ptr(x), None, None,
            _lib.stream_ptr()))
    with pytest.raises(ValueError, match="n must be positive"):
        _lib.check(_lib.lib().pyto_oc_update_noalloc(
            _lib.ptr(x), _lib.ptr(x), 0, 0.2, 0.5, 0.2, 0.001, 1.0, 10, _lib.ptr(x), None, None,
            _lib.stream_ptr()))
