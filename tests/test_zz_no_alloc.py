"""cbspmv_backward_stored_noalloc (no reference counterpart: an entry point
for callers that cannot allocate device memory during the call) against
cbspmv_backward_stored, called through the C ABI.

Both run the same fused pass over the transposed plan, so stored-order dvals
and dx must agree bit for bit: twice in a row, the second time on a smaller
matrix, with dx wanted and not wanted (then the pass writes dx into the
library's static device array).  Also: the size limit of that array and the
state checks are reported as errors, not faults.
"""
import ctypes

import numpy as np
import pytest

from cbspmv import Plan, gen
from cbspmv import _lib

pytestmark = pytest.mark.gpu


def _noalloc():
    vp = ctypes.c_void_p
    return _lib.bind("cbspmv_backward_stored_noalloc", ctypes.c_int, [vp, vp, vp, vp, vp, vp])


def _call(plan, g, x, dx, dvals, stream):
    f = _noalloc()
    return f(plan.handle, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(x.data_ptr()),
             ctypes.c_void_p(dx.data_ptr() if dx is not None else 0), ctypes.c_void_p(dvals.data_ptr()),
             ctypes.c_void_p(stream.cuda_stream))


def _plan(m, n, per_row, seed, dtype, long_row=False):
    r, c, v = gen.random_coo(m, n, per_row, seed=seed, dup=True)
    if long_row:   # one row longer than four work items: cut into pieces
        rng = np.random.default_rng(seed)
        lc = np.sort(rng.choice(n, size=min(n, 20000), replace=False)).astype(np.int32)
        at = np.searchsorted(r, m // 2)
        r = np.concatenate([r[:at], np.full(lc.size, m // 2, np.int32), r[at:]])
        c = np.concatenate([c[:at], lc, c[at:]])
        v = np.concatenate([v[:at], rng.uniform(-1, 1, lc.size), v[at:]])
    p = Plan(m, n, r, c, v, dtype=dtype)
    p.build_transpose()
    p.upload()
    return p


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_matches_backward_stored(gpu, dtype):
    import torch

    tdt = torch.float64 if dtype == "f64" else torch.float32
    s = torch.cuda.current_stream()
    sizes = [(6000, 30000, 12, 1, True), (700, 900, 5, 2, False)]   # second call: smaller input
    for m, n, per_row, seed, long_row in sizes:
        print(f"noalloc {dtype}: m={m} n={n}", flush=True)
        p = _plan(m, n, per_row, seed, dtype, long_row)
        rng = np.random.default_rng(seed + 10)
        g = torch.tensor(rng.uniform(-1, 1, m), dtype=tdt, device=gpu)
        x = torch.tensor(rng.uniform(-1, 1, n), dtype=tdt, device=gpu)
        k = p.stored_entries()
        want_dv = torch.empty(k, dtype=tdt, device=gpu)
        want_dx = torch.empty(n, dtype=tdt, device=gpu)
        p.backward_stored(g, x, want_dv, want_dx)
        for with_dx in (False, True):
            dv = torch.full((k,), float("nan"), dtype=tdt, device=gpu)
            dx = torch.full((n,), float("nan"), dtype=tdt, device=gpu) if with_dx else None
            rc = _call(p, g, x, dx, dv, s)
            assert rc == 0, _lib.lib().cbspmv_last_error()
            torch.cuda.synchronize()
            assert torch.equal(dv, want_dv), f"dvals differ (dx given: {with_dx})"
            if with_dx:
                assert torch.equal(dx, want_dx)
        p.close()


def test_noalloc_errors(gpu):
    import torch

    s = torch.cuda.current_stream()
    # 2.2M FP64 columns need 17.6 MB of dx: above the 16 MiB array
    m, n = 64, 2_200_000
    r = np.arange(m, dtype=np.int32)
    c = (np.arange(m, dtype=np.int64) * 30011 % n).astype(np.int32)
    p = Plan(m, n, r, c, np.ones(m))
    p.build_transpose()
    g = torch.ones(m, dtype=torch.float64, device=gpu)
    x = torch.ones(n, dtype=torch.float64, device=gpu)
    dv = torch.empty(max(1, p.stored_entries()), dtype=torch.float64, device=gpu)
    assert _call(p, g, x, None, dv, s) == _lib.ERR_ARG
    dx = torch.empty(n, dtype=torch.float64, device=gpu)
    # not uploaded yet: a state error, nothing launched
    assert _call(p, g, x, dx, dv, s) == _lib.ERR_STATE
    p.upload()
    assert _call(p, g, x, dx, dv, s) == 0
    torch.cuda.synchronize()
    want = torch.empty_like(dv)
    p.backward_stored(g, x, want)
    torch.cuda.synchronize()
    assert torch.equal(dv, want)
    p.close()
    # no transposed plan
    q = Plan(m, 100, r, c % 100, np.ones(m))
    q.upload()
    assert _call(q, g, torch.ones(100, dtype=torch.float64, device=gpu), None, dv, s) == _lib.ERR_STATE
    q.close()
