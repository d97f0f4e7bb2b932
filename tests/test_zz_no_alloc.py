"""dtwax_distances_noalloc: the allocation-free variant of dtwax_distances
(no counterpart in the reference).  Called through the C ABI, two calls in a
row (the second smaller) give the same dists and last rows as
dtwax_distances on the same inputs, bit for bit; device memory in use does
not change across its calls; sizes beyond its limit are refused."""
import ctypes as C

import numpy as np
import pytest

import dtwax_amd as D
from helpers import half2_words, query_words_h2, tile_reference

pytestmark = pytest.mark.gpu

REF_LEN = 48128
LIMIT = 16384


def _mem_used():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value


def _call(fn, ctx, d_ref, d_q, d_out, n, d_lr, quirks):
    rc = fn(ctx.handle, d_ref.ptr, d_q.ptr, d_out.ptr, n, 0, None, d_lr.ptr if d_lr else None, quirks)
    assert rc == 0, D.lib().dtwax_last_error()
    assert D.lib().dtwax_device_sync() == 0


def test_noalloc_equals_distances():
    L = D.lib()
    ctx = D.Context(0)
    rng = np.random.default_rng(771)
    flat = rng.standard_normal((2, REF_LEN)).astype(np.float16).astype(np.float32)
    d_ref = D.DeviceArray.from_host(tile_reference(half2_words(flat[0], flat[1])))
    cases = []  # (n, quirks): the second call smaller than the first
    for n, quirks in ((64, 1), (5, 0), (3, 2)):
        qs = rng.standard_normal((n, 2048)).astype(np.float32)
        d_q = D.DeviceArray.from_host(query_words_h2(qs))
        cases.append((n, quirks, d_q))
    # the allocating entry point first (its buffers are then in place), and
    # one call of each form so every kernel is loaded before memory is read
    want = []
    for n, quirks, d_q in cases:
        d_out = D.DeviceArray((n,), np.uint32)
        d_lr = D.DeviceArray.from_host(np.zeros(n * REF_LEN, np.uint32))
        _call(L.dtwax_distances, ctx, d_ref, d_q, d_out, n, d_lr, quirks)
        want.append((d_out.to_host(), d_lr.to_host()))
        _call(L.dtwax_distances_noalloc, ctx, d_ref, d_q, d_out, n, d_lr, quirks)
    outs = [(D.DeviceArray((n,), np.uint32), D.DeviceArray.from_host(np.zeros(n * REF_LEN, np.uint32)))
            for n, _, _ in cases]
    used0 = _mem_used()
    for (n, quirks, d_q), (d_out, d_lr) in zip(cases, outs):
        _call(L.dtwax_distances_noalloc, ctx, d_ref, d_q, d_out, n, d_lr, quirks)
    used1 = _mem_used()
    assert used1 == used0, (used0, used1)
    for (n, quirks, _), (d_out, d_lr), (w_out, w_lr) in zip(cases, outs, want):
        assert np.array_equal(d_out.to_host(), w_out), (n, quirks)
        assert np.array_equal(d_lr.to_host(), w_lr), (n, quirks)
    # beyond the limit: refused before any work
    n, quirks, d_q = cases[0]
    rc = L.dtwax_distances_noalloc(ctx.handle, d_ref.ptr, d_q.ptr, outs[0][0].ptr, LIMIT + 1, 0, None, None, 1)
    assert rc == -1 and b"allocation-free limit" in L.dtwax_last_error()
    ctx.close()
