"""The allocation-free sort, sphx_sort_by_key_noalloc, called through the C ABI: bit for bit
the result of sphx_sort_by_key (keys and the stable order of equal keys) on two calls in a
row, the second smaller, with nothing taken from the context's scratch arena; sizes beyond
its maximum are refused with SPHX_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NA_MAX = 1 << 21


def _sorted_with(ctx, fn, keys, vals):
    k, v = keys.clone(), vals.clone()
    ctx.check(fn(ctx.h, C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), k.numel()))
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint64), v.cpu().numpy().view(np.uint32)


def _arena_bytes(ctx):
    from sphexa_amd import _lib

    b = C.c_double()
    ctx.check(ctx.lib.sphx_get_stat(ctx.h, _lib.STAT_SCRATCH_BYTES, C.byref(b)))
    return b.value


def _keys(n, seed):
    """Hilbert-like 63-bit keys with many ties (values drawn from n // 4 distinct keys) and a
    run of identical keys, so stability and every digit pass matter."""
    rng = np.random.default_rng(seed)
    distinct = rng.integers(0, 2 ** 63, max(1, n // 4), dtype=np.uint64)
    k = distinct[rng.integers(0, len(distinct), n)]
    k[: min(n, 1000)] = distinct[0]
    return torch.from_numpy(k.view(np.int64)).to("cuda"), torch.arange(n, dtype=torch.int32, device="cuda")


def test_sort_noalloc_matches_sort_by_key():
    from sphexa_amd import hydro

    na = hydro.Context(0)  # a fresh context: its arena must stay empty
    ref = hydro.Context(0)
    try:
        for n, seed in ((NA_MAX, 1), (100_003, 2)):
            print("case sort_by_key_noalloc n", n, flush=True)
            keys, vals = _keys(n, seed)
            got_k, got_v = _sorted_with(na, na.lib.sphx_sort_by_key_noalloc, keys, vals)
            want_k, want_v = _sorted_with(ref, ref.lib.sphx_sort_by_key, keys, vals)
            assert np.array_equal(got_k, want_k)
            assert np.array_equal(got_v, want_v)
            assert np.all(got_k[1:] >= got_k[:-1])
            assert _arena_bytes(na) == 0
    finally:
        na.close()
        ref.close()


def test_sort_noalloc_limits():
    from sphexa_amd import _lib, hydro

    ctx = hydro.Context(0)
    try:
        k = torch.zeros(NA_MAX + 1, dtype=torch.int64, device="cuda")
        v = torch.zeros(NA_MAX + 1, dtype=torch.int32, device="cuda")
        st = ctx.lib.sphx_sort_by_key_noalloc(ctx.h, C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), NA_MAX + 1)
        assert st == _lib.SPHX_ERR_ARG
        assert b"maximum" in ctx.lib.sphx_last_error(ctx.h)
        for n in (0, 1):  # no-ops
            assert ctx.lib.sphx_sort_by_key_noalloc(ctx.h, C.c_void_p(k.data_ptr()), C.c_void_p(v.data_ptr()), n) == 0
        assert _arena_bytes(ctx) == 0
    finally:
        ctx.close()
