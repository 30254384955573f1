"""The allocation-free lookup entry point (egtb_lookup_batch_noalloc)
against egtb_lookup_batch on the same table: identical rates (bit for bit,
NaN where a move is invalid) and statuses on two calls in a row, the
second with fewer boards, and a clear error above its 2^20-board limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_BOARDS = 1 << 20


def lookup(pkg, fn, t, boards):
    b = np.ascontiguousarray(boards, dtype=np.uint64)
    out = np.full((b.size, 4), 7.0, dtype=np.float32)
    st = np.full(b.size, 9, dtype=np.int32)
    rc = fn(t.handle, b.ctypes.data_as(pkg.native.u64p), b.size, out.ctypes.data_as(pkg.native.f32p),
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return rc, out, st


def test_noalloc_lookup_equals_lookup(pkg):
    from importlib import import_module

    serving = import_module("2048endgametablebase_amd.serving")
    print("no-alloc lookup: 442@256 table", flush=True)
    res = pkg.build_in_memory("442", 8)
    t = pkg.Table.from_result(res, "442", 8)
    L = pkg.native.lib()
    try:
        pool = serving.make_query_pool(res, "442", n=300000, seed=4)
        for n in (300000, 4099):  # the second call smaller: stale staging must not leak through
            boards = pool[:n] if n == pool.size else pool[-n:]
            rc_a, a, sa = lookup(pkg, L.egtb_lookup_batch, t, boards)
            rc_b, b, sb = lookup(pkg, L.egtb_lookup_batch_noalloc, t, boards)
            assert rc_a == 0 and rc_b == 0, pkg.native.last_error()
            assert np.array_equal(sa, sb)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))  # bit for bit, NaNs included
            assert (sb == 0).any() and (sb == 1).any()
        # n = 0 is a no-op; above the limit is an argument error
        rc, _, _ = lookup(pkg, L.egtb_lookup_batch_noalloc, t, np.zeros(0, dtype=np.uint64))
        assert rc == 0
        big = np.zeros(MAX_BOARDS + 1, dtype=np.uint64)
        rc, _, _ = lookup(pkg, L.egtb_lookup_batch_noalloc, t, big)
        assert rc == -1  # EGTB_ERR_ARG
        assert "at most" in pkg.native.last_error()
    finally:
        t.close()
        res.close()
