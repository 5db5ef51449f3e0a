"""tm_match_rows_static, the entry point that allocates nothing: its rows
equal those of a KEY_BITMAP ring (the existing entry point) on calls in a
row of decreasing and increasing size, which share the static scratch array;
sizes above its limits are refused."""
import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def ring_rows(fib, q):
    r = tm.Ring(fib, len(q), mode=tm.MODE_KEY_BITMAP)
    r.submit(q)
    r.wait()
    rows = r.key_bitmaps()
    r.close()
    return rows


def test_static_entry_point_equals_ring(gpu):
    cw, cko, ck, q, _ = tm.workload(60000, 20000, seed=1301)
    fib, *_ = tm.build_fib(cw, cko, ck, 6000)
    kw = (int(fib.stats()["n_distinct_keys"]) + 31) // 32
    nmax = len(q)
    dq = tm.DeviceBuffer(nmax * 32)
    drows = tm.DeviceBuffer(nmax * kw * 4)
    lib = tm.load()
    d = fib.dictionary()
    # a compacted step, a small segmented one, then a large one again: the
    # second call has a smaller input, the third reuses what it left
    for n in (20000, 777, 20000, 64):
        log(f"static entry point: {n} queries")
        qp = np.zeros((n, 8), np.uint32)
        qp[:, :6] = q[:n]
        dq.upload(qp)
        drows.upload(np.full(n * kw, 0xA5A5A5A5, np.uint32))   # stale contents must not leak
        tm.check(lib.tm_match_rows_static(fib.h, tm.C.c_void_p(dq.ptr), n, tm.C.c_void_p(drows.ptr), None),
                 "tm_match_rows_static")
        got = drows.download(np.zeros((n, kw), np.uint32))
        exp = ring_rows(fib, q[:n])
        assert np.array_equal(got, exp), n
        if n <= 777:
            eo, ev = orc.match(cw, cko, ck, q[:n], mode=0)
            off, vals = tm.expand_key_bitmaps(got, d)
            assert np.array_equal(off, eo) and np.array_equal(vals, ev)
    # limits and arguments
    assert lib.tm_match_rows_static(fib.h, tm.C.c_void_p(dq.ptr), tm.STATIC_MAX_QUERIES + 1,
                                    tm.C.c_void_p(drows.ptr), None) == tm.TM_ERR_INVALID
    assert lib.tm_match_rows_static(fib.h, None, 10, tm.C.c_void_p(drows.ptr), None) == tm.TM_ERR_INVALID
    assert lib.tm_match_rows_static(fib.h, tm.C.c_void_p(dq.ptr), 0, None, None) == tm.TM_OK
    dq.close()
    drows.close()
