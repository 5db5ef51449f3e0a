"""KEY_BITMAP results: per-query rows over the sorted key dictionary, read
back or delivered into pinned host memory at the end of each step, and
expanded on the host; equal to KEYS_UNIQUE (and to the oracle) for every
schedule, step size and graph replay."""
import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def table(gpu):
    cw, cko, ck, q, _ = tm.workload(60000, 20000, seed=1201)
    fib, *_ = tm.build_fib(cw, cko, ck, 6000)
    return cw, cko, ck, q, fib


def test_rows_expand_to_keys_unique(table):
    cw, cko, ck, q, fib = table
    d = fib.dictionary()
    assert np.all(np.diff(d.astype(np.int64)) > 0) and np.array_equal(d, np.unique(ck))
    for nq in (20000, 1500, 37):          # compacted, plain-size and segmented steps
        eo, ev = orc.match(cw, cko, ck, q[:nq], mode=0)
        for sched in (tm.SCHED_PERSISTENT, tm.SCHED_PLAIN):
            log(f"nq={nq} schedule={sched}")
            ring = tm.Ring(fib, nq, mode=tm.MODE_KEY_BITMAP, schedule=sched)
            ring.submit(q[:nq])
            st = ring.wait()
            rows = ring.key_bitmaps()
            assert rows.shape == (nq, (len(d) + 31) // 32)
            off, vals = tm.expand_key_bitmaps(rows, d)
            assert np.array_equal(off, eo) and np.array_equal(vals, ev), (nq, sched)
            assert st.n_results == 0
            with pytest.raises(tm.TagMatchError):
                ring.results()                     # rows, not lists
            # the KEYS_UNIQUE ring holds the same rows
            ring.configure(mode=tm.MODE_KEYS_UNIQUE, schedule=sched)
            ring.submit_resident(nq)
            ring.wait()
            assert np.array_equal(ring.key_bitmaps(), rows)
            o2, v2 = ring.results()
            assert np.array_equal(o2, eo) and np.array_equal(v2, ev)
            ring.close()


def test_host_rows_delivery_and_graph_replay(table):
    cw, cko, ck, q, fib = table
    d = fib.dictionary()
    kw = (len(d) + 31) // 32
    nq = 20000
    eo, ev = orc.match(cw, cko, ck, q[:nq], mode=0)
    for graph in (0, 1):
        rings = [tm.Ring(fib, nq, mode=tm.MODE_KEY_BITMAP, use_graph=graph) for _ in range(2)]
        bufs = [tm.PinnedRows(nq * kw) for _ in range(2)]
        for r, b in zip(rings, bufs):
            r.set_host_rows(b)
            r.submit(q[:nq])
            r.wait()
        # two rings in flight, alternating, as a pipelined caller runs them
        for rep in range(4):
            n = (nq, 12345, nq, 999)[rep]
            for r in rings:
                r.submit_resident(n)
            for r, b in zip(rings, bufs):
                r.wait()
                off, vals = tm.expand_key_bitmaps(b.rows(n, kw), d)
                a = int(eo[n])
                assert np.array_equal(off, eo[:n + 1]) and np.array_equal(vals, ev[:a]), (graph, rep)
        for r in rings:
            r.close()
        # a submission larger than the host buffer is refused
        small = tm.PinnedRows(100 * kw)
        r = tm.Ring(fib, nq, mode=tm.MODE_KEY_BITMAP)
        r.set_host_rows(small)
        with pytest.raises(tm.TagMatchError) as e:
            r.submit(q[:101])
        assert e.value.code == tm.TM_ERR_OVERFLOW
        r.submit(q[:100])
        r.wait()
        assert np.array_equal(small.rows(100, kw), r.key_bitmaps())
        r.close()
        for b in bufs + [small]:
            b.close()


def test_host_rows_only_for_bitmap_rings(table):
    cw, cko, ck, q, fib = table
    ring = tm.Ring(fib, 64)
    buf = tm.PinnedRows(64 * 64)
    with pytest.raises(tm.TagMatchError) as e:
        ring.set_host_rows(buf)
    assert e.value.code == tm.TM_ERR_STATE
    ring.close()
    buf.close()
    with pytest.raises(tm.TagMatchError):
        tm.Server(fib, n_slots=1, slot_queries=64, mode=tm.MODE_KEY_BITMAP)


def test_packed_host_rows(gpu):
    """Packed delivery (tm_ring_set_host_rows_packed): a table where a
    one-bit filter carries every key, so the queries holding that bit match
    the whole dictionary and the others do not.  Per step: the entries number
    the other queries' rows in order, the packed rows are theirs, and the
    expansion equals the oracle; steps whose packed rows outgrow what the
    previous step copied with itself, graph replays, switching forms and a
    capacity error."""
    cw, cko, ck, q, _ = tm.workload(60000, 20000, seed=1211)
    every = np.unique(ck)
    freq = np.array([(q[:, b // 32] >> np.uint32(b % 32) & 1).mean() for b in range(192)])
    bit = int(np.argmin(np.abs(freq - 0.5)))
    w1 = np.zeros((1, 6), np.uint32)
    w1[0, bit // 32] = np.uint32(1 << (bit % 32))
    words = np.vstack([cw, w1])
    ko = np.append(cko.astype(np.uint64), np.uint64(cko[-1]) + np.uint64(every.size))
    cw2, cko2, ck2 = tm.canonicalize(words, ko, np.concatenate([ck, every]))
    fib, *_ = tm.build_fib(cw2, cko2, ck2, 6000)
    d = fib.dictionary()
    kw = (len(d) + 31) // 32
    log(f"bit {bit} in {freq[bit]:.2f} of the queries, K={len(d)}")
    # slices whose rows not full grow from step to step
    steps = [(0, 300), (300, 3000), (3300, 12000), (15300, 40), (15340, 4000)]
    for graph in (0, 1):
        ring = tm.Ring(fib, 12000, mode=tm.MODE_KEY_BITMAP, use_graph=graph)
        buf = tm.PackedRows(12000, kw)
        ring.set_host_rows_packed(buf)
        for a, n in steps:
            qa = q[a:a + n]
            eo, ev = orc.match(cw2, cko2, ck2, qa, mode=0)
            ring.submit(qa)
            st = ring.wait()
            full = np.diff(eo.astype(np.int64)) == len(d)
            assert 0.1 < full.mean() < 0.95 or n < 100
            npk = int(st.n_results)
            row_of = buf.row_of.array[:n].copy()
            assert npk == int((~full).sum())
            assert np.array_equal(row_of == tm.ROW_FULL, full)
            assert np.array_equal(row_of[~full], np.arange(npk))
            packed = buf.rows.array[:npk * kw].reshape(npk, kw)
            assert np.array_equal(buf.unpack(n, npk, len(d)), ring.key_bitmaps())
            off, vals = tm.expand_key_rows_packed(row_of, packed, d)
            assert np.array_equal(off, eo) and np.array_equal(vals, ev), (graph, a, n)
        # the other form replaces this one
        rows = tm.PinnedRows(12000 * kw)
        ring.set_host_rows(rows)
        buf.row_of.array[:] = 7
        ring.submit(q[:500])
        ring.wait()
        assert np.all(buf.row_of.array[:500] == 7)
        assert np.array_equal(rows.rows(500, kw), ring.key_bitmaps())
        ring.set_host_rows(None)
        ring.close()
        rows.close()
        buf.close()
    # capacity below the submission
    ring = tm.Ring(fib, 100, mode=tm.MODE_KEY_BITMAP)
    small = tm.PackedRows(50, kw)
    ring.set_host_rows_packed(small)
    with pytest.raises(tm.TagMatchError):
        ring.submit(q[:100])
    ring.set_host_rows_packed(None)
    ring.close()
    small.close()
    fib.close()
