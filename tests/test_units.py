"""Explicit (partition, query batch) units without the lookup
(tm_submit_units): Zipf-sized partitions cut from the sorted table and
log-uniform batches per partition, both schedules, every output mode,
against the oracle run on each unit's partition."""
import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu

MODES = [(tm.MODE_KEYS_UNIQUE, 0), (tm.MODE_KEYS_ALL, 1), (tm.MODE_HITS, 2)]


def log(msg):
    print(msg, flush=True)


def expected(cw, ck, cv, off, q, up, uo, om):
    """Per query: the oracle over the unit's partition only."""
    outs, vals = [0], []
    for u in range(len(up)):
        a, b = int(off[up[u]]), int(off[up[u] + 1])
        qa = q[uo[u]:uo[u + 1]]
        if len(qa) == 0:
            continue
        ko = ck[a:b + 1] - ck[a]
        eo, ev = orc.match(cw[a:b], ko, cv[ck[a]:ck[b]], qa, mode=om)
        if om == 2:
            ev = ev + a                                   # ids are positions in the whole table
        for i in range(len(qa)):
            vals.append(np.sort(ev[eo[i]:eo[i + 1]]))
            outs.append(outs[-1] + len(vals[-1]))
    return np.array(outs, np.uint64), (np.concatenate(vals) if vals else np.zeros(0, np.uint32))


@pytest.mark.parametrize("reserve", [0, 13])
def test_uneven_units_vs_oracle(gpu, reserve):
    cw, ck, cv, _, _ = tm.workload(60000, 1, seed=801)
    off = tm.zipf_runs(len(cw), 100, 12000, seed=3)
    fib, masks = tm.fib_from_runs(cw, ck, cv, off, reserve=reserve)
    up, uo = tm.uneven_units(len(off) - 1, 1, 700, seed=4)
    up = np.concatenate([up, up[:3]])                     # a partition may get several units
    uo = np.concatenate([uo, uo[-1] + np.array([5, 300, 301], np.uint32)])
    nq = int(uo[-1])
    g = tm.Generator(seed=802)
    q, _ = g.queries(nq, len(cw))
    log(f"{len(off) - 1} partitions, {len(up)} units, {nq} queries")
    for mode, om in MODES:
        eo, ev = expected(cw, ck, cv, off, q, up, uo, om)
        for sched in (tm.SCHED_PERSISTENT, tm.SCHED_PLAIN):
            ring = tm.Ring(fib, nq, mode=mode, schedule=sched)
            ring.submit_units(q, up, uo)
            st = ring.wait()
            o, v = ring.results()
            assert np.array_equal(o, eo), (mode, sched)
            got = [np.sort(v[o[i]:o[i + 1]]) for i in range(nq)]
            assert np.array_equal(np.concatenate(got) if got else v, ev), (mode, sched)
            # resident queries, a second time (HITS / KEYS_ALL lists are unordered)
            ring.submit_units(None, up, uo)
            ring.wait()
            o2, v2 = ring.results()
            assert np.array_equal(o, o2)
            got2 = [np.sort(v2[o2[i]:o2[i + 1]]) for i in range(nq)]
            assert all(np.array_equal(x, y) for x, y in zip(got, got2))
            ring.close()


def test_units_argument_errors(gpu):
    cw, ck, cv, q, _ = tm.workload(5000, 10, seed=803)
    off = tm.zipf_runs(len(cw), 100, 1000, seed=5)
    fib, _ = tm.fib_from_runs(cw, ck, cv, off)
    ring = tm.Ring(fib, 10)
    with pytest.raises(tm.TagMatchError):
        ring.submit_units(q, [0], [0, 11])              # past the queries
    with pytest.raises(tm.TagMatchError):
        ring.submit_units(q, [len(off)], [0, 10])       # no such partition
    with pytest.raises(tm.TagMatchError):
        ring.submit_units(q, [0, 1], [0, 6, 5])         # not monotone
    ring.submit_units(q, [0, 1], [0, 4, 10])
    ring.wait()


def test_bucket_segments_do_not_change_results(gpu):
    """Work items cut by partition-size bucket (tm_ring_set_bucket_segments)
    and the tuner over them: every table gives the same results."""
    cw, ck, cv, _, _ = tm.workload(80000, 1, seed=811)
    off = tm.zipf_runs(len(cw), 100, 40000, seed=7)        # 2..625 blocks: buckets 0-3
    fib, _ = tm.fib_from_runs(cw, ck, cv, off)
    up, uo = tm.uneven_units(len(off) - 1, 1, 900, seed=8)
    nq = int(uo[-1])
    q, _ = tm.Generator(seed=812).queries(nq, len(cw))
    eo, ev = expected(cw, ck, cv, off, q, up, uo, 0)
    ring = tm.Ring(fib, nq, mode=tm.MODE_KEYS_UNIQUE)
    assert not ring.bucket_segments().any()
    ring.submit_units(q, up, uo)
    ring.wait()
    o, v = ring.results()
    assert np.array_equal(o, eo) and np.array_equal(v, ev)
    for table in ([1] * 6, [1 << 20] * 6, [3, 7, 20, 100, 0, 5], [2]):
        ring.set_bucket_segments(table)
        assert list(ring.bucket_segments()) == (table + [0] * 6)[:6]
        ring.submit_units(None, up, uo)
        ring.wait()
        o2, v2 = ring.results()
        assert np.array_equal(o2, eo) and np.array_equal(v2, ev), table
    ring.set_bucket_segments([])
    base, best, table = ring.autotune_buckets(up, uo, candidates=(2, 8, 64), reps=1)
    log(f"bucket table {list(table)}: {base:.3f} -> {best:.3f} ms")
    assert best <= base and all(int(x) in (0, 2, 8, 64) for x in table)
    ring.submit_units(None, up, uo)
    ring.wait()
    o2, v2 = ring.results()
    assert np.array_equal(o2, eo) and np.array_equal(v2, ev)
    with pytest.raises(tm.TagMatchError):
        ring.set_bucket_segments([1] * 7)
    ring.close()
