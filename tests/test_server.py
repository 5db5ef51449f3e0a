"""The serving loop (tm_server): each request's result is the same alone and
in the mix, admission under a tight HBM budget, and inserts ordered with
requests."""
import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def table(gpu):
    cw, ck, cv, q, _ = tm.workload(100000, 30000, seed=601)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 5000)
    return cw, ck, cv, q, fib


def request_sizes(rng, n, total):
    s = np.exp(rng.uniform(0, np.log(2000), n)).astype(int) + 1
    return s[np.cumsum(s) <= total]


@pytest.mark.parametrize("mode", [tm.MODE_KEYS_UNIQUE, tm.MODE_HITS])
def test_same_alone_and_in_mix(table, mode):
    cw, ck, cv, q, fib = table
    rng = np.random.default_rng(7)
    sizes = request_sizes(rng, 80, len(q))
    bounds = np.concatenate([[0], np.cumsum(sizes)])
    srv = tm.Server(fib, n_slots=4, slot_queries=8192, mode=mode)
    ids = [srv.submit(q[bounds[i]:bounds[i + 1]]) for i in range(len(sizes))]
    got = []
    for rid in ids:
        assert srv.wait(rid, 60000.0)
        got.append(srv.result(rid))
    st = srv.stats()
    srv.close()
    log(f"mode {mode}: {len(sizes)} requests in {st['n_batches']} steps, "
        f"{st['requests_per_batch_sum'] / max(st['n_batches'], 1):.2f} requests/step, "
        f"max latency {st['latency_max_ms']:.1f} ms")
    assert st["n_requests_done"] == len(sizes)
    assert st["n_batches"] < len(sizes)          # requests were batched together
    for i, (info, off, vals) in enumerate(got):
        qa = q[bounds[i]:bounds[i + 1]]
        eo, ev, _ = tm.match(fib, qa, mode=mode)
        assert info.status == 0 and info.n_queries == len(qa)
        assert info.latency_ms >= info.queue_ms >= 0
        assert np.array_equal(off, eo), i
        if mode == tm.MODE_KEYS_UNIQUE:
            assert np.array_equal(vals, ev), i
        else:
            for j in range(len(qa)):
                assert np.array_equal(np.sort(vals[off[j]:off[j + 1]]), np.sort(ev[eo[j]:eo[j + 1]])), (i, j)
    # and the oracle on a sample of requests
    for i in (0, len(sizes) // 2, len(sizes) - 1):
        qa = q[bounds[i]:bounds[i + 1]]
        eo, ev = orc.match(cw, ck, cv, qa, mode=0 if mode == tm.MODE_KEYS_UNIQUE else 2)
        info, off, vals = got[i]
        assert np.array_equal(off, eo)


def test_tight_budget_admits_one_at_a_time(table):
    cw, ck, cv, q, fib = table
    srv = tm.Server(fib, n_slots=4, slot_queries=4096, hbm_budget=1)
    ids = [srv.submit(q[i * 500:(i + 1) * 500]) for i in range(12)]
    res = []
    for rid in ids:
        assert srv.wait(rid, 60000.0)
        res.append(srv.result(rid))
    st = srv.stats()
    srv.close()
    log(f"budget 1 byte: {st['n_batches']} steps, {st['n_budget_waits']} deferred admissions")
    assert st["n_batches"] == 12                 # one request per step, one step in flight
    assert st["n_budget_waits"] > 0
    eo, ev, _ = tm.match(fib, q[:6000])
    for i, (info, off, vals) in enumerate(res):
        a, b = eo[i * 500], eo[(i + 1) * 500]
        assert np.array_equal(off, eo[i * 500:(i + 1) * 500 + 1] - a)
        assert np.array_equal(vals, ev[a:b])


def test_inserts_ordered_with_requests(gpu):
    cw, ck, cv, q, _ = tm.workload(40000, 600, seed=602)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 4000)
    g = tm.Generator(seed=603)
    bw, bko, bk = g.filters(3000)
    srv = tm.Server(fib, n_slots=3, slot_queries=1024)
    before = srv.submit(q)
    ins = srv.insert(bw, bko, bk, maxp=4000)
    after = srv.submit(q)
    for rid in (before, ins, after):
        assert srv.wait(rid, 60000.0)
    _, o1, v1 = srv.result(before)
    assert srv.info(ins).status == 0
    _, o2, v2 = srv.result(after)
    st = srv.stats()
    srv.close()
    assert st["n_inserts"] == 1
    eo, ev = orc.match(cw, ck, cv, q, mode=0)
    assert np.array_equal(o1, eo) and np.array_equal(v1, ev)
    w = np.concatenate([cw, bw])
    ko = np.concatenate([ck, ck[-1] + bko[1:]]).astype(np.uint64)
    kv = np.concatenate([cv, bk])
    eo2, ev2 = orc.match(w, ko, kv, q, mode=0)
    assert np.array_equal(o2, eo2) and np.array_equal(v2, ev2)
    assert not np.array_equal(o1, o2)            # the insert changed some results


def test_latency_flush(table):
    """flush_queries / flush_us: a free slot waits until enough queries are
    waiting (one step for ten requests), or until the oldest request has
    waited the latency limit, or an insert is queued behind the requests."""
    import time
    cw, ck, cv, q, fib = table
    eo, ev, _ = tm.match(fib, q[:640])
    srv = tm.Server(fib, n_slots=2, slot_queries=4096, flush_queries=640, flush_us=5_000_000)
    ids = [srv.submit(q[i * 64:(i + 1) * 64]) for i in range(10)]
    res = []
    for rid in ids:
        assert srv.wait(rid, 60000.0)
        res.append(srv.result(rid))
    st = srv.stats()
    log(f"flush at 640 queries: {st['n_batches']} steps for {st['n_requests']} requests")
    assert st["n_batches"] == 1 and st["requests_per_batch_sum"] == 10
    for i, (info, off, vals) in enumerate(res):
        a, b = eo[i * 64], eo[(i + 1) * 64]
        assert info.step_queries == 640
        assert np.array_equal(off, eo[i * 64:(i + 1) * 64 + 1] - a) and np.array_equal(vals, ev[a:b])
    srv.close()

    # 1.5 x flush_queries over two slots: the first slot takes a full step,
    # what it leaves behind waits for the latency limit in the second
    srv = tm.Server(fib, n_slots=2, slot_queries=640, flush_queries=640, flush_us=200_000)
    ids = [srv.submit(q[i * 64:(i + 1) * 64]) for i in range(15)]
    eo15, ev15, _ = tm.match(fib, q[:960])
    for i, rid in enumerate(ids):
        assert srv.wait(rid, 60000.0)
        info, off, vals = srv.result(rid)
        a, b = eo15[i * 64], eo15[(i + 1) * 64]
        assert np.array_equal(off, eo15[i * 64:(i + 1) * 64 + 1] - a) and np.array_equal(vals, ev15[a:b])
        if i >= 10:
            assert info.step_queries == 320 and info.queue_ms >= 180.0, (i, info.step_queries, info.queue_ms)
    st = srv.stats()
    log(f"1.5 x flush_queries: {st['n_batches']} steps")
    assert st["n_batches"] == 2
    srv.close()

    # below the threshold: the step starts at the latency limit
    srv = tm.Server(fib, n_slots=2, slot_queries=4096, flush_queries=4096, flush_us=50_000)
    t = time.time()
    rid = srv.submit(q[:64])
    assert srv.wait(rid, 60000.0)
    dt = time.time() - t
    info, off, vals = srv.result(rid)
    log(f"latency limit 50 ms: waited {info.queue_ms:.1f} ms in the queue, {dt * 1e3:.1f} ms in all")
    assert info.queue_ms >= 49.0 and dt < 5.0
    assert np.array_equal(off, eo[:65]) and np.array_equal(vals, ev[:eo[64]])

    # an insert behind a held request flushes it
    fib2, *_ = tm.build_fib(cw, ck, cv, 5000)      # the insert changes this table only
    srv2 = tm.Server(fib2, n_slots=2, slot_queries=4096, flush_queries=4096, flush_us=30_000_000)
    g = tm.Generator(seed=604)
    bw, bko, bk = g.filters(10)
    t = time.time()
    rid = srv2.submit(q[:64])
    iid = srv2.insert(bw, bko, bk)
    assert srv2.wait(rid, 60000.0) and srv2.wait(iid, 60000.0)
    assert time.time() - t < 10.0
    info, off, vals = srv2.result(rid)
    assert np.array_equal(off, eo[:65]) and np.array_equal(vals, ev[:eo[64]])
    srv2.close()
    srv.close()
