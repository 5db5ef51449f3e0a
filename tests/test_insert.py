"""Incremental inserts into a resident FIB (tm_fib_insert) against the
oracle run on the whole table: spare lanes, new partitions, new keys,
graphs captured before the insert, and the busy check."""
import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu

MODES = [(tm.MODE_KEYS_UNIQUE, 0), (tm.MODE_KEYS_ALL, 1), (tm.MODE_HITS, 2)]


def log(msg):
    print(msg, flush=True)


def check_all(fib, words, ko, kv, q, **kw):
    for mode, om in MODES:
        eo, ev = orc.match(words, ko, kv, q, mode=om)
        off, vals, st = tm.match(fib, q, mode=mode, **kw)
        assert np.array_equal(off, eo), (mode, kw)
        if mode == tm.MODE_KEYS_UNIQUE:
            assert np.array_equal(vals, ev), (mode, kw)
        else:
            for i in range(len(q)):
                assert np.array_equal(np.sort(vals[off[i]:off[i + 1]]), ev[eo[i]:eo[i + 1]]), (mode, i, kw)


def gen(n_filters, n_queries, seed):
    g = tm.Generator(seed=seed)
    w, ko, k = g.filters(n_filters)
    q, _ = g.queries(n_queries, n_filters)
    return w, ko, k, q


def concat(a_w, a_ko, a_k, b_w, b_ko, b_k):
    w = np.concatenate([a_w, b_w])
    ko = np.concatenate([a_ko, a_ko[-1] + b_ko[1:]]).astype(np.uint64)
    k = np.concatenate([a_k, b_k])
    return w, ko, k


@pytest.mark.parametrize("reserve", [13, 0])
def test_insert_matches_oracle(gpu, reserve):
    cw, ck, cv, q, _ = tm.workload(30000, 400, seed=401)
    bw, bko, bk, _ = gen(4000, 1, seed=402)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 3000, reserve=reserve)
    p0 = fib.stats()["n_parts"]
    log(f"reserve {reserve}: {p0} partitions, inserting {len(bw)}")
    fib.insert(bw, bko, bk, maxp=3000)
    st = fib.stats()
    log(f"after insert: {st['n_parts']} partitions, {st['n_filters']} filters")
    assert st["n_filters"] == len(cw) + len(bw)
    if reserve == 0:
        assert st["n_parts"] > p0
    w, ko, k = concat(cw, ck, cv, bw, bko, bk)
    for sched in (tm.SCHED_PERSISTENT, tm.SCHED_PLAIN):
        check_all(fib, w, ko, k, q, schedule=sched)


def test_insert_new_keys_and_overflow_lists(gpu):
    cw, ck, cv, q, _ = tm.workload(20000, 300, seed=403)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 2000)
    bw, _, _, _ = gen(500, 1, seed=404)
    rng = np.random.default_rng(5)
    counts = rng.integers(1, 9, len(bw))            # up to 8 keys: some overflow lists
    bko = np.zeros(len(bw) + 1, np.uint32)
    np.cumsum(counts, out=bko[1:])
    bk = np.concatenate([np.sort(rng.choice(np.arange(7_000_000, 7_000_040), c, replace=False))
                         for c in counts]).astype(np.uint32)   # keys unknown to the table
    k0 = fib.stats()["n_distinct_keys"]
    fib.insert(bw, bko, bk)
    assert fib.stats()["n_distinct_keys"] == k0 + len(np.unique(bk))
    w, ko, k = concat(cw, ck, cv, bw, bko, bk)
    check_all(fib, w, ko, k, q)


def test_ring_across_inserts_with_graph(gpu):
    cw, ck, cv, q, _ = tm.workload(30000, 500, seed=405)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 3000)
    ring = tm.Ring(fib, len(q), use_graph=1, batch_size=64)
    w, ko, k = cw, ck, cv
    for step in range(4):
        eo, ev = orc.match(w, ko, k, q, mode=0)
        for rep in range(2):
            ring.submit(q)
            ring.wait()
            o, v = ring.results()
            assert np.array_equal(o, eo) and np.array_equal(v, ev), (step, rep)
        bw, bko, bk, _ = gen(1500, 1, seed=500 + step)
        # a submission in flight blocks the insert
        ring.submit(q)
        with pytest.raises(tm.TagMatchError):
            fib.insert(bw, bko, bk, maxp=3000)
        ring.wait()
        fib.insert(bw, bko, bk, maxp=3000)
        w, ko, k = concat(w, ko, k, bw, bko, bk)
        log(f"step {step}: {fib.stats()['n_parts']} partitions")


def test_live_insert_is_seen_only_by_later_submissions(gpu):
    cw, ck, cv, q, _ = tm.workload(100000, 2000, seed=420)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 8000)
    _, _, _, big, _ = tm.workload(100000, 200000, seed=421)
    bw, bko, bk, _ = gen(3000, 1, seed=422)
    # lane epochs start with the first insert, made with nothing in flight
    b0w, b0ko, b0k, _ = gen(50, 1, seed=423)
    ring0 = tm.Ring(fib, len(q))
    ring0.submit(q)
    assert not fib.insert_live(b0w, b0ko, b0k)          # declined: a submission without epochs runs
    ring0.wait()
    ring0.close()
    assert fib.insert_live(b0w, b0ko, b0k)
    cw, ck, cv = concat(cw, ck, cv, b0w, b0ko, b0k)
    w2, ko2, k2 = concat(cw, ck, cv, bw, bko, bk)
    for mode, om in MODES:
        eo1, ev1 = orc.match(cw, ck, cv, q, mode=om)
        eo2, ev2 = orc.match(w2, ko2, k2, q, mode=om)
        ring = tm.Ring(fib, len(big), mode=mode)
        ring.submit(np.concatenate([q, big])[:len(big)])   # q first, then a long tail
        # the insert lands while that submission runs; it must not see it
        if mode == tm.MODE_KEYS_UNIQUE:
            assert fib.insert_live(bw, bko, bk)
            w_seen, ko_seen, k_seen = cw, ck, cv
        else:
            w_seen, ko_seen, k_seen = w2, ko2, k2          # inserted by the KEYS_UNIQUE round
        ring.wait()
        o, v = ring.results()
        nq = len(q)
        eo, ev = orc.match(w_seen, ko_seen, k_seen, q, mode=om)
        assert np.array_equal(o[:nq + 1], eo), mode
        for i in range(nq):
            assert np.array_equal(np.sort(v[o[i]:o[i + 1]]), ev[eo[i]:eo[i + 1]]), (mode, i)
        ring.close()
        # every later submission sees it
        off2, vals2, st = tm.match(fib, q, mode=mode)
        assert np.array_equal(off2, eo2), mode
        for i in range(len(q)):
            assert np.array_equal(np.sort(vals2[off2[i]:off2[i + 1]]), ev2[eo2[i]:eo2[i + 1]]), (mode, i)


def test_live_insert_declines_and_graph_follows_epochs(gpu):
    cw, ck, cv, q, _ = tm.workload(30000, 500, seed=430)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 3000)
    bw, bko, bk, _ = gen(200, 1, seed=431)
    # a key the table does not hold: declined, nothing changes
    bad = bk.copy()
    bad[0] = 0xFFFFFFF0
    n0 = fib.stats()["n_filters"]
    assert not fib.insert_live(bw, bko, bad)
    assert fib.stats()["n_filters"] == n0
    ring = tm.Ring(fib, len(q), use_graph=1)
    w, ko, k = cw, ck, cv
    for step in range(3):
        for rep in range(2):
            ring.submit(q)
            ring.wait()
            o, v = ring.results()
            eo, ev = orc.match(w, ko, k, q, mode=0)
            assert np.array_equal(o, eo) and np.array_equal(v, ev), (step, rep)
        bw, bko, bk, _ = gen(300, 1, seed=440 + step)
        assert fib.insert_live(bw, bko, bk)
        w, ko, k = concat(w, ko, k, bw, bko, bk)


def test_image_round_trip(gpu, tmp_path):
    """tm_fib_save / tm_fib_load: a table built, then grown by inserts (new
    keys and partitions), saved and loaded into a new FIB answers as the
    oracle on the whole table; the loaded FIB takes further inserts; rich
    filters travel in the image."""
    cw, ck, cv, q, _ = tm.workload(30000, 400, seed=411)
    fib, *_ = tm.build_fib(cw, ck, cv, 3000, reserve=0)
    bw, _, _, _ = gen(3000, 1, seed=412)
    rng = np.random.default_rng(413)
    counts = rng.integers(1, 9, len(bw))
    bko = np.zeros(len(bw) + 1, np.uint32)
    np.cumsum(counts, out=bko[1:])
    bk = np.concatenate([np.sort(rng.choice(np.arange(8_000_000, 8_000_050), c, replace=False))
                         for c in counts]).astype(np.uint32)
    fib.insert(bw, bko, bk, maxp=3000)
    w, ko, k = concat(cw, ck, cv, bw, bko, bk)
    path = tmp_path / "t.fib"
    fib.save(str(path))
    log(f"image: {path.stat().st_size} bytes")
    g = tm.Fib.from_image(str(path))
    a, b = fib.stats(), g.stats()
    for f in ("n_filters", "n_parts", "n_blocks", "n_keys", "n_distinct_keys", "max_part_size"):
        assert a[f] == b[f], f
    for sched in (tm.SCHED_PERSISTENT, tm.SCHED_PLAIN):
        check_all(g, w, ko, k, q, schedule=sched)
    # a loaded FIB grows like a built one
    cw3, ck3, cv3, _ = gen(2000, 1, seed=414)
    g.insert(cw3, ck3, cv3, maxp=3000)
    w2, ko2, k2 = concat(w, ko, k, cw3, ck3, cv3)
    check_all(g, w2, ko2, k2, q)
    # loading over a FIB replaces its table
    fib.load(str(path))
    check_all(fib, w, ko, k, q)

    # rich filters (every key of the table) are part of the image
    K = np.unique(cv)
    extra = np.zeros((1, 6), np.uint32)
    extra[0, 1] = 0b1011
    rw = np.concatenate([cw, extra])
    rko = np.concatenate([ck, [ck[-1] + len(K)]]).astype(np.uint64)
    rk = np.concatenate([cv, K]).astype(np.uint32)
    rw, rko, rk = tm.canonicalize(rw, rko, rk)
    rf, *_ = tm.build_fib(rw, rko, rk, 3000)
    rq = q.copy()
    rq[::2, 1] |= 0b1011
    rf.save(str(path))
    rg = tm.Fib.from_image(str(path))
    eo, ev = orc.match(rw, rko, rk, rq, mode=0)
    for fe in (1, 2):
        off, vals, _ = tm.match(rg, rq, full_exit=fe)
        assert np.array_equal(off, eo) and np.array_equal(vals, ev), fe


def test_image_errors(gpu, tmp_path):
    cw, ck, cv, q, _ = tm.workload(5000, 10, seed=415)
    fib, *_ = tm.build_fib(cw, ck, cv, 1000)
    p = tmp_path / "t.fib"
    fib.save(str(p))
    data = p.read_bytes()
    (tmp_path / "cut.fib").write_bytes(data[: len(data) // 2])
    (tmp_path / "bad.fib").write_bytes(b"x" * 200)
    v2 = bytearray(data)
    v2[8] = 99                                   # image version
    (tmp_path / "v99.fib").write_bytes(bytes(v2))
    for name, what in (("cut.fib", "truncated"), ("bad.fib", "not a FIB image"), ("v99.fib", "version"),
                       ("missing.fib", "cannot open")):
        with pytest.raises(tm.TagMatchError) as e:
            tm.Fib.from_image(str(tmp_path / name))
        assert e.value.code == tm.TM_ERR_INVALID and what in str(e.value), name
    # images whose sizes are right but whose indexes point outside the
    # table: refused before anything is uploaded
    arrays = {}
    pos = 88                                     # header: magic, version, 9 sizes
    for name, item in (("words", 4), ("info", 4), ("slots", 8), ("fid", 4), ("rich", 4), ("rich_bm", 4),
                       ("ovf", 4), ("dict", 4), ("pinfo", 4), ("pstat", 4), ("psize", 4), ("pcap", 4),
                       ("order", 4)):
        n = int.from_bytes(data[pos:pos + 8], "little")
        arrays[name] = (pos + 8, n)
        pos += 8 + n * item
    assert pos == len(data)
    n_blocks = arrays["info"][1] // 8

    def corrupt(name, index, value, item=4):
        b = bytearray(data)
        at = arrays[name][0] + index * item
        b[at:at + item] = int(value).to_bytes(item, "little")
        return bytes(b)

    order1 = int.from_bytes(data[arrays["order"][0] + 4:arrays["order"][0] + 8], "little")
    cases = {"order.fib": (corrupt("order", 0, order1), "bit order"),
             "part.fib": (corrupt("pinfo", 6, n_blocks + 5), "outside the table"),
             "slot.fib": (corrupt("slots", 0, 0xFFFFFFFFFFFFFFFE, 8), "key id out of range"),
             "ovf.fib": (corrupt("slots", 0, 0x0000FFFFFFFFFFFF, 8), "overflow list out of range")}
    for name, (blob, what) in cases.items():
        (tmp_path / name).write_bytes(blob)
        with pytest.raises(tm.TagMatchError) as e:
            tm.Fib.from_image(str(tmp_path / name))
        assert e.value.code == tm.TM_ERR_INVALID and what in str(e.value), name
    # a failed load leaves the FIB as it was
    for name in ("cut.fib", "part.fib"):
        with pytest.raises(tm.TagMatchError):
            fib.load(str(tmp_path / name))
    check_all(fib, cw, ck, cv, q)
    empty = tm.Fib.empty()
    with pytest.raises(tm.TagMatchError) as e:
        empty.save(str(tmp_path / "e.fib"))
    assert e.value.code == tm.TM_ERR_STATE


def test_inserts_reach_compacted_steps(gpu, tmp_path):
    """Compacted steps skip groups of 8 blocks by the AND of their masks.
    Filters inserted live, then at a barrier, then through a saved and
    loaded image are made to match queries whose key sets are still
    incomplete, each carrying a key its query lacks: every compacted step
    after an insert must find those keys (a group mask left stale would hide
    them), and equal the oracle on the grown table."""
    cw, ck, cv, _, _ = tm.workload(100000, 1, seed=440)
    K = np.unique(cv)
    # one filter holding every key, so that the rich-filter pass (and with
    # it compaction) is on
    extra = np.zeros((1, 6), np.uint32)
    extra[0, 2] = 0b110101
    w = np.concatenate([cw, extra])
    ko = np.concatenate([ck, [ck[-1] + len(K)]]).astype(np.uint64)
    kv = np.concatenate([cv, K]).astype(np.uint32)
    w, ko, kv = tm.canonicalize(w, ko, kv)
    fib, perm, off, masks = tm.build_fib(w, ko, kv, 8000)
    m = np.asarray(masks, np.uint32).reshape(-1, 6)
    q, _ = tm.Generator(seed=440).queries(20000, 100000, first=1 << 20)
    nq = len(q)                                   # 313 tiles: a compacted step
    rng = np.random.default_rng(441)

    def step():
        o, v, st = tm.match(fib, q, mode=tm.MODE_KEYS_UNIQUE)
        return o, v

    def batch(o, v, n):
        """n filters, each contained in a query whose key set is incomplete
        and carrying one key that query lacks."""
        cnt = np.diff(o.astype(np.int64))
        partial = np.nonzero(cnt < len(K))[0]
        assert len(partial) >= n, len(partial)
        fw, fk = [], []
        for i in rng.choice(partial, n, replace=False):
            have = v[int(o[i]):int(o[i + 1])]
            missing = np.setdiff1d(K, have)
            sel = np.nonzero(((m & ~q[i][None, :]) == 0).all(axis=1))[0]   # partitions the query selects
            f = m[rng.choice(sel)].copy()
            extra_bits = q[i] & rng.integers(0, 2**32, 6, dtype=np.uint64).astype(np.uint32)
            f |= extra_bits & rng.integers(0, 2**32, 6, dtype=np.uint64).astype(np.uint32)
            assert ((f & ~q[i]) == 0).all()
            fw.append(f)
            fk.append(int(rng.choice(missing)))
        fko = np.arange(n + 1, dtype=np.uint32)
        return np.array(fw, np.uint32), fko, np.array(fk, np.uint32)

    o, v = step()
    eo, ev = orc.match(w, ko, kv, q, mode=0)
    assert np.array_equal(o, eo) and np.array_equal(v, ev)
    # live, at a barrier, then through an image
    for how in ("live", "barrier", "image"):
        bw, bko, bk = batch(o, v, 150)
        if how == "live":
            assert fib.insert_live(bw, bko, bk)
        elif how == "barrier":
            fib.insert(bw, bko, bk, maxp=8000)
        w, ko, kv = concat(w, ko, kv, bw, bko, bk)
        if how == "image":
            path = tmp_path / "g.fib"
            fib.insert(bw, bko, bk, maxp=8000)
            fib.save(str(path))
            fib = tm.Fib.from_image(str(path))
        o, v = step()
        eo, ev = orc.match(w, ko, kv, q, mode=0)
        log(f"{how}: {int(o[-1])} keys, {int(eo[-1])} expected")
        assert np.array_equal(o, eo) and np.array_equal(v, ev), how
