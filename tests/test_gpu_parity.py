"""Parity of the HIP engine with the CPU oracle (bit-exact), on the GPU.

Every test calls the engine through the C ABI (libtagmatch_amd.so) and
compares with the oracle on the same seeded inputs, or with the reference's
own outputs (tests/golden)."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu

MODES = {tm.MODE_KEYS_UNIQUE: 0, tm.MODE_KEYS_ALL: 1, tm.MODE_HITS: 2}


def log(msg):
    print(msg, flush=True)


def lists(off, vals):
    return [vals[int(off[i]):int(off[i + 1])] for i in range(len(off) - 1)]


def assert_same(got_off, got_vals, exp_off, exp_vals, sort_got, what=""):
    g = lists(got_off, got_vals)
    e = lists(exp_off, exp_vals)
    assert len(g) == len(e), what
    for i, (x, y) in enumerate(zip(g, e)):
        x = np.sort(x) if sort_got else x
        if not np.array_equal(x, y):
            raise AssertionError(f"{what}: query {i}: {len(x)} values vs {len(y)} expected")


def run_all_modes(fib, words, key_off, keys, queries, schedules=(tm.SCHED_PERSISTENT, tm.SCHED_PLAIN), **kw):
    exp = {m: orc.match(words, key_off, keys, queries, mode=o) for m, o in MODES.items()}
    for sched in schedules:
        for mode, (eo, ev) in exp.items():
            log(f"  mode={mode} schedule={sched} nq={len(queries)} {kw}")
            off, vals, st = tm.match(fib, queries, mode=mode, schedule=sched, **kw)
            # KEYS_UNIQUE comes sorted; HITS / KEYS_ALL are unordered per query
            assert_same(off, vals, eo, ev, sort_got=(mode != tm.MODE_KEYS_UNIQUE), what=f"mode {mode} sched {sched}")
            if mode == tm.MODE_HITS:
                assert st["n_hits"] == int(eo[-1])


# ----------------------------------------------------------------- partitioner

@pytest.mark.parametrize("n,maxp", [(1, 10), (50, 10), (3000, 64), (20000, 1000), (60000, 5000), (20000, 30000)])
def test_partition_bit_exact_vs_oracle(gpu, n, maxp):
    log(f"partition n={n} maxp={maxp}")
    cw, _, _, _, _ = tm.workload(n * 2, 1, seed=1000 + n)
    cw = cw[:n]
    perm, off, masks = tm.partition(cw, maxp)
    rc, operm, ooff, omasks = orc.partition(cw, maxp)
    assert rc == 0
    assert np.array_equal(perm, operm)
    assert np.array_equal(off, ooff)
    assert np.array_equal(masks, omasks)


def test_partition_vs_reference_fixture(gpu, golden):
    g = golden.partition
    for maxp in (64, 500, 2000, 10000):
        log(f"partition fixture maxp={maxp}")
        perm, off, masks = tm.partition(g["words"], maxp)
        assert np.array_equal(perm, g[f"order_{maxp}"])
        pid = g[f"pid_{maxp}"]
        starts = sorted({int(np.nonzero(pid == p)[0].min()) for p in np.unique(pid)})
        ours = sorted({int(off[i]) for i in range(len(off) - 1) if off[i + 1] > off[i]})
        assert ours == starts


def test_partition_reference_test_cases(gpu, golden):
    v = golden.vectors
    Fp = v["F"][:10] + [v["F10_partitioning"]]
    tables = {"small_fib_no_part": v["Q"][:3], "small_fib_part": Fp[:3], "large_fib_part": Fp,
              "large_fib_permutation": Fp + v["Q"]}
    for name, strings in tables.items():
        case = v["partitioning"][name]
        words = np.stack([orc.from_string(s) for s in strings])
        perm, off, masks = tm.partition(words, case["maxp"])
        assert len(off) - 1 == case["n_partitions"], name
        groups = sorted(sorted(int(x) for x in perm[off[i]:off[i + 1]]) for i in range(len(off) - 1))
        assert groups == case["groups"], name
        assert list(map(int, perm)) == case["order"], name


def test_partition_all_zero_filters(gpu):
    # a canonical table holds at most one all-zero filter; the reference's
    # nonzero-mask loop never ends on it, the engine gives it a zero mask
    words = np.zeros((5, 6), np.uint32)
    words[1, 0] = 1
    words[2, 2] = 4
    words[3, 1] = 1 << 8
    words[4, 0] = 3
    for maxp in (2, 10):
        perm, off, masks = tm.partition(words, maxp)
        assert np.array_equal(np.sort(perm), np.arange(5))
        zero_part = [i for i in range(len(off) - 1) if 0 in perm[off[i]:off[i + 1]]]
        assert len(zero_part) == 1 and not masks[zero_part[0]].any()
        for i in range(len(off) - 1):
            assert off[i + 1] - off[i] <= maxp
    perm, off, masks = tm.partition(words, 2)
    # every filter in exactly one partition; the zero filters share a zero mask
    assert np.array_equal(np.sort(perm), np.arange(5))
    for i in range(len(off) - 1):
        m = np.bitwise_and.reduce(words[perm[off[i]:off[i + 1]]], axis=0) if off[i + 1] > off[i] else np.full(6, 0xFFFFFFFF, np.uint32)
        assert np.array_equal(m, masks[i])


# ----------------------------------------------------------------- matcher

def test_match_golden_reference_outputs(gpu, golden):
    g = golden.match
    fib, perm, off, masks = tm.build_fib(g["words"], g["key_off"], g["keys"], 256)
    for sched in (tm.SCHED_PERSISTENT, tm.SCHED_PLAIN):
        for mode, name in ((tm.MODE_KEYS_UNIQUE, "unique"), (tm.MODE_KEYS_ALL, "all"), (tm.MODE_HITS, "hits")):
            log(f"golden sched={sched} mode={name}")
            o, v, st = tm.match(fib, g["queries"], mode=mode, schedule=sched)
            assert_same(o, v, g[name + "_off"], g[name + "_vals"], sort_got=(mode != tm.MODE_KEYS_UNIQUE),
                        what=f"golden {name} {sched}")


def test_match_reference_test_vectors(gpu, golden):
    v = golden.vectors
    F = np.stack([orc.from_string(s) for s in v["F"]])
    Q = np.stack([orc.from_string(s) for s in v["Q"]])
    ko = np.arange(len(F) + 1, dtype=np.uint64)
    keys = np.arange(len(F), dtype=np.uint32)
    cw, cko, ck = tm.canonicalize(F, ko, keys)
    for maxp in (1, 3, 100):
        fib, *_ = tm.build_fib(cw, cko, ck, maxp)
        off, vals, st = tm.match(fib, Q)
        assert [list(map(int, x)) for x in lists(off, vals)] == v["misc_subsets_keys"]


@pytest.mark.parametrize("n,nq,maxp", [(30000, 1500, 2000), (30000, 700, 64 * 1024), (5000, 333, 100)])
def test_match_generated_vs_oracle(gpu, n, nq, maxp):
    log(f"generated n={n} nq={nq} maxp={maxp}")
    cw, cko, ck, q, s = tm.workload(n, nq, seed=n + nq + maxp)
    fib, *_ = tm.build_fib(cw, cko, ck, maxp)
    run_all_modes(fib, cw, cko, ck, q)


@pytest.mark.parametrize("batch", [64, 128, 256])
def test_batch_sizes_and_chunks(gpu, batch):
    cw, cko, ck, q, s = tm.workload(20000, 900, seed=batch)
    fib, *_ = tm.build_fib(cw, cko, ck, 3000)
    eo, ev = orc.match(cw, cko, ck, q, mode=0)
    for sched in (tm.SCHED_PERSISTENT, tm.SCHED_PLAIN):
        for chunk in (1, 7, 256):
            log(f"batch={batch} sched={sched} chunk={chunk}")
            o, v, st = tm.match(fib, q, schedule=sched, batch_size=batch, chunk_blocks=chunk)
            assert_same(o, v, eo, ev, sort_got=False, what=f"b{batch} s{sched} c{chunk}")


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6, 7, 10, 18])
def test_inner_loop_variants_vs_oracle(gpu, variant):
    """Every persistent inner-loop variant (scalar or LDS query words, key
    log, early exit), large submissions (one tile per item) and small ones
    (block segments), all output modes."""
    cw, cko, ck, q, s = tm.workload(30000, 3000, seed=910 + variant)
    fib, *_ = tm.build_fib(cw, cko, ck, 2500)
    for nq in (3000, 37):
        log(f"variant={variant} nq={nq}")
        run_all_modes(fib, cw, cko, ck, q[:nq], schedules=(tm.SCHED_PERSISTENT,), variant=variant,
                      batch_size=64 if variant % 2 else 128)


def test_long_key_lists(gpu):
    """Filters with 17 to 300 keys (overflow lists that the whole wave walks
    together in KEYS_UNIQUE, rich filters from 32 keys) in every output mode,
    one segment and 8-block segments, against the oracle."""
    cw, cko, ck = tm.workload(40000, 1, seed=981)[:3]
    rng = np.random.default_rng(982)
    pool = np.unique(rng.integers(0, 2**31, 5000).astype(np.uint32))
    klists = []
    for i in range(len(cw)):
        n = int(rng.integers(17, 301)) if i % 5 == 0 else int(rng.integers(1, 5))
        klists.append(np.sort(rng.choice(pool, n, replace=False)))
    ko = np.zeros(len(cw) + 1, np.uint64)
    np.cumsum([len(k) for k in klists], out=ko[1:])
    keys = np.concatenate(klists).astype(np.uint32)
    q = rng.integers(0, 2**32, size=(1500, 6), dtype=np.uint64).astype(np.uint32)
    fib, *_ = tm.build_fib(cw, ko, keys, 5000)
    for mode in (0, 1, 2):
        eo, ev = orc.match(cw, ko, keys, q, mode=mode)
        for seg in (1 << 30, 8):
            for fe in ((1, 2) if mode == 0 else (2,)):   # 1: rich-filter pass and exit
                o, v, st = tm.match(fib, q, mode=mode, chunk_blocks=seg, full_exit=fe)
                assert_same(o, v, eo, ev, sort_got=mode != 0, what=f"mode {mode} segment {seg} full_exit {fe}")


def test_rich_filter_pass(gpu):
    """Rich filters (the most keys) fill the rows of the queries that
    contain them before the match; a filter whose key set is the whole
    dictionary answers such a query outright.  Results must equal the
    oracle's with the pass (one segment and 8-block segments) and without it
    (full_exit 2); the pass must also cut the hits processed."""
    cw, cko, ck = tm.workload(60000, 1, seed=971)[:3]
    K = np.unique(ck)
    # two complete filters with 3 and 5 bits, the rest unchanged
    extra = np.zeros((2, 6), np.uint32)
    extra[0, 0] = 0b111
    extra[1, 2] = 0b11111 << 7
    words = np.concatenate([cw, extra])
    klists = [ck[cko[i]:cko[i + 1]] for i in range(len(cw))] + [K, K[::-1]]
    ko = np.zeros(len(words) + 1, np.uint64)
    np.cumsum([len(k) for k in klists], out=ko[1:])
    keys = np.concatenate(klists).astype(np.uint32)
    cw2, cko2, ck2 = tm.canonicalize(words, ko, keys)
    rng = np.random.default_rng(972)
    q = rng.integers(0, 2**32, size=(3000, 6), dtype=np.uint64).astype(np.uint32)
    q[::3, 0] |= 0b111                   # a third contain the first complete filter
    q[1::7, 2] |= 0b11111 << 7
    fib, *_ = tm.build_fib(cw2, cko2, ck2, 8000)
    eo, ev = orc.match(cw2, cko2, ck2, q, mode=0)
    for seg in (1 << 30, 8):
        res = {}
        for fe in (1, 2):
            o, v, st = tm.match(fib, q, full_exit=fe, chunk_blocks=seg)
            assert_same(o, v, eo, ev, sort_got=False, what=f"full_exit {fe} segment {seg}")
            res[fe] = st["n_hits"]
        log(f"rich pass, segment {seg}: hits processed {res[1]} with it, {res[2]} without")
        assert res[1] < res[2]
    # a step large enough to be compacted: only the queries whose rows the
    # pass left incomplete are matched, rows and flags through the list
    q2 = rng.integers(0, 2**32, size=(20000, 6), dtype=np.uint64).astype(np.uint32)
    q2[::2, 0] |= 0b111
    eo2, ev2 = orc.match(cw2, cko2, ck2, q2, mode=0)
    for seg in (0, 8):
        o, v, st = tm.match(fib, q2, chunk_blocks=seg)
        assert_same(o, v, eo2, ev2, sort_got=False, what=f"compacted step, segment {seg}")


@pytest.mark.parametrize("n_keys", [5, 37, 64, 100])
def test_full_key_set_exit(gpu, n_keys):
    """With few distinct keys most queries collect every key early; the exit
    that stops testing such queries must not change any result (key counts
    around word boundaries of the bitmap row: 5, 37, 64, 100)."""
    cw, cko, ck, q, s = tm.workload(150000, 2000, seed=950 + n_keys)
    ck = (ck % n_keys).astype(np.uint32) * 7 + 3          # n_keys distinct keys
    # keys per filter must stay distinct for the canonical table
    key_lists = [np.unique(ck[cko[i]:cko[i + 1]]) for i in range(len(cw))]
    cko = np.zeros(len(cw) + 1, np.uint64)
    np.cumsum([len(k) for k in key_lists], out=cko[1:])
    ck = np.concatenate(key_lists).astype(np.uint32)
    fib, *_ = tm.build_fib(cw, cko, ck, 10000)
    eo, ev = orc.match(cw, cko, ck, q, mode=0)
    # one segment (bitmaps accumulate in LDS), and 8-block segments (the
    # rows other segments merged into HBM count too)
    for seg in (1 << 30, 8):
        for var in (2, 18):
            res = {}
            for fe in (1, 2):
                o, v, st = tm.match(fib, q, full_exit=fe, chunk_blocks=seg, variant=var)
                assert_same(o, v, eo, ev, sort_got=False, what=f"full_exit {fe} segment {seg} variant {var}")
                res[fe] = st["n_hits"]
            log(f"n_keys {n_keys} segment {seg} variant {var}: hits processed {res[1]} with the exit, "
                f"{res[2]} without")
            assert res[1] < res[2]


def bitf(*bits):
    x = np.zeros(6, np.uint32)
    for b in bits:
        x[b // 32] |= np.uint32(1 << (b % 32))
    return x


def test_edge_cases(gpu):
    ones = np.full(6, 0xFFFFFFFF, np.uint32)
    zero = np.zeros(6, np.uint32)
    # single-bit filters and nested prefixes: ALL_ONES finds all 192
    single = np.stack([bitf(i) for i in range(192)])
    nested = np.stack([bitf(*range(i + 1)) for i in range(192)])
    for tab in (single, nested):
        ko = np.arange(193, dtype=np.uint64)
        cw, cko, ck = tm.canonicalize(tab, ko, np.arange(192, dtype=np.uint32))
        for maxp in (1, 5, 1000):
            fib, *_ = tm.build_fib(cw, cko, ck, maxp)
            off, vals, st = tm.match(fib, np.stack([ones, zero, bitf(0), bitf(191)]))
            got = lists(off, vals)
            exp_off, exp_vals = orc.match(cw, cko, ck, np.stack([ones, zero, bitf(0), bitf(191)]), mode=0)
            assert_same(off, vals, exp_off, exp_vals, sort_got=False, what="edge")
            assert list(got[0]) == list(range(192))
    # an empty batch
    cw, cko, ck, q, s = tm.workload(2000, 5, seed=3)
    fib, *_ = tm.build_fib(cw, cko, ck, 100)
    off, vals, st = tm.match(fib, np.zeros((0, 6), np.uint32))
    assert len(off) == 1 and off[0] == 0


def test_all_zero_filter_and_overflow_keys(gpu):
    # an all-zero filter (matches every query), filters with > 4 keys, keys
    # at the ends of the uint32 range
    rng = np.random.default_rng(8)
    n = 3000
    words = (rng.integers(0, 2**32, (n, 6), dtype=np.uint64) & rng.integers(0, 2**32, (n, 6), dtype=np.uint64)
             & rng.integers(0, 2**32, (n, 6), dtype=np.uint64) & rng.integers(0, 2**32, (n, 6), dtype=np.uint64)).astype(np.uint32)
    words[0] = 0
    cnt = rng.integers(1, 12, n)
    ko = np.zeros(n + 1, np.uint64)
    ko[1:] = np.cumsum(cnt)
    keys = rng.integers(0, 300, int(ko[-1])).astype(np.uint32)
    keys[:5] = [0, 0xFFFFFFFF, 0xFFFFFFFE, 0xFFFF, 0x10000]
    cw, cko, ck = tm.canonicalize(words, ko, keys)
    q = (rng.integers(0, 2**32, (500, 6), dtype=np.uint64) | rng.integers(0, 2**32, (500, 6), dtype=np.uint64)).astype(np.uint32)
    q[0] = 0
    q[1] = 0xFFFFFFFF
    for maxp in (200, 5000):
        fib, *_ = tm.build_fib(cw, cko, ck, maxp)
        run_all_modes(fib, cw, cko, ck, q)


def test_many_distinct_keys(gpu):
    # 70000 distinct keys: KEYS_UNIQUE is refused (bitmap limit 65536);
    # KEYS_ALL and HITS use the overflow key lists
    rng = np.random.default_rng(9)
    cw, cko, _, q, s = tm.workload(40000, 300, seed=44)
    nk = int(cko[-1])
    keys = (rng.permutation(400000)[:nk] * 7 + 1).astype(np.uint32)
    for i in range(len(cw)):      # keys must be distinct within a filter
        seg = keys[int(cko[i]):int(cko[i + 1])]
        keys[int(cko[i]):int(cko[i + 1])] = np.sort(seg)
    fib, *_ = tm.build_fib(cw, cko, keys, 4000)
    assert fib.stats()["n_distinct_keys"] > 65536
    with pytest.raises(tm.TagMatchError):
        tm.match(fib, q, mode=tm.MODE_KEYS_UNIQUE)
    for mode in (tm.MODE_KEYS_ALL, tm.MODE_HITS):
        eo, ev = orc.match(cw, cko, keys, q, mode=MODES[mode])
        o, v, st = tm.match(fib, q, mode=mode)
        assert_same(o, v, eo, ev, sort_got=True, what=f"many keys {mode}")


def test_overflow_retry_small_capacity(gpu):
    cw, cko, ck, q, s = tm.workload(30000, 800, seed=12)
    fib, *_ = tm.build_fib(cw, cko, ck, 2000)
    eo, ev = orc.match(cw, cko, ck, q, mode=2)
    o, v, st = tm.match(fib, q, mode=tm.MODE_HITS, hit_capacity=1000)
    assert st["overflow"] == 1 and st["hit_capacity"] >= int(eo[-1])
    assert_same(o, v, eo, ev, sort_got=True, what="overflow retry")


def test_ring_reuse_graph_replay(gpu):
    cw, cko, ck, q, s = tm.workload(40000, 3000, seed=21)
    fib, *_ = tm.build_fib(cw, cko, ck, 4000)
    eo, ev = orc.match(cw, cko, ck, q, mode=0)
    ring = tm.Ring(fib, 3000, use_graph=1)
    eager = tm.Ring(fib, 3000)
    for rep in range(4):
        for part in (slice(0, 3000), slice(0, 1234), slice(1000, 3000)):
            log(f"graph rep={rep} part={part}")
            sub = q[part]
            ring.submit(sub)
            ring.wait()
            o, v = ring.results()
            eager.submit(sub)        # eager work between replays
            eager.wait()
            o2, v2 = eager.results()
            assert np.array_equal(o, o2) and np.array_equal(v, v2)
            lo = np.asarray(eo[part.start:part.stop + 1] - eo[part.start], np.uint64)
            assert_same(o, v, lo, ev[int(eo[part.start]):int(eo[part.stop])], sort_got=False, what="graph")
    ring.close()
    eager.close()


def test_resident_queries_and_async_rings(gpu):
    cw, cko, ck, q, s = tm.workload(40000, 4000, seed=31)
    fib, *_ = tm.build_fib(cw, cko, ck, 5000)
    eo, ev = orc.match(cw, cko, ck, q, mode=0)
    rings = [tm.Ring(fib, 1000) for _ in range(4)]
    for i, r in enumerate(rings):          # four submissions in flight
        r.submit(q[i * 1000:(i + 1) * 1000])
    for i, r in enumerate(rings):
        r.wait()
        o, v = r.results()
        lo = np.asarray(eo[i * 1000:(i + 1) * 1000 + 1] - eo[i * 1000], np.uint64)
        assert_same(o, v, lo, ev[int(eo[i * 1000]):int(eo[(i + 1) * 1000])], sort_got=False, what=f"ring {i}")
        r.close()


def test_larger_table_spot_check(gpu):
    # 500k filters: plain == persistent on all 5000 queries; a sample of 150
    # queries against the oracle; HITS count consistent with KEYS_UNIQUE
    cw, cko, ck, q, s = tm.workload(500000, 5000, seed=0x7461676D61746302)
    fib, *_ = tm.build_fib(cw, cko, ck, 64 * 1024)
    o1, v1, st1 = tm.match(fib, q, schedule=tm.SCHED_PERSISTENT)
    o2, v2, st2 = tm.match(fib, q, schedule=tm.SCHED_PLAIN)
    assert np.array_equal(o1, o2) and np.array_equal(v1, v2)
    oh, vh, sth = tm.match(fib, q[:150], mode=tm.MODE_HITS)
    eo, ev = orc.match(cw, cko, ck, q[:150], mode=2)
    assert_same(oh, vh, eo, ev, sort_got=True, what="hits sample")
    ku, kv = orc.match(cw, cko, ck, q[:150], mode=0)
    assert_same(o1[:151], v1[:int(o1[150])], ku, kv, sort_got=False, what="keys sample")
    # the default persistent step stops testing queries whose key set is
    # complete; without that exit it counts every match, as the plain form
    o3, v3, st3 = tm.match(fib, q, schedule=tm.SCHED_PERSISTENT, full_exit=2)
    assert np.array_equal(o1, o3) and np.array_equal(v1, v3)
    assert st3["n_hits"] == st2["n_hits"] >= st1["n_hits"]


def test_cpp_api_reference_cases(gpu, golden):
    v = golden.vectors
    path = os.path.join(tempfile.mkdtemp(), "vectors.txt")
    with open(path, "w") as f:
        for s in v["F"]:
            f.write(f"F {s}\n")
        for s in v["Q"]:
            f.write(f"Q {s}\n")
    exe = os.path.join(os.path.dirname(tm.LIB_PATH), "tm_api_tests")
    r = subprocess.run([exe, path], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


@pytest.mark.parametrize("nq", [100, 20000])
def test_lookup_counts_vs_bruteforce(gpu, nq):
    """Per-partition query counts of the lookup (tm_ring_lookup_counts) and
    the (query, partition) pair count against a brute force over the
    partition masks: a small step, and a compacted one (counted by
    k_lookup_hist over every query), in each output mode and schedule."""
    cw, cko, ck, q, _ = tm.workload(500000, nq, seed=0x7461676D61746302)
    fib, perm, off, masks = tm.build_fib(cw, cko, ck, 64 * 1024)
    m = np.asarray(masks, np.uint32).reshape(-1, 6)
    exp = ((m[None, :, :] & ~q[:, None, :]) == 0).all(axis=2).sum(axis=0)
    for mode in (tm.MODE_KEYS_UNIQUE, tm.MODE_KEY_BITMAP, tm.MODE_HITS):
        for sched in (tm.SCHED_PERSISTENT, tm.SCHED_PLAIN):
            ring = tm.Ring(fib, nq, mode=mode, schedule=sched)
            ring.submit(q)
            st = ring.wait()
            got = ring.lookup_counts()
            assert np.array_equal(got[:len(exp)], exp), (mode, sched)
            # the spare partition(s) after the table's have an empty mask:
            # every query selects them
            assert np.all(got[len(exp):] == nq)
            assert st.n_pairs == int(got.sum()), (mode, sched)
            ring.close()
    fib.close()


def test_compact_tiles_threshold(gpu):
    """compact_tiles: steps of more tiles than the threshold match only the
    queries the rich-filter pass leaves incomplete.  A 2000-query step (32
    tiles, below the default 128) compacted and not, in both row modes,
    equals the oracle."""
    cw, cko, ck, q, _ = tm.workload(200000, 2000, seed=0x7461676D61746302)
    # a filter holding every key: the rich-filter pass (and compaction) is on
    K = np.unique(ck)
    extra = np.zeros((1, 6), np.uint32)
    extra[0, 4] = 0b1001101
    w = np.concatenate([cw, extra])
    ko = np.concatenate([cko, [cko[-1] + len(K)]]).astype(np.uint64)
    kv = np.concatenate([ck, K]).astype(np.uint32)
    cw, cko, ck = tm.canonicalize(w, ko, kv)
    fib, *_ = tm.build_fib(cw, cko, ck, 20000)
    eo, ev = orc.match(cw, cko, ck, q, mode=0)
    for ct in (0, 1, 1000):
        o, v, st = tm.match(fib, q, compact_tiles=ct)
        assert np.array_equal(o, eo) and np.array_equal(v, ev), ct
        ring = tm.Ring(fib, len(q), mode=tm.MODE_KEY_BITMAP, compact_tiles=ct)
        ring.submit(q)
        ring.wait()
        off, vals = tm.expand_key_bitmaps(ring.key_bitmaps(), fib.dictionary())
        assert np.array_equal(off, eo) and np.array_equal(vals, ev), ct
        ring.close()
