"""The CPU oracle, pinned against the reference's own vectors and programs.

Each oracle function is checked against data the reference fixes: the
vectors and expected answers of its unit tests (tests/golden/ref_vectors.json),
outputs of the reference programs compiled from its sources
(tests/golden/*.npz, and oracle/_ref live when present), and invariants."""
import os
import tempfile

import numpy as np
import pytest

import oracle as orc


def w(s):
    return orc.from_string(s)


# ------------------------------------------------------------- bit layout / I/O

def test_binary_encoding_test_io():
    # reference test/test_io.cc:50-68: the 24-byte image of a filter
    zero = "0" * 192
    cases = [("1" + "0" * 191, {0: 0x01}), (zero, {}), ("0" * 191 + "1", {23: 0x80}),
             ("01" + "0" * 190, {0: 0x02}), ("0" * 9 + "1" + "0" * 182, {1: 0x02})]
    for s, nz in cases:
        b = w(s).tobytes()
        exp = bytearray(24)
        for k, v in nz.items():
            exp[k] = v
        assert b == bytes(exp), s


def test_bit_set_positions_test_bitvector():
    # reference test/test_bitvector.cc:92-136 (set_bit vs string positions)
    for bits in ((2, 10), (66, 74), (130, 138)):
        s = ["0"] * 192
        for b in bits:
            s[b] = "1"
        x = w("".join(s))
        y = np.zeros(6, np.uint32)
        for b in bits:
            y[b // 32] |= np.uint32(1 << (b % 32))
        assert np.array_equal(x, y)


def test_bitvector_cases(golden):
    # reference test/test_bitvector.cc:426-590, read as data
    cases = golden.vectors["bitvector_cases"]
    assert len(cases) >= 30
    for c in cases:
        a, b = w(c["a"]), w(c["b"])
        if c["fn"] == "subset_of":
            assert orc.subset(a, b) == c["expected"], c
        elif c["fn"] == "prefix_less_than":
            assert (orc.compare(a, b) < 0) == c["expected"], c
        elif c["fn"] == "leftmost_diff":
            assert orc.leftmost_diff(a, b) == c["expected"], c
        elif c["fn"] == "equals":
            assert (orc.compare(a, b) == 0) == c["expected"], c


def test_bitvector_pairs_vs_reference(golden):
    g = golden.bitvector
    a, b = g["a"], g["b"]
    n = a.shape[0]
    sub = np.array([orc.subset(a[i], b[i]) for i in range(n)])
    cmp_ = np.array([orc.compare(a[i], b[i]) for i in range(n)])
    lmd = np.array([orc.leftmost_diff(a[i], b[i]) for i in range(n)])
    assert np.array_equal(sub, g["subset"])
    assert np.array_equal(cmp_ < 0, g["less"])
    assert np.array_equal(cmp_ > 0, g["greater"])
    assert np.array_equal(lmd, g["leftmost_diff"])
    # memory image: uint32_value(j) of the reference == our word j
    assert np.array_equal(g["words"], a)
    pc = np.unpackbits(a.view(np.uint8), axis=1).sum(1)
    assert np.array_equal(pc, g["popcount"])
    # the pairs exercise both outcomes of each predicate
    assert g["subset"].any() and (~g["subset"]).any()
    assert g["less"].any() and g["greater"].any() and (lmd == 192).any()


def test_subset_truth_discriminates_planted_mistakes(golden):
    # A plausible mistake in the subset predicate (testing f & q instead of
    # f & ~q, or testing q within f) must disagree with the reference pairs.
    g = golden.bitvector
    a, b = g["a"].astype(np.uint64), g["b"].astype(np.uint64)
    wrong1 = np.all((a & b) == 0, axis=1)
    wrong2 = np.all((b & ~a & 0xFFFFFFFF) == 0, axis=1)
    assert not np.array_equal(wrong1, g["subset"])
    assert not np.array_equal(wrong2, g["subset"])
    right = np.all((a & ~b & 0xFFFFFFFF) == 0, axis=1)
    assert np.array_equal(right, g["subset"])


# ------------------------------------------------------------- matcher

def _lists(off, vals):
    return [vals[int(off[i]):int(off[i + 1])] for i in range(len(off) - 1)]


def test_match_vs_reference_test_matcher(golden):
    g = golden.match
    for mode, key in ((0, "unique"), (1, "all"), (2, "hits")):
        if mode == 2:
            off, vals = orc.match(g["words"], g["key_off"], g["keys"], g["queries"], mode=2)
        else:
            off, vals = orc.match(g["words"], g["key_off"], g["keys"], g["queries"], mode=mode)
        exp = _lists(g[key + "_off"], g[key + "_vals"])
        got = _lists(off, vals)
        assert len(got) == len(exp)
        for i, (x, y) in enumerate(zip(got, exp)):
            assert np.array_equal(np.sort(x), y), (mode, i)
    # the fixture is not vacuous: many queries match, not all keys everywhere
    n_hits = np.diff(g["hits_off"])
    assert (n_hits > 0).mean() > 0.5 and n_hits.max() > 10


def test_misc_subsets_known_answers(golden):
    # reference test/test_tagmatch.cc:484-534 and the Q1 case {2}
    v = golden.vectors
    F = np.stack([w(s) for s in v["F"]])
    Q = np.stack([w(s) for s in v["Q"]])
    off, vals = orc.match(F, np.arange(len(F) + 1, dtype=np.uint64), np.arange(len(F), dtype=np.uint32), Q, mode=0)
    got = [list(map(int, x)) for x in _lists(off, vals)]
    assert got == v["misc_subsets_keys"]
    assert got[1] == [2] and got[12] == list(range(11)) and got[0] == []


def test_match_corner_cases():
    # reference test/test_tagmatch.cc:536-638
    def f(*bits):
        x = np.zeros(6, np.uint32)
        for b in bits:
            x[b // 32] |= np.uint32(1 << (b % 32))
        return x
    ones = np.full(6, 0xFFFFFFFF, np.uint32)
    # corner_case_filters
    F = np.stack([f(0), f(191)])
    ko = np.arange(3, dtype=np.uint64)
    k = np.array([1, 2], np.uint32)
    Q = np.stack([f(1, 10, 100, 190), f(0, 10, 100, 190), f(1, 10, 100, 191), f(0, 10, 100, 191)])
    off, vals = orc.match(F, ko, k, Q, mode=0)
    assert [list(map(int, x)) for x in _lists(off, vals)] == [[], [1], [2], [1, 2]]
    # single_bit_filters and deepest_trie: ALL_ONES finds all 192 keys
    single = np.stack([f(i) for i in range(192)])
    nested = np.stack([f(*range(i + 1)) for i in range(192)])
    for tab in (single, nested):
        off, vals = orc.match(tab, np.arange(193, dtype=np.uint64), np.arange(192, dtype=np.uint32), ones[None], mode=0)
        assert list(vals) == list(range(192))
    # ALL_ZEROS finds nothing non-empty
    off, vals = orc.match(single, np.arange(193, dtype=np.uint64), np.arange(192, dtype=np.uint32),
                          np.zeros((1, 6), np.uint32), mode=0)
    assert off[-1] == 0


def test_match_brute_force_numpy():
    # closed form keys(q) = sorted(unique(U keys(f) : f subset q))
    rng = np.random.default_rng(1)
    F = (rng.integers(0, 2**32, (300, 6), dtype=np.uint64) & rng.integers(0, 2**32, (300, 6), dtype=np.uint64)
         & rng.integers(0, 2**32, (300, 6), dtype=np.uint64)).astype(np.uint32)
    Q = (rng.integers(0, 2**32, (40, 6), dtype=np.uint64) | rng.integers(0, 2**32, (40, 6), dtype=np.uint64)).astype(np.uint32)
    Q[:10] = F[:10] | Q[:10]
    cnt = rng.integers(1, 5, 300)
    ko = np.zeros(301, np.uint64)
    ko[1:] = np.cumsum(cnt)
    keys = rng.integers(0, 50, int(ko[-1])).astype(np.uint32)
    off, vals = orc.match(F, ko, keys, Q, mode=0)
    hoff, hvals = orc.match(F, ko, keys, Q, mode=2)
    for i in range(40):
        m = np.all((F & ~Q[i]) == 0, axis=1)
        exp = np.unique(np.concatenate([keys[int(ko[j]):int(ko[j + 1])] for j in np.nonzero(m)[0]] + [np.zeros(0, np.uint32)]))
        assert np.array_equal(vals[int(off[i]):int(off[i + 1])], exp)
        assert np.array_equal(hvals[int(hoff[i]):int(hoff[i + 1])], np.nonzero(m)[0])


# ------------------------------------------------------------- partitioner

def _ranges(off):
    return [(int(off[i]), int(off[i + 1])) for i in range(len(off) - 1)]


def test_partition_vs_reference_program(golden):
    g = golden.partition
    words = g["words"]
    for maxp in (64, 500, 2000, 10000):
        rc, perm, off, masks = orc.partition(words, maxp)
        assert rc == 0
        # same final positions as the reference's fib vector
        assert np.array_equal(perm, g[f"order_{maxp}"])
        # same partitions: reference partition ids grouped by position
        pid = g[f"pid_{maxp}"]
        ref = {}
        for pos, p in enumerate(pid):
            ref.setdefault(int(p), []).append(pos)
        ref_ranges = sorted((min(v), max(v) + 1) for v in ref.values())
        ours = [r for r in _ranges(off) if r[1] > r[0]]
        assert ours == ref_ranges
        n_empty_ref = int((g[f"sizes_{maxp}"] == 0).sum())
        assert len(_ranges(off)) == len(ref_ranges) + n_empty_ref
        # masks equal to the reference's for every non-empty partition
        ref_masks = {(min(v), max(v) + 1): g[f"masks_{maxp}"][p] for p, v in ref.items()}
        for i, r in enumerate(_ranges(off)):
            if r[1] > r[0]:
                assert np.array_equal(masks[i], ref_masks[r])


def test_partition_reference_test_cases(golden):
    # reference test/test_partitioning.cc:107-246
    v = golden.vectors
    Fp = v["F"][:10] + [v["F10_partitioning"]]
    tables = {"small_fib_no_part": v["Q"][:3], "small_fib_part": Fp[:3], "large_fib_part": Fp,
              "large_fib_permutation": Fp + v["Q"]}
    for name, strings in tables.items():
        case = v["partitioning"][name]
        words = np.stack([w(s) for s in strings])
        rc, perm, off, masks = orc.partition(words, case["maxp"])
        assert rc == 0
        assert len(off) - 1 == case["n_partitions"], name
        groups = sorted(sorted(int(x) for x in perm[a:b]) for a, b in _ranges(off))
        assert groups == case["groups"], name
        assert list(map(int, perm)) == case["order"], name
        for i, (a, b) in enumerate(_ranges(off)):
            for x in perm[a:b]:
                assert orc.subset(masks[i], words[x])
    # test_partitioning.cc:126-150: masks {F0, F1 & F2}
    words = np.stack([w(s) for s in Fp[:3]])
    rc, perm, off, masks = orc.partition(words, 5)
    got = {tuple(m) for m in masks}
    assert got == {tuple(words[0]), tuple(words[1] & words[2])}


def test_partition_invariants():
    import tagmatch_amd as tm
    cw, cko, ck, q, s = tm.workload(20000, 10, seed=77)
    for maxp in (100, 1000, 7000):
        rc, perm, off, masks = orc.partition(cw, maxp)
        assert rc == 0
        assert np.array_equal(np.sort(perm), np.arange(len(cw)))   # each filter exactly once
        for i, (a, b) in enumerate(_ranges(off)):
            assert b - a <= maxp
            if b > a:
                m = np.bitwise_and.reduce(cw[perm[a:b]], axis=0)
                assert np.array_equal(m, masks[i])
                assert m.any()                                  # non-empty mask
                assert np.all(np.diff(perm[a:b].astype(np.int64)) > 0)   # input order kept


def test_partition_all_zero_filter_reported():
    words = np.zeros((3, 6), np.uint32)
    words[1, 0] = 1
    rc, perm, off, masks = orc.partition(words, 10)
    assert rc == -1   # the reference would not terminate


# ------------------------------------------------------------- lookup

def test_lookup_is_exactly_mask_subset():
    rng = np.random.default_rng(3)
    masks = np.zeros((300, 6), np.uint32)
    for i in range(300):
        for b in rng.choice(192, size=rng.integers(1, 6), replace=False):
            masks[i, b // 32] |= np.uint32(1 << (b % 32))
    lk = orc.Lookup(masks)
    for _ in range(200):
        q = (rng.integers(0, 2**32, 6, dtype=np.uint64) | rng.integers(0, 2**32, 6, dtype=np.uint64)).astype(np.uint32)
        exp = np.nonzero(np.all((masks & ~q) == 0, axis=1))[0]
        assert np.array_equal(lk.query(q), exp)


def test_schedule_count_equals_bruteforce():
    import tagmatch_amd as tm
    cw, cko, ck, q, s = tm.workload(8000, 150, seed=5)
    rc, perm, off, masks = orc.partition(cw, 500)
    pw = cw[perm]
    assert orc.schedule_count(pw, off, masks, q) == orc.bruteforce_count(cw, q)


def test_schedule_match_equals_brute_force_keys():
    """The CPU schedule with key output (lookup + per-partition brute force +
    sort/dedup) gives the brute force's per-query keys, in all three modes
    (HITS ids mapped back through the permutation); a planted error in the
    partition table (one partition's mask widened by a bit its filters lack)
    loses matches, so the comparison can fail."""
    import tagmatch_amd as tm
    cw, cko, ck, q, s = tm.workload(8000, 200, seed=6)
    rc, perm, off, masks = orc.partition(cw, 500)
    cnt = (cko[1:] - cko[:-1]).astype(np.int64)[perm]
    ko = np.zeros(len(perm) + 1, np.uint64)
    np.cumsum(cnt, out=ko[1:])
    idx = np.repeat(cko[:-1][perm].astype(np.int64) - ko[:-1].astype(np.int64), cnt) + np.arange(int(ko[-1]))
    pk = ck[idx]
    for mode in (0, 1, 2):
        eo, ev = orc.match(cw, cko, ck, q, mode=mode)
        go, gv = orc.schedule_match(cw[perm], ko, pk, off, masks, q, mode=mode, threads=3)
        assert np.array_equal(go, eo), mode
        for i in range(len(q)):
            a = gv[int(go[i]):int(go[i + 1])]
            b = ev[int(eo[i]):int(eo[i + 1])]
            if mode == 2:
                a = np.sort(perm[a])
            elif mode == 1:
                a = np.sort(a)
            assert np.array_equal(a, b), (mode, i)
    bad = masks.copy()
    big = int(np.argmax(np.diff(off)))
    free_bit = next(b for b in range(192) if not (bad[big, b // 32] >> (b % 32)) & 1)
    bad[big, free_bit // 32] |= np.uint32(1 << (free_bit % 32))
    go, gv = orc.schedule_match(cw[perm], ko, pk, off, bad, q, mode=2)
    eo, ev = orc.match(cw, cko, ck, q, mode=2)
    assert int(go[-1]) < int(eo[-1])


# ------------------------------------------------------------- consolidate

def _ops(entries):
    rows = []
    for s, k, op in entries:
        r = np.zeros(8, np.uint32)
        r[:6] = w(s)
        r[6] = k
        r[7] = op
        rows.append(r)
    return np.stack(rows)


def test_consolidate_merges_identical_filters(golden):
    F = golden.vectors["F"]
    # reference test_tagmatch.cc:339-364 add_same_filter -> {2,3,4}
    words, off, keys = orc.consolidate(_ops([(F[1], 2, 0), (F[1], 3, 0), (F[1], 4, 0), (F[1], 3, 0)]))
    assert words.shape[0] == 1 and list(keys) == [2, 3, 4]
    # lexicographic order of the table
    words, off, keys = orc.consolidate(_ops([(F[i], i, 0) for i in range(11)]))
    for i in range(len(words) - 1):
        assert orc.compare(words[i], words[i + 1]) < 0


def test_consolidate_matches_engine_host_code():
    import tagmatch_amd as tm
    g = tm.Generator(seed=99)
    rw, rko, rk = g.filters(5000)
    cw, cko, ck = tm.canonicalize(rw, rko, rk)
    idx = np.repeat(np.arange(5000), np.diff(rko).astype(np.int64))
    ops = np.zeros((len(rk), 8), np.uint32)
    ops[:, :6] = rw[idx]
    ops[:, 6] = rk
    ow, oko, ok = orc.consolidate(ops)
    assert np.array_equal(ow, cw) and np.array_equal(oko, cko) and np.array_equal(ok, ck)


# ------------------------------------------------------------- live reference

@pytest.mark.skipif(not orc.ref_available("test_matcher"), reason="oracle/_ref not built")
def test_oracle_vs_live_reference():
    import tagmatch_amd as tm
    d = tempfile.mkdtemp()
    cw, cko, ck, q, s = tm.workload(4000, 60, seed=2024)
    ref = orc.ref_test_matcher(cw, cko, ck, q, d)
    off, vals = orc.match(cw, cko, ck, q, mode=0)
    for i, r in enumerate(ref):
        assert np.array_equal(np.unique(r), vals[int(off[i]):int(off[i + 1])])
    order, pid, masks, sizes = orc.ref_partition(cw, 300, d, threads=3)
    rc, perm, poff, pmasks = orc.partition(cw, 300)
    assert np.array_equal(order, perm)


@pytest.mark.skipif(not orc.ref_available("libref_patricia.so"), reason="oracle/_ref not built")
def test_oracle_vs_reference_patricia():
    """The reference compact Patricia-trie matcher (the CPU baseline) and the
    oracle give the same match_unique results; a table with repeated filters
    merges their keys, and results do not depend on the thread count."""
    import tagmatch_amd as tm
    cw, cko, ck, q, s = tm.workload(20000, 300, seed=2025)
    t = orc.RefPatricia(cw, cko, ck)
    assert t.n_nodes == len(cw)
    off, vals = orc.match(cw, cko, ck, q, mode=0)
    po, pv = t.match_unique(q, threads=1)
    assert np.array_equal(po, off.astype(np.uint64)) and np.array_equal(pv, vals)
    po3, pv3 = t.match_unique(q, threads=3)
    assert np.array_equal(po3, po) and np.array_equal(pv3, pv)
    # raw (not canonical) table: every filter twice, keys split between copies
    g = tm.Generator(seed=7)
    rw, rko, rk = g.filters(3000)
    n = len(rw)
    half = rko[:-1] + (rko[1:] - rko[:-1]) // 2
    w2 = np.concatenate([rw, rw])
    k1 = [rk[int(rko[i]):int(half[i])] for i in range(n)]
    k2 = [rk[int(half[i]):int(rko[i + 1])] for i in range(n)]
    lists = k1 + k2
    ko2 = np.zeros(2 * n + 1, np.uint64)
    np.cumsum([len(x) for x in lists], out=ko2[1:])
    kv2 = np.concatenate(lists).astype(np.uint32)
    t2 = orc.RefPatricia(w2, ko2, kv2)
    cw2, cko2, ck2 = tm.canonicalize(rw, rko, rk)
    assert t2.n_nodes == len(cw2)
    q2, _ = g.queries(200, 3000)
    o2, v2 = orc.match(cw2, cko2, ck2, q2, mode=0)
    po, pv = t2.match_unique(q2, threads=2)
    assert np.array_equal(po, o2.astype(np.uint64)) and np.array_equal(pv, v2)
