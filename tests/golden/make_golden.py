"""Write the golden fixtures of tests/golden from the reference's own programs.

Run in the build container (it needs oracle/_ref, compiled from the
reference sources by `make -C oracle ref`):

    python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs, nothing else):
  ref_vectors.json      filter strings of test_tagmatch.cc / test_partitioning.cc
                        and the expected answers of their cases; the
                        bitvector cases of test_bitvector.cc (strings and
                        expected results)
  bitvector_pairs.npz   random filter pairs and the reference bitvector<192>
                        results (subset, order, leftmost_diff, popcount, words)
  match_small.npz       a generated table + queries; the reference
                        test_matcher's output (keys and, with key = filter
                        index, the matched filter ids)
  partition_small.npz   a generated table; the reference partitioner's
                        final order and partition ranges at several maxp
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as orc  # noqa: E402
import tagmatch_amd as tm  # noqa: E402

# Filter and query strings of the reference tests (test_tagmatch.cc:100-128;
# test_partitioning.cc:40-68 is the same table except F[10], which differs).
F = [
    "000000000000100000001000000000000001000000010001000000000000000000000000000000000000000000000000001000000000000000000000001000000000000000000010000000000001000000000000010000000000001010010001",
    "000000000000010000000100000000000000100000001000100000000000000000000000000000000000000000000000000100000000000000000000000100000000000000000001000000000000100000000000001000000000000101001000",
    "000000000000010000000000000000000000100000000000100000000010000000000010010000000000000100010000000000000000100000000000000000000000000000100000000100000010000101000000000110000010000100000010",
    "000000000001000000000001000100000000001000000010000000000000000001000000000000000100000000000000000000000000000010100000010000000011000000010000000010000000000000000010000000001000000000101000",
    "000010000100100000000000001000000100000000000000000000000000000000000000000101000000000000010000000000000000000000000000001000000100000000101000000000000010000000000000010000000000000000000000",
    "000110000010000100100000000000000000000000010000000000001001000000100000000000000000100010000000000000000000000000000000000000101000000000000011000000001000000000000100000000000001001000000010",
    "011000000010000000000000000000000001000000000000000000000000000000000001000010000000000000001000000000000010000000000010000000000000010000000100000000100000000000000000000000000000101000000000",
    "000010000000010000000000000010010000000000000001000100000000000000000000000000010000000000000000100000000000000000000000000000000000000000001000000000000000000000000000000000000000111000000000",
    "000000000000000000000000000000001100000000000000000100000000000000010000000010001000000000000000000000001000000000000000010000000001100000000000000000000000000000000000000000000100000100100001",
    "100000100000000000000010000000000000000000000000000000000010000010000000011010010000000101000000000000000000000000000000001001000000010100000000000000000000000000000000001010001100000000000000",
    "000000000000000100000000100000010000000010000000000010100000000000000000100000000000000000010011000101000000000100000000000001000000000001000000010100100000000100000000000000011000000000000000",
]
F10_PARTITIONING = "000000000000000100000000100000010000000010000000000010100000000000000000100000000000000000010011000101000000000100000000000001000000000001000000010100100000000100000000000000011000001000000000"
Q = [
    "000000000000010000000100000000000000100000001000100000000000000000000000000000000000000000000000000100000000000000000000000100000000000000000001000000000000100000000000000000000000000101001000",
    "100000100000010001000010000000000000100000000000100011000010000010000010011010010000000101010000000000000001100000000000001001000000010100100000000100000010000111000000000111000110000100000010",
    "000000000000010000000000000000000000100000000000100000000010000000000010010000000000000100010000000000000000100000000000000000000000000000100000000100000010000101000000000110000010000100000010",
    "000000000001000000000001000100000000001000000010000000000000000001000000000000000100000000000000000000000000000010100000010000000011000000010000000010000000000000000010000000001000000000101000",
    "100010100111110101100010000000000000100000011111010001111101000001100001001101001000100010101000000000000000110000000000001100100100001010111000000001000011100001110000111000001110000100000010",
    "000010000100100000000000001000000100000000000000000000000000000000000000000101000000000000010000000000000000000000000000001000000100000000101000000000000010000000000000010000000000000000000000",
    "000110000010000100100000000000000000000000010000000000001001000000100000000000000000100010000000000000000000000000000000000000101000000000000011000000001000000000000100000000000001001000000010",
    "011000000010000000000000000000000001000000000000000000000000000000000001000010000000000000001000000000000010000000000010000000000000010000000100000000100000000000000000000000000000101000000000",
    "000010000000010000000000000010010000000000000001000100000000000000000000000000010000000000000000100000000000000000000000000000000000000000001000000000000000000000000000000000000000111000000000",
    "000000000000000000000000000000001100000000000000000100000000000000010000000010001000000000000000000000001000000000000000010000000001100000000000000000000000000000000000000000000100000100100001",
    "000000100000000000000010001010011100000000000001000100000010000010010011001110011000000101000011100000001111000000000000011001000001110100001100000000000000000000000000001010001100111100110001",
    "000000000000000100000000100000010000000010000000000010100000000000000000100000000000000000010011000101000000000100000000000001000000000001000000010100100000000100000000000000011000000000000000",
    "111111111111111111111111111111111111111011111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111111",
]


def words_of(strings):
    return np.stack([orc.from_string(s) for s in strings])


def table_of(strings, keys=None):
    w = words_of(strings)
    n = w.shape[0]
    ko = np.arange(n + 1, dtype=np.uint64)
    k = np.arange(n, dtype=np.uint32) if keys is None else np.asarray(keys, np.uint32)
    return w, ko, k


def bitvector_cases(path):
    """Cases of the reference's bitvector unit test, read as data: calls of
    subset_of / prefix_less_than / leftmost_diff / equals with their two
    filter strings and expected result (test_bitvector.cc:426-590)."""
    import re
    text = open(path).read()
    cases = []
    pat = re.compile(r'BOOST_CHECK\(\s*(subset_of|prefix_less_than|leftmost_diff|equals)\(\s*"([01]+)"\s*,\s*"([01]+)"\s*\)'
                     r'\s*==\s*(true|false|filter_t::WIDTH|\d+)\s*\)', re.S)
    for m in pat.finditer(text):
        fn, a, b, exp = m.groups()
        if exp in ("true", "false"):
            val = exp == "true"
        elif exp == "filter_t::WIDTH":
            val = 192
        else:
            val = int(exp)
        cases.append({"fn": fn, "a": a, "b": b, "expected": val})
    return cases


def main():
    work = tempfile.mkdtemp()
    assert orc.ref_available("test_matcher"), "build the reference programs: make -C oracle ref"

    # --- test_tagmatch vectors: reference test_matcher with key = index
    w, ko, k = table_of(F)
    qw = words_of(Q)
    ref = orc.ref_test_matcher(w, ko, k, qw, work)
    misc = [sorted(set(int(x) for x in r)) for r in ref]

    # --- test_partitioning vectors through the reference partitioner
    Fp = F[:10] + [F10_PARTITIONING]
    parts = {}
    for name, strings, maxp in (("small_fib_no_part", Q[:3], 5), ("small_fib_part", Fp[:3], 5),
                                ("large_fib_part", Fp, 3), ("large_fib_permutation", Fp + Q, 4)):
        order, pid, masks, sizes = orc.ref_partition(words_of(strings), maxp, work)
        groups = {}
        for pos, p in enumerate(pid):
            groups.setdefault(int(p), []).append(int(order[pos]))
        parts[name] = {
            "maxp": maxp,
            "n_partitions": int(len(masks)),
            "groups": sorted(sorted(v) for v in groups.values()),
            "order": [int(x) for x in order],
        }
    ref_root = os.environ.get("REF", "/root/reference")
    vec = {
        "bitvector_cases": bitvector_cases(os.path.join(ref_root, "test", "test_bitvector.cc")),
        "F": F, "F10_partitioning": F10_PARTITIONING, "Q": Q,
        "misc_subsets_keys": misc,
        "partitioning": parts,
        "source": "reference test/test_tagmatch.cc:100-128, test/test_partitioning.cc:40-68; "
                  "answers from the reference test_matcher and partitioner",
    }
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(vec, f, indent=1)

    # --- bitvector pairs
    rng = np.random.default_rng(5)
    n = 4000
    a = rng.integers(0, 2**32, size=(n, 6), dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, size=(n, 6), dtype=np.uint64).astype(np.uint32)
    # sparse filters, supersets, equal pairs and single-word differences
    a[: n // 4] &= rng.integers(0, 2**32, size=(n // 4, 6), dtype=np.uint64).astype(np.uint32)
    a[: n // 4] &= rng.integers(0, 2**32, size=(n // 4, 6), dtype=np.uint64).astype(np.uint32)
    b[: n // 4] = a[: n // 4] | rng.integers(0, 2**32, size=(n // 4, 6), dtype=np.uint64).astype(np.uint32)
    b[n // 4: n // 4 + 200] = a[n // 4: n // 4 + 200]
    for i in range(n // 4 + 200, n // 4 + 800):
        b[i] = a[i]
        j = rng.integers(0, 192)
        b[i, j // 32] ^= np.uint32(1 << (j % 32))
    a[n - 10:] = 0
    pairs = np.concatenate([a, b], axis=1)
    r = orc.ref_bitvector(pairs, work)
    np.savez_compressed(os.path.join(HERE, "bitvector_pairs.npz"), a=a, b=b, **r)

    # --- generated table + queries through the reference matcher
    cw, cko, ck, q, s = tm.workload(3000, 120, seed=0x7461676D61746301)
    ref_keys = orc.ref_test_matcher(cw, cko, ck, q, work)
    idx_keys = np.arange(cw.shape[0], dtype=np.uint32)
    ref_ids = orc.ref_test_matcher(cw, np.arange(cw.shape[0] + 1, dtype=np.uint64), idx_keys, q, work)

    def pack(lists):
        off = np.zeros(len(lists) + 1, np.uint64)
        off[1:] = np.cumsum([len(x) for x in lists])
        vals = np.concatenate([np.asarray(x, np.uint32) for x in lists]) if off[-1] else np.zeros(0, np.uint32)
        return off, vals

    ko_u, v_u = pack([np.unique(x) for x in ref_keys])
    ko_a, v_a = pack([np.sort(x) for x in ref_keys])
    ko_h, v_h = pack([np.sort(x) for x in ref_ids])
    np.savez_compressed(os.path.join(HERE, "match_small.npz"), words=cw, key_off=cko, keys=ck, queries=q,
                        unique_off=ko_u, unique_vals=v_u, all_off=ko_a, all_vals=v_a, hits_off=ko_h, hits_vals=v_h)

    # --- generated table through the reference partitioner
    pw, _, _, _, _ = tm.workload(6000, 1, seed=0x7461676D61746302)
    out = {"words": pw}
    for maxp in (64, 500, 2000, 10000):
        order, pid, masks, sizes = orc.ref_partition(pw, maxp, work, threads=4)
        out[f"order_{maxp}"] = order
        out[f"pid_{maxp}"] = pid
        out[f"masks_{maxp}"] = masks
        out[f"sizes_{maxp}"] = sizes
    np.savez_compressed(os.path.join(HERE, "partition_small.npz"), **out)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
