"""Parity at the full sizes of the benchmark workloads, on the GPU.

  C1: 1M generated filters, 100k queries, partitions of at most 64k filters:
      every query against the oracle.
  C2: the benchmark's own workload (seed 0x7461676D61746302, 10M filters,
      1M queries, partitions of at most 256k filters): the default step over
      all 1M queries (rich-filter pass, compaction, full-key-set exit) against
      the oracle on a 10k-query sample, a 500-query sample in HITS mode, and
      the whole run cross-checked between the default step, the step without
      the exit (every match enumerated), other tile shapes and the plain
      per-(partition, 256-query batch) form on a prefix.
  C3: the same 10M table cut into Zipf(1.2) partitions of 1k-1M filters,
      batches of 1-4096 queries per partition: the persistent launch against
      the plain form on every unit, and the oracle on a sample of units.

The oracle is the brute force of reference test_matcher.cc:74-81 with the
sort and dedup of tagmatch_query.hh:105-109.  Whole-run comparisons use a
digest of the offsets and values (xxh3-128) so that the 8 GB result arrays
of a C2 step are never held twice.
"""
import time

import numpy as np
import pytest
import xxhash

import oracle as orc
import tagmatch_amd as tm

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

SEED_C1 = 0x7461676D61746301
SEED_C2 = 0x7461676D61746302


def log(msg):
    print(msg, flush=True)


def digest(off, vals):
    h = xxhash.xxh3_128()
    h.update(memoryview(np.ascontiguousarray(off)))
    h.update(memoryview(np.ascontiguousarray(vals)))
    return h.hexdigest()


def compare_sample(off, vals, idx, eo, ev, what):
    """Results of queries idx (from a whole-run off/vals) against the
    oracle's results for those queries, in order."""
    cnt = (off[idx + 1] - off[idx]).astype(np.int64)
    ecnt = np.diff(eo).astype(np.int64)
    bad = np.nonzero(cnt != ecnt)[0]
    assert len(bad) == 0, f"{what}: query {int(idx[bad[0]])}: {cnt[bad[0]]} keys vs {ecnt[bad[0]]} expected"
    starts = off[idx].astype(np.int64)
    pos = np.repeat(starts - eo[:-1].astype(np.int64), ecnt) + np.arange(int(eo[-1]), dtype=np.int64)
    got = vals[pos]
    if not np.array_equal(got, ev):
        first = int(np.nonzero(got != ev)[0][0])
        q = int(np.searchsorted(eo, first, side="right")) - 1
        raise AssertionError(f"{what}: query {int(idx[q])} differs from the oracle")


def generate(seed, n_filters, n_queries):
    t = time.time()
    g = tm.Generator(seed=seed)
    w, ko, k = g.filters(n_filters)
    q, _ = g.queries(n_queries, n_filters)
    cw, cko, ck = tm.canonicalize(w, ko, k)
    log(f"generated {n_filters} filters ({len(cw)} distinct), {n_queries} queries in {time.time() - t:.1f}s")
    return cw, cko, ck, q


def step(ring, nq, **cfg):
    ring.configure(tm.config(mode=tm.MODE_KEYS_UNIQUE, **cfg))
    ring.submit_resident(nq)
    st = ring.wait()
    assert st.error == 0          # an overflow retry (st.overflow) is allowed: results are complete
    off, vals = ring.results()
    return off, vals, st


# ------------------------------------------------------------------ C1

def test_c1_every_query_vs_oracle(gpu):
    cw, cko, ck, q = generate(SEED_C1, 1_000_000, 100_000)
    fib, perm, poff, masks = tm.build_fib(cw, cko, ck, 64 * 1024)
    log(f"C1: {len(poff) - 1} partitions")
    ring = tm.Ring(fib, len(q))
    ring.submit(q)
    st = ring.wait()
    off, vals = ring.results()
    t = time.time()
    eo, ev = orc.match(cw, cko, ck, q, mode=0)
    log(f"C1 oracle: {len(q)} queries in {time.time() - t:.1f}s, {int(eo[-1])} keys")
    assert np.array_equal(off, eo), "C1: key counts differ from the oracle"
    assert np.array_equal(vals, ev), "C1: keys differ from the oracle"
    # the plain form, the step without the exit and 128-query tiles agree
    ref = digest(off, vals)
    for cfg in (dict(schedule=tm.SCHED_PLAIN, batch_size=256), dict(full_exit=2),
                dict(batch_size=128, waves_per_group=4)):
        o, v, _ = step(ring, len(q), **cfg)
        log(f"C1 {cfg}: {int(o[-1])} keys")
        assert digest(o, v) == ref, f"C1 {cfg} differs from the default step"
    ring.close()
    fib.close()


# ------------------------------------------------------------------ C2

@pytest.fixture(scope="module")
def c2(gpu):
    cw, cko, ck, q = generate(SEED_C2, 10_000_000, 1_000_000)
    t = time.time()
    fib, perm, poff, masks = tm.build_fib(cw, cko, ck, 256 * 1024)
    log(f"C2: {len(poff) - 1} partitions, partition + upload {time.time() - t:.1f}s")
    ring = tm.Ring(fib, len(q))
    ring.submit(q)
    ring.wait()
    yield dict(cw=cw, cko=cko, ck=ck, q=q, fib=fib, ring=ring)
    ring.close()
    fib.close()


def test_c2_default_step_vs_oracle_sample(c2):
    q, ring = c2["q"], c2["ring"]
    off, vals, st = step(ring, len(q))
    log(f"C2 default step: {st.n_items} items, {st.n_hits} hits enumerated, {int(off[-1])} keys, "
        f"{st.device_ms:.2f} ms on the device")
    idx = np.arange(0, len(q), 100)                      # 10k queries spread over the run
    t = time.time()
    eo, ev = orc.match(c2["cw"], c2["cko"], c2["ck"], q[idx], mode=0)
    log(f"C2 oracle: {len(idx)} queries in {time.time() - t:.1f}s")
    compare_sample(off, vals, idx, eo, ev, "C2 default step")
    c2["default_digest"] = digest(off, vals)
    c2["default_off"] = off
    c2["default_prefix"] = vals[:int(off[32768])].copy()


def test_c2_full_run_cross_check(c2):
    q, ring = c2["q"], c2["ring"]
    if "default_digest" not in c2:
        off, vals, _ = step(ring, len(q))
        c2["default_digest"] = digest(off, vals)
        c2["default_off"] = off
        c2["default_prefix"] = vals[:int(off[32768])].copy()
        del vals
    ref = c2["default_digest"]
    # the autotuner's other shapes, and the step that enumerates every match
    for cfg in (dict(batch_size=64, waves_per_group=8, chunk_blocks=4096),
                dict(batch_size=128, waves_per_group=4, chunk_blocks=0),
                dict(full_exit=2)):
        o, v, st = step(ring, len(q), **cfg)
        log(f"C2 {cfg}: {st.n_hits} hits enumerated, {int(o[-1])} keys")
        assert digest(o, v) == ref, f"C2 {cfg} differs from the default step"
        del o, v
    # the plain per-(partition, 256-query batch) form on a prefix
    n = 32768
    o, v, st = step(ring, n, schedule=tm.SCHED_PLAIN, batch_size=256)
    log(f"C2 plain form: {n} queries, {st.n_items} launches")
    assert np.array_equal(o, c2["default_off"][:n + 1])
    assert np.array_equal(v, c2["default_prefix"])


def test_c2_packed_delivery(c2):
    """The headline's result form at full size: packed key bitmap rows in
    pinned host memory (graph replay and not), expanded on the host, equal
    the default KEYS_UNIQUE step on every query."""
    q, fib = c2["q"], c2["fib"]
    if "default_digest" not in c2:
        off, vals, _ = step(c2["ring"], len(q))
        c2["default_digest"] = digest(off, vals)
        del off, vals
    d = fib.dictionary()
    kw = (len(d) + 31) // 32
    for graph in (0, 1):
        ring = tm.Ring(fib, len(q), tm.config(mode=tm.MODE_KEY_BITMAP, use_graph=graph))
        buf = tm.PackedRows(len(q), kw)
        ring.set_host_rows_packed(buf)
        for _ in range(2):                    # the second step copies its rows with itself
            ring.submit(q)
            st = ring.wait()
        n = int(st.n_results)
        row_of = buf.row_of.array[:len(q)]
        assert int((row_of != tm.ROW_FULL).sum()) == n
        off, vals = tm.expand_key_rows_packed(row_of, buf.rows.array[:n * kw].reshape(n, kw), d)
        log(f"C2 packed (graph {graph}): {n} rows not the whole dictionary, {int(off[-1])} keys")
        assert digest(off, vals) == c2["default_digest"], f"C2 packed delivery (graph {graph})"
        del off, vals
        ring.close()
        buf.close()


def test_c2_hits_sample_vs_oracle(c2):
    q = c2["q"]
    idx = np.arange(7, len(q), 2000)                     # 500 queries
    o, v, st = tm.match(c2["fib"], q[idx], mode=tm.MODE_HITS)
    eo, ev = orc.match(c2["cw"], c2["cko"], c2["ck"], q[idx], mode=2)
    assert st["n_hits"] == int(eo[-1])
    for i in range(len(idx)):
        got = np.sort(v[int(o[i]):int(o[i + 1])])
        assert np.array_equal(got, ev[int(eo[i]):int(eo[i + 1])]), f"C2 HITS: query {int(idx[i])}"
    log(f"C2 HITS: {len(idx)} queries, {int(eo[-1])} matched filters")


# ------------------------------------------------------------------ C3

def test_c3_uneven_units(c2):
    cw, cko, ck = c2["cw"], c2["cko"], c2["ck"]
    poff = tm.zipf_runs(len(cw), 1000, 1_000_000, s=1.2, seed=3)
    sizes = np.diff(poff)
    assert sizes.min() >= 1000 and sizes.max() <= 1_000_000
    fib, masks = tm.fib_from_runs(cw, cko, ck, poff)
    parts, offs = [], [0]
    for r in range(2):
        up, uo = tm.uneven_units(len(poff) - 1, 1, 4096, seed=100 + r)
        parts.append(up)
        offs.extend((offs[-1] + uo[1:].astype(np.int64)).tolist())
    up = np.concatenate(parts).astype(np.uint32)
    uo = np.array(offs, np.int64)
    nq = int(uo[-1])
    g = tm.Generator(seed=SEED_C2)
    q, _ = g.queries(nq, len(cw), first=5_000_000)
    log(f"C3: {len(poff) - 1} partitions ({sizes.min()}..{sizes.max()} filters), {len(up)} units, {nq} queries")
    ring = tm.Ring(fib, nq)
    res = {}
    for name, sched in (("persistent", tm.SCHED_PERSISTENT), ("plain", tm.SCHED_PLAIN)):
        ring.configure(tm.config(schedule=sched))
        ring.submit_units(q, up, uo.astype(np.uint32))
        st = ring.wait()
        assert st.error == 0
        res[name] = ring.results()
        log(f"C3 {name}: {st.n_items} items, {int(res[name][0][-1])} keys")
    off, vals = res["persistent"]
    assert np.array_equal(off, res["plain"][0]) and np.array_equal(vals, res["plain"][1])
    # the oracle per unit: the unit's queries against its partition's run
    rng = np.random.default_rng(33)
    work, checked = 0, 0
    for u in rng.permutation(len(up)):
        p = int(up[u])
        a, b = int(poff[p]), int(poff[p + 1])
        qa, qb = int(uo[u]), int(uo[u + 1])
        if (b - a) * (qb - qa) > 2e9 or work + (b - a) * (qb - qa) > 1.5e10:
            continue
        work += (b - a) * (qb - qa)
        ko = (cko[a:b + 1] - cko[a]).astype(np.uint64)
        eo, ev = orc.match(cw[a:b], ko, ck[int(cko[a]):int(cko[b])], q[qa:qb], mode=0)
        compare_sample(off, vals, np.arange(qa, qb), eo, ev, f"C3 unit {u} (partition {p}, {b - a} filters)")
        checked += 1
    log(f"C3 oracle: {checked} units, {work:.2e} pair tests")
    assert checked >= 40
    ring.close()
    fib.close()
