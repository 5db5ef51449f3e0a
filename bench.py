#!/usr/bin/env python3
"""Throughput benchmark of the MI355X TagMatch engine on one GPU.

Workload (the metric's configuration): 10M generated 192-bit filters (1-10
Zipf tags x 7 bits, 1-4 keys each, identical filters merged), partitioned on
the GPU into partitions of at most 256k filters; 1M generated queries (10-30
Zipf tags, 30% supersets of a filter), matched in batches of 256 queries per
partition.  A step is one pass of the whole engine over the 1M queries,
which are resident in HBM when timing starts: on-device partition lookup,
batched subset match, key expansion and per-query dedup, and the delivery of
every query's sorted unique key set to host memory (as a key bitmap row over
the FIB's sorted key dictionary, 256 bytes per query at 2048 keys).

Beside the headline step the run measures, on the same resident table:
  * the device-resident step with sorted key lists left in HBM, the same
    lists copied to host memory, and the rows expanded into lists on the host;
  * parity outside the timed region: the delivered rows against the key-list
    step's rows and the rows of a step that enumerates every match, and a
    sample of queries against the oracle;
  * the plain single-launch form (one launch per (partition, 256-query
    batch), the reference's dispatch shape) on a prefix of the queries;
  * the learned cost model: trained on the plain form's own unit logs,
    held-out MAPE, and the persistent step priced by it;
  * uneven batching: Zipf(1.2) partitions of 1k-1M filters, batches of
    1-4096 queries, persistent launch vs the plain form;
  * the serving loop: Poisson arrivals of 64-query requests at 30/70/95% of
    its sustained peak, with 10k-filter inserts every 100 ms; p50/p99/p999;
  * the CPU baselines: the oracle's restatement of the reference's CPU
    schedule and brute force, and the reference's Patricia-trie matcher, on
    bounded samples.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this engine runs
on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import tagmatch_amd as tm  # noqa: E402

METRIC = "queries/s and matches/s at 10M 192-bit filters; p99 latency; % HBM roofline"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12   # int32 lane-ops/s, non-packed
OPS_PER_PAIR = 7                            # 6 and-not/or + 1 compare per pair-test
PCIE_PEAK_GBS = 63.0                        # PCIe 5.0 x16, one direction
# (queries per work item, waves per workgroup, blocks per segment (0 = automatic), replay from a
# captured graph)
AUTOTUNE_CANDIDATES = [(64, 4, 0, 0), (64, 8, 0, 0), (64, 4, 2048, 0), (64, 4, 512, 0), (128, 4, 0, 0),
                       (128, 8, 0, 0), (64, 4, 0, 1), (64, 8, 0, 1)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--filters", type=int, default=10_000_000)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--maxp", type=int, default=256 * 1024)
    ap.add_argument("--batch", type=int, default=256, help="queries per (partition, batch) of the plain form")
    ap.add_argument("--tile", type=int, default=0, help="queries per persistent work item (0: autotuned)")
    ap.add_argument("--autotune", type=int, default=1, help="autotune tile size and waves before timing (0: off)")
    ap.add_argument("--chunk", type=int, default=0, help="blocks per work item (0: default)")
    ap.add_argument("--groups-per-cu", type=int, default=0)
    ap.add_argument("--schedule", choices=["persistent", "plain"], default="persistent")
    ap.add_argument("--graph", action="store_true", help="replay steps from a hipGraph")
    ap.add_argument("--plain-steps", type=int, default=1, help="steps of the plain single-launch form")
    ap.add_argument("--plain-queries", type=int, default=16384,
                    help="queries of the plain form's step (a prefix of the resident queries)")
    ap.add_argument("--model", type=int, default=1, help="cost model leg (0: skip)")
    ap.add_argument("--uneven-rounds", type=int, default=8,
                    help="batches per partition in the uneven-batching leg (0: skip the leg)")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="seconds per serving load level (0: skip)")
    ap.add_argument("--serve-request", type=int, default=64, help="queries per serving request")
    ap.add_argument("--serve-slots", type=int, default=3, help="serving slots (device steps in flight)")
    ap.add_argument("--serve-slot-queries", type=int, default=1024,
                    help="queries per serving step at most (small steps keep the latency low)")
    ap.add_argument("--serve-graphs", type=int, default=0,
                    help="serving steps from captured graphs: 0 default (on), 1 on, 2 off")
    ap.add_argument("--serve-insert", type=int, default=10000, help="filters per insert (0: none)")
    ap.add_argument("--serve-insert-ms", type=float, default=100.0, help="interval between inserts")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--parity-queries", type=int, default=1000,
                    help="queries checked against the oracle after the timed steps (0: skip)")
    ap.add_argument("--host-key-steps", type=int, default=2,
                    help="steps timed with full sorted key lists copied to host memory (0: skip)")
    ap.add_argument("--seed", type=lambda x: int(x, 0), default=0x7461676D61746302)
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    ap.add_argument("--leg", choices=["headline", "c4", "c5"], default="headline",
                    help="c4: the learned cost model at 50M filters (plain unit logs, MAPE, throughput against the "
                         "plain form); c5: the serving loop at 200M filters; both print one JSON line of their own")
    ap.add_argument("--c4-filters", type=int, default=50_000_000)
    ap.add_argument("--c4-rows", type=int, default=2_000_000, help="plain unit log rows to train on")
    ap.add_argument("--c4-log-queries", type=int, default=4_000_000,
                    help="queries the plain form logs units over (drawn apart from the timed ones)")
    ap.add_argument("--c4-log-seconds", type=float, default=660.0, help="time limit of the logging runs")
    ap.add_argument("--c5-filters", type=int, default=200_000_000)
    ap.add_argument("--c5-step-queries", type=int, default=1_000_000,
                    help="c5 leg: queries of the packed-delivery step timed on the big table (0 skips it)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run an elementwise device copy of MIB MiB (read + write) three times, "
                         "to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE, and exit")
    return ap.parse_args()


def setup(args):
    t0 = time.time()
    g = tm.Generator(seed=args.seed)
    w, ko, k = g.filters(args.filters)
    q, sup = g.queries(args.queries, args.filters)
    t1 = time.time()
    cw, cko, ck = tm.canonicalize(w, ko, k)
    del w, ko, k
    t2 = time.time()
    times = {}
    fib, perm, off, masks = tm.build_fib(cw, cko, ck, args.maxp, times=times)
    t3 = time.time()
    log(f"setup: generate {t1 - t0:.1f}s canonicalize {t2 - t1:.1f}s partition (GPU) {times['partition']:.2f}s "
        f"upload {times['upload']:.2f}s; {cw.shape[0]} distinct filters, {len(off) - 1} partitions")
    return cw, cko, ck, q, fib, perm, off, masks, {"total": t3 - t0, "generate": t1 - t0, "canonicalize": t2 - t1,
                                                    **times}


def _sized(fn, q, budget_s, n=16):
    """Runs fn on a prefix of q sized to about budget_s seconds; returns
    (queries run, seconds, fn's result)."""
    t = time.time()
    fn(q[:n])
    dt = max(time.time() - t, 1e-3)
    n2 = int(min(len(q), max(n, n * budget_s / dt)))
    t = time.time()
    r = fn(q[:n2])
    return n2, time.time() - t, r


def cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"hardware_concurrency": os.cpu_count(), "cpu_model": model}


def cpu_baseline(cw, cko, ck, perm, part_off, masks, q, budget_s, maxp):
    """CPU baselines on bounded samples of the same queries, on the host's
    cores, all with the reference's match_unique output (each query's keys
    collected, sorted and deduplicated):
      * primary: the reference's CPU schedule, restated by the oracle --
        partition lookup (front_end.cc:465-477) and per-partition brute force
        (back_end.cc:490-516), all threads;
      * the brute force of test_matcher.cc:74-81 over the whole table, all
        threads and one thread;
      * the reference's compact Patricia-trie matcher built from its sources
        into oracle/_ref, when present (kind "reference")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    cnt = (cko[1:] - cko[:-1]).astype(np.int64)[perm]
    ko = np.zeros(len(perm) + 1, np.uint64)
    np.cumsum(cnt, out=ko[1:])
    idx = np.repeat(cko[:-1][perm].astype(np.int64) - ko[:-1].astype(np.int64), cnt) + np.arange(int(ko[-1]))
    pw, pk = cw[perm], ck[idx]
    del idx
    share = budget_s / 4
    n1, dt1, (o1, _) = _sized(lambda x: orc.schedule_match(pw, ko, pk, part_off, masks, x, mode=0, threads=threads),
                              q, share)
    n2, dt2, (o2, _) = _sized(lambda x: orc.match(cw, cko, ck, x, mode=0, threads=threads), q, share)
    n3, dt3, _ = _sized(lambda x: orc.match(cw, cko, ck, x, mode=0, threads=1), q, share / 2, n=2)
    out = {"value": n1 / dt1, "unit": "queries/s", "cores": threads, "kind": "port",
           "sample": f"first {n1} of the {len(q)} queries against all {cw.shape[0]} filters: oracle restatement of "
                     f"the reference's CPU schedule (partition lookup + per-partition brute force), keys sorted and "
                     f"deduplicated per query, {threads} threads, {dt1:.1f}s",
           "keys_per_query": float(o1[-1]) / n1, "compiler": "gcc " + orc.compiler(), **cpu_info(),
           "bruteforce": {"value": n2 / dt2, "unit": "queries/s", "cores": threads, "queries": n2, "seconds": dt2},
           "bruteforce_single_thread": {"value": n3 / dt3, "unit": "queries/s", "cores": 1, "queries": n3,
                                        "seconds": dt3}}
    # the partitioning of the whole table on the host: the oracle's
    # restatement (one thread) and, when built, the reference's own program
    # (all threads; its time includes writing the table to a file and
    # reading the result back)
    t = time.time()
    _, _, poff, _ = orc.partition(cw, maxp)
    out["partition"] = {"port_seconds": time.time() - t, "port_threads": 1, "partitions": int(len(poff) - 1)}
    if orc.ref_available("ref_partition"):
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            t = time.time()
            orc.ref_partition(cw, maxp, d, threads=threads)
            out["partition"].update({"reference_seconds": time.time() - t, "reference_threads": threads})
    if orc.ref_available("libref_patricia.so"):
        t = time.time()
        trie = orc.RefPatricia(cw, cko, ck)
        tb = time.time() - t
        n4, dt4, (c4, tot4) = _sized(lambda x: trie.count(x, threads), q, share)
        trie.close()
        out["reference_patricia"] = {
            "value": n4 / dt4, "unit": "queries/s", "cores": threads, "kind": "reference",
            "sample": f"first {n4} queries; the reference's compact Patricia trie built from its sources "
                      f"(oracle/_ref), built in {tb:.1f}s, keys collected, sorted and deduplicated, {dt4:.1f}s"}
    log(f"cpu baseline: {out}")
    return out


def host_key_lists(ring, nq, steps):
    """KEYS_UNIQUE steps whose sorted key lists are copied to pinned host
    memory (tm_read_results) before the step counts as done."""
    if steps <= 0:
        return None
    ring.submit_resident(nq)
    st = ring.wait()
    n = int(st.n_results)
    buf = tm.PinnedRows(n)
    off = np.zeros(nq + 1, np.uint64)
    lib = tm.load()
    tt = []
    for _ in range(steps):
        t = time.perf_counter()
        ring.submit_resident(nq)
        ring.wait()
        tm.check(lib.tm_read_results(ring.h, tm._ptr(off), tm.C.c_void_p(buf.ptr), n), "tm_read_results")
        tt.append(time.perf_counter() - t)
    buf.close()
    ms = 1e3 * float(np.mean(tt))
    res = {"ms_per_step": ms, "queries_per_s": nq / (ms / 1e3), "keys": n, "bytes": 4 * n + 8 * (nq + 1),
           "gb_per_s": (4 * n + 8 * (nq + 1)) / (ms / 1e3) / 1e9,
           "note": "KEYS_UNIQUE step, then offsets and every key copied to pinned host memory"}
    log(f"full key lists to the host: {res}")
    return res


def profile_figures():
    """HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)
    and VALU instructions per launch (SQ_INSTS_VALU) of the match kernel in
    the default step and in the step without the full-key-set exit, from
    the summaries under profiles/ (empty when absent)."""
    out = {}
    for name in ("match_traffic.json", "match_counters.json"):
        p = os.path.join(ROOT, "profiles", name)
        if os.path.exists(p):
            try:
                out.update(json.load(open(p)))
            except Exception:
                pass
    return out


def unit_logs(fib, q, batches):
    """Plain-form unit logs (features, hits, ticks) over q at each batch size."""
    ring = tm.Ring(fib, len(q), schedule=tm.SCHED_PLAIN, log_units=1)
    cols = [[], [], [], []]
    for bs in batches:
        ring.configure(schedule=tm.SCHED_PLAIN, batch_size=bs, log_units=1)
        ring.submit(q)
        ring.wait()
        for c, x in zip(cols, ring.unit_log()[:4]):
            c.append(x)
    ring.close()
    return tuple(np.concatenate(c) for c in cols)


def cost_model_leg(fib, q, ring, nq, steps, base_ms, n_log):
    """Train the MLP on the plain form's unit logs (first n_log queries),
    report held-out MAPE (next n_log queries) against a rate model, and time
    the persistent step priced by the model."""
    t0 = time.time()
    fa, ha, ka, ta = unit_logs(fib, q[:n_log], (256, 128))
    fb, hb, kb, tb = unit_logs(fib, q[n_log:2 * n_log], (256,))
    t1 = time.time()
    m = tm.MLP(seed=5)
    loss = m.train(fa, tm.encode_targets(ha, ka, ta), epochs=150, batch=256, lr=2e-3)
    m.calibrate(fa, ha, ka, ta)
    t2 = time.time()
    ph, pk, pt = tm.decode_targets(m.forward(fb))

    def size(f):
        return (np.exp2(f[:, 2] * 12) - 1) * (np.exp2(f[:, 0] * 20) - 1)

    rate_h = ha.sum() / size(fa).sum()
    rate_t = ta.sum() / size(fa).sum()
    res = {
        "train_units": int(len(fa)), "heldout_units": int(len(fb)), "train_loss": loss,
        "mape_hits": tm.mape(ph, hb, floor=16), "mape_keys": tm.mape(pk, kb, floor=16),
        "mape_ticks": tm.mape(pt, tb, floor=16),
        "rate_model_mape_hits": tm.mape(rate_h * size(fb), hb, floor=16),
        "rate_model_mape_ticks": tm.mape(rate_t * size(fb), tb, floor=16),
        "total_hits_error": float(ph.sum() / max(hb.sum(), 1) - 1.0),
        "log_s": t1 - t0, "train_s": t2 - t1,
    }
    log(f"cost model: {res}")
    # the headline step priced by the model (tile order + pre-sizing)
    ring.set_cost_model(m)
    ring.submit_resident(nq)
    st = ring.wait()
    tt = []
    for i in range(steps):
        ts = time.perf_counter()
        ring.submit_resident(nq)
        st = ring.wait()
        tt.append((time.perf_counter() - ts) * 1e3)
    ring.set_cost_model(None)
    res.update({"ms_per_step_with_model": float(np.mean(tt)), "ms_per_step_without": base_ms,
                "speedup_from_model": base_ms / float(np.mean(tt)), "pricing_ms": st.model_ms,
                "pred_hits_error": st.pred_hits / max(st.n_hits, 1) - 1.0})
    log(f"model-priced step: {np.mean(tt):.1f} ms vs {base_ms:.1f} ms")
    m.close()
    return res


def uneven_leg(cw, cko, ck, args):
    """Uneven batching at 10M filters: partitions are runs of the sorted
    table with Zipf(1.2) sizes in [1k, 1M], each receives several batches of
    log-uniform size in [1, 4096] (explicit units, no lookup).  The
    persistent launch and the plain per-(partition, batch) form run the same
    units."""
    t0 = time.time()
    off = tm.zipf_runs(len(cw), 1000, 1_000_000, s=1.2, seed=3)
    fib, masks = tm.fib_from_runs(cw, cko, ck, off)
    parts, offs = [], [0]
    for r in range(args.uneven_rounds):
        up, uo = tm.uneven_units(len(off) - 1, 1, 4096, seed=100 + r)
        parts.append(up)
        offs.extend((offs[-1] + uo[1:].astype(np.int64)).tolist())
    up = np.concatenate(parts).astype(np.uint32)
    uo = np.array(offs, np.uint64).astype(np.uint32)
    nq = int(uo[-1])
    g = tm.Generator(seed=args.seed)
    q, _ = g.queries(nq, len(cw), first=5_000_000)
    log(f"uneven: {len(off) - 1} partitions (sizes {int(np.diff(off).min())}..{int(np.diff(off).max())}), "
        f"{len(up)} units, {nq} queries, setup {time.time() - t0:.1f}s")
    res = {"partitions": int(len(off) - 1), "units": int(len(up)), "queries": nq,
           "partition_sizes": [int(np.diff(off).min()), int(np.median(np.diff(off))), int(np.diff(off).max())]}
    ring = tm.Ring(fib, nq)
    ring.submit_units(q, up, uo)
    ring.wait()
    ref = ring.results()
    for name, sched, steps in (("persistent", tm.SCHED_PERSISTENT, 3), ("persistent_tuned", tm.SCHED_PERSISTENT, 3),
                               ("plain", tm.SCHED_PLAIN, 1)):
        ring.configure(schedule=sched)
        if name == "persistent_tuned":
            # blocks per work item by partition-size bucket, tuned on these units
            b0, b1, table = ring.autotune_buckets(up, uo, reps=2)
            res["bucket_table"] = [int(x) for x in table]
            res["bucket_tune_device_ms"] = [b0, b1]
            log(f"uneven bucket table {res['bucket_table']}: device {b0:.2f} -> {b1:.2f} ms")
        ring.submit_units(None, up, uo)
        ring.wait()
        tt = []
        for i in range(steps):
            t = time.perf_counter()
            ring.submit_units(None, up, uo)
            st = ring.wait()
            tt.append(time.perf_counter() - t)
        o, v = ring.results()
        same = bool(np.array_equal(o, ref[0]) and np.array_equal(v, ref[1]))
        res[name] = {"queries_per_s": nq / float(np.mean(tt)), "ms_per_step": 1e3 * float(np.mean(tt)),
                     "match_ms": st.match_ms, "launches": st.n_items, "same_results": same}
        log(f"uneven {name}: {res[name]}")
    res["speedup_persistent_vs_plain"] = res["persistent"]["queries_per_s"] / res["plain"]["queries_per_s"]
    res["speedup_tuned_vs_persistent"] = (res["persistent_tuned"]["queries_per_s"] /
                                          res["persistent"]["queries_per_s"])
    ring.close()
    fib.close()
    return res


def serving_leg(fib, args, n_filters):
    """Poisson arrivals of fixed-size requests through tm_server at 30/70/95%
    of the sustained peak, with inserts interleaved."""
    req = args.serve_request
    # the table's generator (the seed also fixes each tag's bits), 20% of the
    # queries supersets of a filter, drawn past the headline's queries
    g = tm.Generator(seed=args.seed, superset_frac=0.2)
    pool, _ = g.queries(1 << 17, n_filters, first=1 << 30)
    gi = tm.Generator(seed=args.seed)
    srv = tm.Server(fib, n_slots=args.serve_slots, slot_queries=args.serve_slot_queries,
                    step_graphs=args.serve_graphs)
    rng = np.random.default_rng(11)

    def take(i):
        a = (i * req) % (len(pool) - req)
        return pool[a:a + req]

    # closed-loop peak: a backlog of requests drained as fast as possible
    # (the first drain warms the slots' buffers and the caches)
    nback = max(64, int(300000 / req))
    for rep in range(2):
        t = time.perf_counter()
        ids = [srv.submit(take(i)) for i in range(nback)]
        for rid in ids:
            srv.wait(rid)
        peak = nback * req / (time.perf_counter() - t)
        for rid in ids:
            srv.release(rid)
    burst_peak = peak
    log(f"serving burst peak: {peak:.0f} queries/s with {req}-query requests")

    def poisson(rate_qps, dur, with_inserts):
        """Open-loop Poisson arrivals of requests at rate_qps queries/s for
        dur seconds (inserts interleaved when asked): latencies, inserts'
        latencies, achieved queries/s."""
        nonlocal ins_first
        rate = rate_qps / req
        n = max(1, int(rate * dur))
        arrive = np.cumsum(rng.exponential(1.0 / rate, n))
        ids, inserts = [], []
        next_ins = args.serve_insert_ms / 1e3
        t0 = time.perf_counter()
        for i in range(n):
            now = time.perf_counter() - t0
            if with_inserts and args.serve_insert and now >= next_ins:
                w, ko, k = gi.filters(args.serve_insert, first=ins_first)
                ins_first += args.serve_insert
                inserts.append(srv.insert(w, ko.astype(np.uint32), k, maxp=args.maxp))
                next_ins += args.serve_insert_ms / 1e3
            while now < arrive[i]:
                now = time.perf_counter() - t0
            ids.append(srv.submit(take(i)))
        for rid in ids + inserts:
            srv.wait(rid)
        wall = time.perf_counter() - t0
        lat = np.array([srv.info(rid).latency_ms for rid in ids])
        ins_lat = np.array([srv.info(rid).latency_ms for rid in inserts]) if inserts else np.zeros(1)
        for rid in ids + inserts:
            srv.release(rid)
        return lat, ins_lat, n * req / wall, len(inserts)

    # Sustained peak: the highest rate of a ladder (fractions of the burst
    # peak) that open-loop arrivals sustain: achieved >= 98% of offered and
    # a p99 within 2x that of the lowest rung.  The burst drain forms full
    # steps that Poisson arrivals do not, so it overstates this rate.
    ins_first = n_filters
    ladder = []
    base_p99 = None
    peak = 0.2 * burst_peak   # when not even the lowest rung is sustained
    for frac in (0.2, 0.25, 0.3, 0.35, 0.4, 0.45, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0):
        lat, _, ach, _ = poisson(frac * burst_peak, 1.0, False)
        p99 = float(np.percentile(lat, 99))
        base_p99 = base_p99 or p99
        ok = ach >= 0.98 * frac * burst_peak and p99 <= 2 * base_p99
        ladder.append({"offered_qps": frac * burst_peak, "achieved_qps": ach, "p99_ms": p99, "sustained": bool(ok)})
        if not ok:
            break
        peak = frac * burst_peak
    log(f"serving sustained peak: {peak:.0f} queries/s (ladder {[round(x['p99_ms'], 2) for x in ladder]})")

    levels = {}
    for load in (0.3, 0.7, 0.95):
        lat, ins_lat, ach, n_ins = poisson(load * peak, args.serve_seconds, True)
        n = len(lat)
        st_l = srv.stats()
        levels[f"{int(load * 100)}%"] = {
            "offered_qps": load * peak, "achieved_qps": ach, "requests": n,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "p999_ms": float(np.percentile(lat, 99.9)), "max_ms": float(lat.max()),
            "inserts": n_ins, "live_inserts_so_far": st_l["n_live_inserts"],
            "insert_p50_ms": float(np.percentile(ins_lat, 50)), "insert_max_ms": float(ins_lat.max())}
        log(f"serving {load:.0%}: {levels[f'{int(load * 100)}%']}")
    st = srv.stats()
    srv.close()
    return {"peak_qps": peak, "burst_peak_qps": burst_peak, "peak_ladder": ladder,
            "peak_note": "peak = highest Poisson rate sustained (achieved >= 98% of offered, p99 within 2x of "
                         "the lowest rung); burst_peak = a drained backlog of requests",
            "request_queries": req, "slots": args.serve_slots,
            "slot_queries": args.serve_slot_queries,
            "device_ms_per_step": st["device_ms_sum"] / max(st["n_steps"], 1),
            "insert_filters": args.serve_insert, "insert_every_ms": args.serve_insert_ms,
            "levels": levels, "steps": st["n_batches"], "budget_waits": st["n_budget_waits"],
            "requests_per_step": st["requests_per_batch_sum"] / max(st["n_batches"], 1)}


def big_table(n_filters, args, n_queries):
    """Generated table of n_filters, canonical and partitioned on the GPU."""
    t0 = time.time()
    g = tm.Generator(seed=args.seed)
    w, ko, k = g.filters(n_filters)
    q, _ = g.queries(n_queries, n_filters)
    log(f"generated {n_filters} filters in {time.time() - t0:.1f}s")
    cw, cko, ck = tm.canonicalize(w, ko, k)
    del w, ko, k
    log(f"canonical: {cw.shape[0]} distinct filters, {time.time() - t0:.1f}s")
    times = {}
    fib, perm, off, masks = tm.build_fib(cw, cko, ck, args.maxp, times=times)
    log(f"partitioned (GPU {times['partition']:.2f}s) and uploaded ({times['upload']:.2f}s): {len(off) - 1} "
        f"partitions, {time.time() - t0:.1f}s")
    return cw, cko, ck, q, fib, perm, off, masks, {"total": time.time() - t0, **times}


def oracle_sample(cw, cko, ck, q, n):
    """The oracle's sorted unique keys of the first n queries (all host threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    return orc.match(cw, cko, ck, q[:n], mode=0, threads=min(16, os.cpu_count() or 1))


def c4_leg(args):
    """Learned cost model at 50M filters: the MLP is trained on the plain
    form's own (partition, 256-query batch) unit logs -- hits, key instances
    and kernel ticks -- and reports held-out MAPE against a rate model.  It
    then prices the persistent step (pre-sized result buffers, tile order);
    throughput is compared with the plain form and with the unpriced step,
    cold (a new ring's first step, where an unpriced ring retries on
    overflow) and warm."""
    nq = args.queries
    cw, cko, ck, q, fib, perm, off, masks, setup_s = big_table(args.c4_filters, args, nq)
    res = {"leg": "c4", "filters": args.c4_filters, "distinct_filters": int(cw.shape[0]),
           "partitions": int(len(off) - 1), "queries": nq, "setup_seconds": setup_s}
    # ---- plain unit logs, chunk by chunk, until the row target or the time
    # limit, over queries of their own (the same generator, past the timed ones)
    ql, _ = tm.Generator(seed=args.seed).queries(args.c4_log_queries, args.c4_filters, first=1 << 32)
    chunk = 16384
    held_chunks = 2
    ring = tm.Ring(fib, chunk, schedule=tm.SCHED_PLAIN, batch_size=256, log_units=1)
    cols, held = [[], [], [], []], [[], [], [], []]
    rows, t0, c = 0, time.time(), 0
    plain_s, plain_q = 0.0, 0
    while c * chunk < len(ql):
        qa = ql[c * chunk:(c + 1) * chunk]
        ts = time.perf_counter()
        ring.submit(qa)
        ring.wait()
        plain_s += time.perf_counter() - ts
        plain_q += len(qa)
        f, h, k, t, _ = ring.unit_log()
        dest = held if c < held_chunks else cols
        for d, x in zip(dest, (f, h, k, t)):
            d.append(x)
        if c >= held_chunks:
            rows += len(f)
        c += 1
        log(f"c4 logs: chunk {c}, {rows} training rows, {time.time() - t0:.0f}s")
        if rows >= args.c4_rows or time.time() - t0 > args.c4_log_seconds:
            break
    ring.close()
    fa, ha, ka, ta = (np.concatenate(x) for x in cols)
    fb, hb, kb, tb = (np.concatenate(x) for x in held)
    res.update({"train_rows": int(len(fa)), "heldout_rows": int(len(fb)), "log_s": time.time() - t0,
                "log_queries": c * chunk, "row_target": args.c4_rows,
                "plain_queries_per_s": plain_q / plain_s})
    t1 = time.time()
    m = tm.MLP(seed=5)
    epochs = int(max(3, min(150, 600_000 // max(len(fa), 1))))
    loss = m.train(fa, tm.encode_targets(ha, ka, ta), epochs=epochs, batch=1024 if len(fa) > 100_000 else 256,
                   lr=2e-3)
    m.calibrate(fa, ha, ka, ta)
    ph, pk, pt = tm.decode_targets(m.forward(fb))

    def size(x):
        return (np.exp2(x[:, 2] * 12) - 1) * (np.exp2(x[:, 0] * 20) - 1)

    rate_h = ha.sum() / size(fa).sum()
    rate_t = ta.sum() / size(fa).sum()
    res.update({"epochs": epochs, "train_s": time.time() - t1, "train_loss": loss,
                "mape_hits": tm.mape(ph, hb, floor=16), "mape_keys": tm.mape(pk, kb, floor=16),
                "mape_ticks": tm.mape(pt, tb, floor=16),
                "rate_model_mape_hits": tm.mape(rate_h * size(fb), hb, floor=16),
                "rate_model_mape_ticks": tm.mape(rate_t * size(fb), tb, floor=16),
                "total_hits_error": float(ph.sum() / max(hb.sum(), 1) - 1.0)})
    log(f"c4 model: {res}")
    del fa, ha, ka, ta
    # ---- the persistent step over every query: cold and warm, with and without the model
    out = {}
    for name, model in (("unpriced", None), ("priced", m)):
        # every hit of every query is a list entry in HITS mode (tens of
        # thousands per query at this size): that step runs a prefix
        for mode, mname, n in ((tm.MODE_KEYS_UNIQUE, "keys_unique", nq), (tm.MODE_HITS, "hits", min(nq, chunk))):
            r = tm.Ring(fib, n, mode=mode)
            if model is not None:
                r.set_cost_model(model)
            ts = time.perf_counter()
            r.submit(q[:n])
            st = r.wait()
            cold = time.perf_counter() - ts
            tt = []
            for i in range(3):
                ts = time.perf_counter()
                r.submit_resident(n)
                st2 = r.wait()
                tt.append(time.perf_counter() - ts)
            key = f"{name}_{mname}"
            out[key] = {"queries": n, "cold_ms": 1e3 * cold, "cold_overflow_retry": int(st.overflow),
                        "warm_ms": 1e3 * float(np.mean(tt)), "queries_per_s": n / float(np.mean(tt)),
                        "pricing_ms": st2.model_ms, "pred_hits": st2.pred_hits, "hits": int(st2.n_hits)}
            if mode == tm.MODE_KEYS_UNIQUE:
                o, v = r.results()
                if name == "unpriced":
                    ref_o, ref_v = o, v
                else:
                    out[key]["same_results"] = bool(np.array_equal(o, ref_o) and np.array_equal(v, ref_v))
                del o, v
            log(f"c4 {key}: {out[key]}")
            r.close()
    res["steps"] = out
    # the oracle on a sample of the queries
    n_or = 64
    eo, ev = oracle_sample(cw, cko, ck, q, n_or)
    res["oracle_equal_sample"] = bool(np.array_equal(ref_o[:n_or + 1], eo) and
                                      np.array_equal(ref_v[:int(eo[-1])], ev))
    res["oracle_sample_queries"] = n_or
    pq = res["plain_queries_per_s"]
    res["speedup_priced_vs_plain"] = out["priced_keys_unique"]["queries_per_s"] / pq
    res["speedup_priced_vs_unpriced"] = (out["priced_keys_unique"]["queries_per_s"] /
                                         out["unpriced_keys_unique"]["queries_per_s"])
    res["cold_speedup_priced_vs_unpriced_hits"] = out["unpriced_hits"]["cold_ms"] / out["priced_hits"]["cold_ms"]
    m.close()
    fib.close()
    return res


def c5_leg(args):
    """The serving loop at 200M filters: burst and sustained peaks, then
    Poisson arrivals at 30/70/95% of the sustained peak with 10k-filter
    inserts every 100 ms; a request's results alone are checked against the
    oracle on a sample first."""
    cw, cko, ck, q, fib, perm, off, masks, setup_s = big_table(args.c5_filters, args,
                                                               max(4096, args.c5_step_queries))
    res = {"leg": "c5", "filters": args.c5_filters, "distinct_filters": int(cw.shape[0]),
           "partitions": int(len(off) - 1), "setup_seconds": setup_s, "device_bytes": int(fib.stats()["device_bytes"])}
    n_or = 32
    srv = tm.Server(fib, n_slots=args.serve_slots, slot_queries=args.serve_slot_queries,
                    step_graphs=args.serve_graphs)
    rid = srv.submit(q[:n_or])
    srv.wait(rid)
    _, o, v = srv.result(rid)
    srv.close()
    eo, ev = oracle_sample(cw, cko, ck, q, n_or)
    res["oracle_equal_sample"] = bool(np.array_equal(o, eo) and np.array_equal(v, ev))
    res["oracle_sample_queries"] = n_or
    log(f"c5 oracle sample: {res['oracle_equal_sample']}")
    del cw, cko, ck
    if args.c5_step_queries:
        # the headline step on this table: every query's key set in host
        # memory, packed, two rings alternating, steps replayed from graphs
        nq = args.c5_step_queries
        kwords = (int(fib.stats()["n_distinct_keys"]) + 31) // 32
        ms, mm, packed, _, prings, pbufs = packed_steps(fib, q[:nq], {"use_graph": 1}, 5, 2, kwords)
        res["step"] = {"queries": nq, "ms_per_step": ms, "queries_per_s": nq / (ms / 1e3),
                       "match_ms_median": float(np.median(mm)) if mm else None,
                       "packed_rows_per_step": packed[-1],
                       "note": "packed key-set delivery to pinned host memory, as the headline step; "
                               "match_ms is 0 for steps replayed from captured graphs"}
        log(f"c5 step: {res['step']}")
        for r in prings:
            r.close()
        del pbufs
    res["serving"] = serving_leg(fib, args, args.c5_filters)
    fib.close()
    return res


def copy_calibration(mib):
    """A kernel of known traffic: y = x + 1 over mib MiB of int32 (reads and
    writes mib MiB each), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = (mib << 20) // 4
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.add(x, 1, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": 3}), flush=True)


def lists_equal_sample(off, vals, idx, eo, ev):
    """Whether the lists of queries idx (in a whole-run off/vals) equal the
    oracle's lists eo/ev for those queries."""
    cnt = (off[idx + 1] - off[idx]).astype(np.int64)
    ecnt = np.diff(eo).astype(np.int64)
    if not np.array_equal(cnt, ecnt):
        return False
    pos = np.repeat(off[idx].astype(np.int64) - eo[:-1].astype(np.int64), ecnt) + np.arange(int(eo[-1]))
    return bool(np.array_equal(vals[pos], ev))


def delivered_steps(fib, q, cfg_kw, steps, warmup, kw):
    """The headline step: KEY_BITMAP rings whose submissions end with every
    query's key bitmap row copied into pinned host memory (tm_ring_set_host_rows).
    Two rings alternate so that one ring's copy overlaps the other's match.
    Returns (ms per step, match kernel ms per step, last ring's rows, rings)."""
    nq = q.shape[0]
    rings, bufs = [], []
    for _ in range(2):
        r = tm.Ring(fib, nq, tm.config(mode=tm.MODE_KEY_BITMAP, **cfg_kw))
        b = tm.PinnedRows(nq * kw)
        r.set_host_rows(b)
        r.submit(q)
        r.wait()
        rings.append(r)
        bufs.append(b)

    def run(k):
        match_ms = []
        rings[0].submit_resident(nq)
        for i in range(k):
            if i + 1 < k:
                rings[(i + 1) % 2].submit_resident(nq)
            st = rings[i % 2].wait()
            match_ms.append(st.match_ms)
        return match_ms, bufs[(k - 1) % 2]

    run(max(warmup, 1))
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    t0 = time.perf_counter()
    match_ms, last = run(steps)
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return ms, match_ms, last, rings, bufs


def packed_steps(fib, q, cfg_kw, steps, warmup, kw):
    """The headline step: KEY_BITMAP rings whose submissions end with every
    query's key set in pinned host memory in packed form
    (tm_ring_set_host_rows_packed): 4 bytes per query, and a row only for
    the queries whose key set is not the whole dictionary.  Two rings
    alternate.  Returns (ms per step, match kernel ms per step, packed rows
    per step, last ring's buffers, rings, buffers)."""
    nq = q.shape[0]
    rings, bufs = [], []
    for _ in range(2):
        r = tm.Ring(fib, nq, tm.config(mode=tm.MODE_KEY_BITMAP, **cfg_kw))
        b = tm.PackedRows(nq, kw)
        r.set_host_rows_packed(b)
        r.submit(q)
        r.wait()
        rings.append(r)
        bufs.append(b)

    def run(k):
        match_ms, packed = [], []
        rings[0].submit_resident(nq)
        for i in range(k):
            if i + 1 < k:
                rings[(i + 1) % 2].submit_resident(nq)
            st = rings[i % 2].wait()
            match_ms.append(st.match_ms)
            packed.append(int(st.n_results))
        return match_ms, packed

    run(max(warmup, 1))
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    t0 = time.perf_counter()
    match_ms, packed = run(steps)
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return ms, match_ms, packed, bufs[(steps - 1) % 2], rings, bufs


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported",
              file=sys.stderr)
        sys.exit(2)
    tm.init(0)
    if args.leg != "headline":
        res = c4_leg(args) if args.leg == "c4" else c5_leg(args)
        line = json.dumps(res)
        print(line, flush=True)
        if args.out:
            with open(args.out, "w") as f:
                f.write(line + "\n")
        return
    cw, cko, ck, q, fib, perm, part_off, masks, setup_s = setup(args)
    nq = q.shape[0]
    sched = tm.SCHED_PERSISTENT if args.schedule == "persistent" else tm.SCHED_PLAIN
    kw = dict(chunk_blocks=args.chunk, groups_per_cu=args.groups_per_cu, use_graph=int(args.graph))
    if sched == tm.SCHED_PLAIN or args.tile:
        kw["batch_size"] = args.batch if sched == tm.SCHED_PLAIN else args.tile
    cfg = tm.config(mode=tm.MODE_KEYS_UNIQUE, schedule=sched, **kw)
    ring = tm.Ring(fib, nq, cfg)
    # upload the queries once; every later step reads them from HBM
    t = time.time()
    ring.submit(q)
    st = ring.wait()
    log(f"first step (with H2D and capacity growth): {time.time() - t:.2f}s, overflow={st.overflow}")
    tuned = None
    if args.autotune and sched == tm.SCHED_PERSISTENT and not args.tile:
        # the autotuner (tm_ring_autotune) picks the tile size, waves per
        # workgroup, segment size and graph replay on this workload, before
        # the timed steps
        cands = [dict(batch_size=b, waves_per_group=w, chunk_blocks=c, use_graph=int(g or args.graph))
                 for b, w, c, g in AUTOTUNE_CANDIDATES]
        best, times = ring.autotune(nq, cands, reps=1)
        cfg = ring.cfg
        kw.update(batch_size=cands[best]["batch_size"], waves_per_group=cands[best]["waves_per_group"],
                  chunk_blocks=cands[best]["chunk_blocks"], use_graph=cands[best]["use_graph"])
        tuned = {"candidates": cands, "step_ms": [float(x) for x in times], "best": cands[best]}
        log(f"autotuned: {cands[best]} ({times[best]:.1f} ms); all {np.round(times, 1).tolist()}")
    for i in range(max(0, args.warmup - 1)):
        ring.submit_resident(nq)
        ring.wait()

    # ---- device-resident step: sorted key lists left in HBM
    counts = ring.lookup_counts()[:len(part_off) - 1]   # the FIB's spare last partition aside
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    dev_ms, stats = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        ring.submit_resident(nq)
        st = ring.wait()
        dev_ms.append((time.perf_counter() - ts) * 1e3)
        stats.append(st.as_dict())
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    device_step_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    st = stats[-1]
    log(f"device-resident KEYS_UNIQUE step: {device_step_ms:.2f} ms ({st['n_results']} keys)")

    # ---- headline: every query's key set delivered to host memory, packed
    n_keys = int(fib.stats()["n_distinct_keys"])
    kwords = (n_keys + 31) // 32
    ms_per_step, hl_match_ms, packed_rows, last_packed, prings, pbufs = packed_steps(fib, q, kw, args.steps,
                                                                                    args.warmup, kwords)
    qps = nq / (ms_per_step / 1e3)
    # the match kernel's duration for the roofline: device-resident steps,
    # where no other step's kernels share the GPU (the headline's two rings
    # overlap theirs).  A replayed graph records no event around the match
    # kernel, so those steps are timed again without the graph: the same
    # kernel over the same queries.
    kernel_runs = [x["match_ms"] for x in stats]
    if kw.get("use_graph"):
        ring.configure(tm.config(mode=tm.MODE_KEYS_UNIQUE, schedule=sched, **{**kw, "use_graph": 0}))
        kernel_runs = []
        for i in range(4):
            ring.submit_resident(nq)
            m = ring.wait().match_ms
            if i > 0:
                kernel_runs.append(m)
        ring.configure(cfg)
    kernel_ms = float(np.mean(kernel_runs))
    n_packed = packed_rows[-1]
    log(f"host-delivered step (packed): {ms_per_step:.2f} ms -> {qps / 1e6:.1f}M queries/s ({n_packed} rows not "
        f"the whole dictionary; match kernel alone {kernel_ms:.2f} ms)")
    # every row copied to the host, for comparison
    full_ms, full_match_ms, host_rows, hrings, hbufs = delivered_steps(fib, q, kw, args.steps, args.warmup, kwords)
    full_rows = {"ms_per_step": full_ms, "queries_per_s": nq / (full_ms / 1e3), "bytes_per_step": float(nq * kwords * 4),
                 "note": "the same steps delivering every query's full key bitmap row (tm_ring_set_host_rows)"}
    log(f"host-delivered step (every row): {full_ms:.2f} ms -> {nq / full_ms / 1e3:.1f}M queries/s")

    # ---- parity outside the timed region: the delivered rows equal the
    # KEYS_UNIQUE step's rows and the rows of the step that enumerates every
    # match; a sample of queries against the oracle
    parity = {}
    rows_ku = ring.key_bitmaps()
    delivered = last_packed.unpack(nq, n_packed, n_keys)
    parity["rows_equal_keys_unique_step"] = bool(np.array_equal(delivered, rows_ku))
    parity["full_rows_equal_keys_unique_step"] = bool(np.array_equal(host_rows.rows(nq, kwords), rows_ku))
    del rows_ku
    ring.configure(tm.config(mode=tm.MODE_KEYS_UNIQUE, schedule=sched, full_exit=2, **kw))
    ring.submit_resident(nq)
    ring.wait()
    ts = time.perf_counter()
    ring.submit_resident(nq)
    st_all = ring.wait()
    all_ms = (time.perf_counter() - ts) * 1e3
    parity["rows_equal_no_exit_step"] = bool(np.array_equal(ring.key_bitmaps(), delivered))
    n_hits = st_all.n_hits
    no_exit = {"ms_per_step": all_ms, "queries_per_s": nq / (all_ms / 1e3), "matches_per_s": n_hits / (all_ms / 1e3),
               "match_ms": st_all.match_ms, "hits_enumerated": st_all.n_hits, "default_hits_enumerated": st["n_hits"]}
    log(f"without the full-key-set exit: {all_ms:.1f} ms per step, {n_hits} matches")
    ring.configure(cfg)
    ring.submit_resident(nq)
    ring.wait()
    dictionary = fib.dictionary()
    if args.parity_queries > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        idx = np.unique(np.linspace(0, nq - 1, min(args.parity_queries, nq)).astype(np.int64))
        t = time.time()
        eo, ev = orc.match(cw, cko, ck, q[idx], mode=0)
        so, sv = tm.expand_key_bitmaps(delivered[idx], dictionary)
        parity["oracle_sample_queries"] = int(len(idx))
        parity["oracle_equal_delivered"] = bool(np.array_equal(so, eo) and np.array_equal(sv, ev))
        off_ku, vals_ku = ring.results()
        parity["oracle_equal_keys_unique"] = lists_equal_sample(off_ku, vals_ku, idx, eo, ev)
        del off_ku, vals_ku
        log(f"parity: {parity} (oracle {time.time() - t:.1f}s)")
    parity["ok"] = all(v for k, v in parity.items() if isinstance(v, bool))
    if not parity["ok"]:
        log("PARITY FAILURE: the delivered results differ from the reference check")

    # ---- the same results as full sorted key lists in host memory, and the
    # rows expanded into key lists on the host
    hk = host_key_lists(ring, nq, args.host_key_steps)
    del delivered
    exp_off = np.zeros(nq + 1, np.uint64)
    lib = tm.load()
    row_of = last_packed.row_of.array[:nq]
    prow = last_packed.rows.array[:max(n_packed, 1) * kwords]

    def expand(values, threads):
        return lib.tm_expand_key_rows_packed(tm._ptr(row_of), tm._ptr(prow), n_packed, nq, kwords,
                                             tm._ptr(dictionary), dictionary.size, tm._ptr(exp_off),
                                             tm._ptr(values) if values is not None else None,
                                             values.size if values is not None else 0, threads)
    tm.check(expand(None, 0), "expand")
    exp_vals = np.empty(int(exp_off[-1]), np.uint32)
    exp_vals[::1024] = 0       # touch the pages before timing
    threads = os.cpu_count() or 1
    tt = []
    for _ in range(2):
        t = time.perf_counter()
        tm.check(expand(exp_vals, threads), "expand")
        tt.append(time.perf_counter() - t)
    host_expand = {"queries_per_s": nq / min(tt), "seconds": min(tt), "threads": threads,
                   "keys": int(exp_off[-1]), "keys_per_s": int(exp_off[-1]) / min(tt),
                   "note": "the packed delivered rows expanded into sorted key lists (tm_expand_key_rows_packed)"}
    log(f"host expansion of the rows: {host_expand}")
    del exp_vals
    for r in hrings + prings:
        r.close()
    for b in hbufs + pbufs:
        b.close()

    # ---- roofline of the dominant kernel (k_match_q) and of the step
    sizes = np.diff(part_off).astype(np.float64)
    batches = np.ceil(counts / args.batch)
    bytes_bruteforce = float(np.sum(batches * 24.0 * sizes) + 24.0 * counts.sum())
    pair_tests = float(np.sum(sizes * counts))
    fs = fib.stats()
    n_blocks = int(fs["n_blocks"])
    # what the algorithm must move per launch: every block of the table once
    # (6 word planes, block info, key slots), the active queries (words and
    # bitmap rows in and out) and the partition table
    alg_bytes = float(n_blocks * (1536 + 32 + 512) + len(part_off) * 64)
    prof = profile_figures()
    traffic = prof.get("bytes_per_launch")
    kernel_s = kernel_ms / 1e3
    achieved = (traffic / kernel_s / 1e9) if traffic else alg_bytes / kernel_s / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_match_q", "kernel_ms": kernel_ms,
            "alg_bytes": alg_bytes, "alg_frac": alg_bytes / kernel_s / 1e9 / HBM_PEAK_GBS,
            "measured_valu_frac": (prof["valu_insts"] * 64 / kernel_s / 1e12 / VALU_PEAK_TOPS)
            if prof.get("valu_insts") else None,
            "bruteforce_bytes": bytes_bruteforce,
            "note": "achieved = HBM traffic of one launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated against a "
                    "copy kernel; profiles/match_traffic.json) / the launch's mean duration, measured here with HIP "
                    "events on the ring's stream over the device-resident timed steps.  alg_bytes = what the "
                    "algorithm must move: every block of the table once (words, info, key slots), the active queries' "
                    "words and rows, the partition table.  bruteforce_bytes = what the reference's "
                    "(partition, 256-query batch) units would read; block masks, the rich-filter pass and the "
                    "full-key-set exit skip most of it exactly"}
    step_bytes = float(nq * 4 + n_packed * kwords * 4)
    step_roof = {"bound": "pcie", "achieved": step_bytes / (ms_per_step / 1e3) / 1e9, "peak": PCIE_PEAK_GBS,
                 "unit": "GB/s", "frac": step_bytes / (ms_per_step / 1e3) / 1e9 / PCIE_PEAK_GBS,
                 "bytes_per_step": step_bytes,
                 "full_rows_frac": full_rows["bytes_per_step"] / (full_ms / 1e3) / 1e9 / PCIE_PEAK_GBS,
                 "note": "the headline step moves 4 bytes per query and a key bitmap row (4 x ceil(K/32) bytes) per "
                         "query whose key set is not the whole dictionary; full_rows_frac is the same for the form "
                         "that copies every row.  PCIe 5.0 x16 is 63 GB/s per direction"}
    for name, rr in (("roofline", roof), ("step_roofline", step_roof)):
        assert rr["frac"] <= 1.0, f"{name}: fraction {rr['frac']:.3f} above 1"
    for name, v in (("alg_frac", roof["alg_frac"]), ("measured_valu_frac", roof["measured_valu_frac"])):
        assert v is None or v <= 1.0, f"roofline {name} {v:.3f} above 1"
    if no_exit["match_ms"] > 0:
        ne_s = no_exit["match_ms"] / 1e3
        # brute-force-equivalent VALU rate / peak: above 1 because block masks
        # skip most pair tests; a ratio of work avoided, not a roofline fraction
        no_exit["bruteforce_equivalent_valu_ratio"] = OPS_PER_PAIR * pair_tests / ne_s / 1e12 / VALU_PEAK_TOPS
        if prof.get("no_full_exit_bytes_per_launch"):
            no_exit["measured_hbm_frac"] = prof["no_full_exit_bytes_per_launch"] / ne_s / 1e9 / HBM_PEAK_GBS
        if prof.get("no_full_exit_valu_insts"):
            no_exit["measured_valu_frac"] = prof["no_full_exit_valu_insts"] * 64 / ne_s / 1e12 / VALU_PEAK_TOPS
        for k in ("measured_hbm_frac", "measured_valu_frac"):
            assert no_exit.get(k) is None or no_exit[k] <= 1.0, f"no_full_exit {k} above 1"

    # ---- the plain single-launch form on a prefix of the same queries
    plain = None
    if args.plain_steps > 0 and sched == tm.SCHED_PERSISTENT:
        ring.configure(tm.config(mode=tm.MODE_KEYS_UNIQUE, schedule=tm.SCHED_PLAIN, batch_size=args.batch))
        npq = min(nq, args.plain_queries)
        tt = []
        for i in range(args.plain_steps):
            ts = time.perf_counter()
            ring.submit_resident(npq)
            pst = ring.wait()
            tt.append(time.perf_counter() - ts)
            log(f"plain step {i}: {npq} queries in {tt[-1] * 1e3:.1f} ms, {pst.n_items} launches")
        plain = {"queries_per_s": npq / np.mean(tt), "ms_per_step": 1e3 * np.mean(tt), "launches": pst.n_items,
                 "queries": npq, "match_ms": pst.match_ms, "matches": pst.n_hits}
    model = None
    if args.model and sched == tm.SCHED_PERSISTENT:
        ring.configure(cfg)
        model = cost_model_leg(fib, q, ring, nq, max(1, args.steps // 2), device_step_ms, 4096)
    ring.close()

    uneven = None
    if args.uneven_rounds > 0:
        uneven = uneven_leg(cw, cko, ck, args)

    serving = None
    if args.serve_seconds > 0:
        serving = serving_leg(fib, args, args.filters)

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(cw, cko, ck, perm, part_off, masks, q, args.cpu_seconds, args.maxp)

    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: seeded generator (Zipf tags, Bloom bits, keys); queries resident in HBM when timing "
                "starts, results delivered to host memory",
        "config": {
            "workload": f"{args.filters // 1_000_000}M filters (x 192-bit), {nq} queries, partitions <= {args.maxp}, "
                        f"batches of {args.batch} queries per partition; a step ends with every query's sorted "
                        f"unique key set in host memory (packed key bitmap rows over the sorted dictionary)",
            "filters": args.filters, "distinct_filters": int(cw.shape[0]), "queries": nq,
            "partitions": int(len(part_off) - 1), "maxp": args.maxp, "batch": args.batch,
            "schedule": args.schedule, "graph": bool(kw.get("use_graph")),
            "result_form": "packed key bitmap rows in host memory",
        },
        "value_note": "queries/s with every query's match_unique result (sorted unique keys) in host memory, packed "
                      f"lossless: a 4-byte entry per query, TM_ROW_FULL when its key set is the whole dictionary, "
                      f"else the index of its {kwords * 4}-byte key bitmap row over the FIB's sorted dictionary; two "
                      "rings alternate so that one step's copies overlap the next step's match",
        "packed_rows_per_step": n_packed,
        "full_rows_delivered": full_rows,
        "device_step_ms": device_step_ms,
        "device_step_queries_per_s": nq / (device_step_ms / 1e3),
        "device_step_note": "KEYS_UNIQUE step whose sorted key lists stay in HBM",
        "host_key_lists": hk,
        "host_expand": host_expand,
        "parity_checked": parity,
        "matches_per_s": no_exit["matches_per_s"],
        "matches_note": "every (query, filter) match enumerated, in a step without the full-key-set exit",
        "keys_per_s": st["n_results"] / (ms_per_step / 1e3),
        "hits_per_query": n_hits / nq,
        "no_full_exit": no_exit,
        "p99_latency_ms": serving["levels"]["70%"]["p99_ms"] if serving else None,
        "latency_note": "p99 of 64-query requests through the serving loop at 70% of its sustained peak, "
                        "10k-filter inserts every 100 ms",
        "step_latency_ms": {"p50": float(np.percentile(dev_ms, 50)), "max": float(np.max(dev_ms))},
        "serving": serving,
        "cost_model": model,
        "uneven_batching": uneven,
        "autotune": tuned,
        "roofline": roof,
        "step_roofline": step_roof,
        "roofline_valu": {"bound": "valu", "pair_tests_per_launch": pair_tests,
                          "bruteforce_ops_per_launch": OPS_PER_PAIR * pair_tests,
                          "measured_frac": roof["measured_valu_frac"],
                          "note": "measured_frac = SQ_INSTS_VALU x 64 / kernel time / 39.3 T lane-ops/s"},
        "plain": plain,
        "speedup_vs_plain": (qps / plain["queries_per_s"]) if plain else None,
        "speedup_vs_plain_same_algorithm": (no_exit["queries_per_s"] / plain["queries_per_s"]) if plain else None,
        "plain_note": "the plain form is the reference's dispatch shape (one launch per partition and 256-query "
                      "batch, no rich-filter pass or full-key-set exit) on a prefix of the queries; the same-"
                      "algorithm ratio compares it with the persistent step without the exit",
        "cpu_baseline": cpu,
        "setup_seconds": setup_s,
        "step_stats": stats[-1],
    }
    line = json.dumps(out)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
