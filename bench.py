#!/usr/bin/env python3
"""Throughput benchmark of the MI355X TagMatch engine on one GPU.

Workload (the metric's configuration): 10M generated 192-bit filters (1-10
Zipf tags x 7 bits, 1-4 keys each, identical filters merged), partitioned on
the GPU into partitions of at most 256k filters; 1M generated queries (10-30
Zipf tags, 30% supersets of a filter), matched in batches of 256 queries per
partition.  A step is one pass of the whole engine over the 1M queries
already resident in HBM: on-device partition lookup, batched subset match,
key expansion and per-query dedup into sorted key lists.

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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--filters", type=int, default=10_000_000)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--maxp", type=int, default=256 * 1024)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--chunk", type=int, default=0, help="blocks per work item (0: default)")
    ap.add_argument("--groups-per-cu", type=int, default=0)
    ap.add_argument("--schedule", choices=["persistent", "plain"], default="persistent")
    ap.add_argument("--graph", action="store_true", help="replay steps from a hipGraph")
    ap.add_argument("--plain-steps", type=int, default=1, help="steps of the plain single-launch form")
    ap.add_argument("--plain-queries", type=int, default=65536,
                    help="queries of the plain form's step (a prefix of the resident queries)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--seed", type=lambda x: int(x, 0), default=0x7461676D61746302)
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
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
    fib, perm, off, masks = tm.build_fib(cw, cko, ck, args.maxp)
    t3 = time.time()
    log(f"setup: generate {t1 - t0:.1f}s canonicalize {t2 - t1:.1f}s partition+upload {t3 - t2:.1f}s; "
        f"{cw.shape[0]} distinct filters, {len(off) - 1} partitions")
    return cw, cko, ck, q, fib, perm, off, masks, t3 - t0


def cpu_baseline(cw, off_parts, masks, perm_words, q, budget_s):
    """The oracle's CPU schedule (partition lookup + per-partition brute
    force, restating the reference's CPU back end) on a bounded sample of
    the same queries, on the host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    n = 16
    t = time.time()
    orc.schedule_count(perm_words, off_parts, masks, q[:n], threads)
    dt = max(time.time() - t, 1e-3)
    n2 = int(min(len(q), max(n, n * budget_s / dt)))
    t = time.time()
    hits = orc.schedule_count(perm_words, off_parts, masks, q[:n2], threads)
    dt = time.time() - t
    return {"value": n2 / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {n2} of the {len(q)} queries against all {cw.shape[0]} filters "
                      f"(oracle lookup + per-partition brute force, {threads} threads, {dt:.1f}s)",
            "hits_per_query": hits / n2}


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "match_traffic.json")
    if os.path.exists(p):
        try:
            return json.load(open(p)).get("bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported",
              file=sys.stderr)
        sys.exit(2)
    tm.init(0)
    cw, cko, ck, q, fib, perm, part_off, masks, setup_s = setup(args)
    nq = q.shape[0]
    sched = tm.SCHED_PERSISTENT if args.schedule == "persistent" else tm.SCHED_PLAIN
    cfg = tm.config(mode=tm.MODE_KEYS_UNIQUE, schedule=sched, batch_size=args.batch, chunk_blocks=args.chunk,
                    groups_per_cu=args.groups_per_cu, use_graph=int(args.graph))
    ring = tm.Ring(fib, nq, cfg)
    # upload the queries once; every later step reads them from HBM
    t = time.time()
    ring.submit(q)
    st = ring.wait()
    log(f"first step (with H2D and capacity growth): {time.time() - t:.2f}s, overflow={st.overflow}")
    for i in range(max(0, args.warmup - 1)):
        ring.submit_resident(nq)
        ring.wait()

    counts = ring.lookup_counts()
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    step_ms, match_ms, stats = [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        ring.submit_resident(nq)
        st = ring.wait()
        step_ms.append((time.perf_counter() - ts) * 1e3)
        match_ms.append(st.match_ms)
        stats.append(st.as_dict())
        log(f"step {i}: {step_ms[-1]:.1f} ms (match kernel {st.match_ms:.1f} ms, device {st.device_ms:.1f} ms, "
            f"{st.n_items} items, {st.n_hits} hits)")
    tm.check(tm.load().tm_dev_synchronize(), "sync")
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    qps = nq / (ms_per_step / 1e3)
    st = stats[-1]
    n_hits = st["n_hits"]

    # algorithmic work of the (partition, batch) units of this workload
    sizes = np.diff(part_off).astype(np.float64)
    batches = np.ceil(counts / args.batch)
    bytes_alg = float(np.sum(batches * 24.0 * sizes) + 24.0 * counts.sum())
    pair_tests = float(np.sum(sizes * counts))
    kernel_s = (np.mean(match_ms) / 1e3) if match_ms[-1] > 0 else ms_per_step / 1e3
    hbm_gbs = bytes_alg / kernel_s / 1e9
    valu_tops = OPS_PER_PAIR * pair_tests / kernel_s / 1e12
    traffic = traffic_from_profiles()

    # the plain single-launch form on the same queries (reference dispatch shape)
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
                 "queries": npq, "match_ms": pst.match_ms}
    ring.close()

    cpu = None
    if args.cpu_seconds > 0:
        # the same partitions, in partitioned order, for the CPU schedule
        cpu = cpu_baseline(cw, part_off, masks, cw[perm], q, args.cpu_seconds)

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
        "data": "synthetic: seeded generator (Zipf tags, Bloom bits, keys), inputs resident in HBM",
        "config": {
            "workload": f"{args.filters // 1_000_000}M filters (x 192-bit), {nq} queries, partitions <= {args.maxp}, "
                        f"batches of {args.batch} queries per partition, KEYS_UNIQUE output",
            "filters": args.filters, "distinct_filters": int(cw.shape[0]), "queries": nq,
            "partitions": int(len(part_off) - 1), "maxp": args.maxp, "batch": args.batch,
            "schedule": args.schedule, "graph": bool(args.graph),
        },
        "matches_per_s": n_hits / (ms_per_step / 1e3),
        "keys_per_s": st["n_results"] / (ms_per_step / 1e3),
        "hits_per_query": n_hits / nq,
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99)),
                       "what": "completion time of a 1M-query step"},
        "roofline": {"bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hbm_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_match", "kernel_ms": kernel_s * 1e3,
                     "bytes_per_launch": bytes_alg},
        "roofline_valu": {"bound": "valu", "achieved": valu_tops, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                          "frac": valu_tops / VALU_PEAK_TOPS, "pair_tests_per_launch": pair_tests,
                          "note": "algorithmic int ops (7 per pair-test over selected partitions); "
                                  "block masks let the kernel skip most of them"},
        "plain": plain,
        "speedup_vs_plain": (qps / plain["queries_per_s"]) if plain else None,
        "cpu_baseline": cpu,
        "setup_s": setup_s,
        "step_stats": stats[-1],
    }
    line = json.dumps(out)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
