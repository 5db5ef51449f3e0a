"""The learned cost model on the GPU: the HIP forward, backward and Adam
kernels against the CPU oracle (oracle/mlp_oracle.py), the plain schedule's
unit log against brute force, and what a trained model gains: per-unit
predictions better than a rate model, no overflow retry on a fresh ring,
and unchanged results."""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("shape", [(16, 64, 3), (5, 7, 2)])
def test_forward_matches_oracle(gpu, shape):
    log(f"mlp forward {shape}")
    m = tm.MLP(*shape, seed=11)
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (777, shape[0])).astype(np.float32)
    y = m.forward(x)
    ye, _ = mo.forward(m.params(), x, shape)
    np.testing.assert_allclose(y, ye, rtol=1e-4, atol=1e-4 * np.abs(ye).max())


@pytest.mark.parametrize("shape,n", [((16, 64, 3), 300), ((5, 7, 2), 41)])
def test_loss_and_gradient_match_oracle(gpu, shape, n):
    log(f"mlp loss/grad {shape} n={n}")
    m = tm.MLP(*shape, seed=12)
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, (n, shape[0])).astype(np.float32)
    t = rng.normal(0, 1, (n, shape[2])).astype(np.float32)
    loss, g = m.loss_grad(x, t)
    le, ge = mo.loss_grad(m.params(), x, t, shape)
    assert loss == pytest.approx(le, rel=1e-4)
    np.testing.assert_allclose(g, ge, rtol=2e-3, atol=2e-5 * np.abs(ge).max())


def test_adam_step_matches_oracle(gpu):
    shape = (16, 64, 3)
    m = tm.MLP(*shape, seed=13)
    rng = np.random.default_rng(3)
    x = rng.normal(0, 1, (256, 16)).astype(np.float32)
    t = rng.normal(0, 1, (256, 3)).astype(np.float32)
    p0 = m.params().astype(np.float64)
    _, g = mo.loss_grad(p0, x, t, shape)
    lr = 1e-3
    m.train(x, t, epochs=1, batch=256, lr=lr)          # one full-batch step
    p1 = m.params()
    pe, _, _ = mo.adam_step(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, lr)
    big = np.abs(g) > 1e-4 * np.abs(g).max()           # away from g = 0, where sign(g) is fragile
    np.testing.assert_allclose(p1[big], pe[big], rtol=0, atol=1e-6)
    assert np.all(np.abs(p1 - p0) <= lr * 1.001 + 1e-7)


def test_training_reduces_loss(gpu):
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (4096, 16)).astype(np.float32)
    t = np.stack([x[:, :4].sum(1) / 4, np.sin(3 * x[:, 5]) * 0.5 + 0.5, x[:, 7] * x[:, 8]], 1).astype(np.float32)
    m = tm.MLP(seed=14)
    l0, _ = m.loss_grad(x, t)
    l1 = m.train(x, t, epochs=30, batch=256, lr=3e-3)
    log(f"loss {l0:.4g} -> {l1:.4g}")
    assert l1 < 0.05 * l0


# ----------------------------------------------------------------- unit log

def small_table(n_filters, maxp, seed, with_perm=False):
    cw, ck, cv, _, _ = tm.workload(n_filters, 1, seed=seed)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, maxp)
    pw = cw[perm]
    if with_perm:
        return cw, ck, cv, fib, pw, off, masks, perm
    return cw, ck, cv, fib, pw, off, masks


def test_plain_unit_log_matches_brute_force(gpu):
    cw, ck, cv, fib, pw, off, masks, perm = small_table(20000, 1500, 301, with_perm=True)
    _, _, _, q, _ = tm.workload(20000, 200, seed=301)
    log(f"unit log: {len(off) - 1} partitions, {len(q)} queries")
    ring = tm.Ring(fib, len(q), schedule=tm.SCHED_PLAIN, batch_size=256, log_units=1)
    ring.submit(q)
    ring.wait()
    f, h, kk, t, u = ring.unit_log()
    orw, cnt, pop = mo.partition_stats(pw, off)
    # 200 queries, batch 256: every partition is one batch of all its queries
    assert len(u) == len(set(u[:, 0].tolist()))
    for k in range(len(u)):
        p, b, nb = (int(v) for v in u[k])
        assert b == 0
        a, e = int(off[p]), int(off[p + 1])
        sel = np.all((masks[p][None, :] & ~q) == 0, axis=1)
        assert nb == int(sel.sum())
        fw = pw[a:e]
        hitmask = np.array([np.all((fw & ~qq) == 0, axis=1) for qq in q[sel]]).reshape(-1, e - a)
        hits = int(hitmask.sum())
        assert int(h[k]) == hits, (p, int(h[k]), hits)
        fid = perm[a:e]                              # canonical ids of the partition's filters
        nkeys = (ck[fid + 1] - ck[fid]).astype(np.int64)
        assert int(kk[k]) == int((hitmask * nkeys[None, :]).sum())
        assert int(t[k]) > 0
        fe = mo.unit_features(masks[p], orw[p], cnt[p], pop[p], q[sel])
        np.testing.assert_allclose(f[k], fe, rtol=1e-5, atol=1e-6)


def collect_logs(fib, queries, batch_sizes=(64, 128, 256)):
    """Plain-schedule unit logs at several batch sizes."""
    feats, hits, keys, ticks = [], [], [], []
    ring = tm.Ring(fib, len(queries), schedule=tm.SCHED_PLAIN, log_units=1)
    for bs in batch_sizes:
        ring.configure(schedule=tm.SCHED_PLAIN, batch_size=bs, log_units=1)
        ring.submit(queries)
        ring.wait()
        f, h, k, t, _ = ring.unit_log()
        feats.append(f)
        hits.append(h)
        keys.append(k)
        ticks.append(t)
    ring.close()
    return tuple(np.concatenate(x) for x in (feats, hits, keys, ticks))


@pytest.fixture(scope="module")
def trained(gpu):
    cw, ck, cv, fib, pw, off, masks = small_table(200000, 8000, 302)
    _, _, _, qa, _ = tm.workload(200000, 4096, seed=302, query_first=0)
    _, _, _, qb, _ = tm.workload(200000, 4096, seed=302, query_first=10**6)
    log(f"cost model fixture: {len(off) - 1} partitions")
    fa, ha, ka, ta = collect_logs(fib, qa)
    fb, hb, kb, tb = collect_logs(fib, qb)
    m = tm.MLP(seed=21)
    loss = m.train(fa, tm.encode_targets(ha, ka, ta), epochs=60, batch=512, lr=2e-3)
    m.calibrate(fa, ha, ka, ta)
    log(f"trained on {len(fa)} units, loss {loss:.4g}; held-out {len(fb)} units")
    return dict(cw=cw, ck=ck, cv=cv, fib=fib, pw=pw, off=off, masks=masks, qa=qa, qb=qb, m=m,
                fa=fa, ha=ha, ka=ka, ta=ta, fb=fb, hb=hb, kb=kb, tb=tb)


def test_model_beats_rate_baseline(trained):
    """Held-out per-unit MAPE of the network against the method without it:
    a rate model hits = r * B * n_p, r fitted on the training units."""
    d = trained
    ph, pk, pt = tm.decode_targets(d["m"].forward(d["fb"]))

    def sizes(f):
        return (np.exp2(f[:, 2] * 12) - 1) * (np.exp2(f[:, 0] * 20) - 1)

    r = d["ha"].sum() / sizes(d["fa"]).sum()
    rt = d["ta"].sum() / sizes(d["fa"]).sum()
    base_h = r * sizes(d["fb"])
    base_t = rt * sizes(d["fb"])
    mh, bh = tm.mape(ph, d["hb"], floor=16), tm.mape(base_h, d["hb"], floor=16)
    mt, bt = tm.mape(pt, d["tb"], floor=16), tm.mape(base_t, d["tb"], floor=16)
    log(f"MAPE hits: model {mh:.3f} vs rate {bh:.3f}; ticks: model {mt:.3f} vs rate {bt:.3f}")
    assert mh < bh
    assert mt < bt


def tile_rows(pw, off, masks, q, nqt):
    orw, cnt, pop = mo.partition_stats(pw, off)
    rows = []
    for t0 in range(0, len(q), nqt):
        qt = q[t0:t0 + nqt]
        for p in range(len(off) - 1):
            sel = np.all((masks[p][None, :] & ~qt) == 0, axis=1)
            rows.append(mo.unit_features(masks[p], orw[p], cnt[p], pop[p], qt[sel]))
    return np.array(rows)


def test_predicted_hits_match_oracle_pricing(trained):
    """The ring's on-device pricing (tile features, fused inference, tile
    sums) against the same pricing done by the oracle."""
    d = trained
    q = d["qb"][:1000]
    ring = tm.Ring(d["fib"], len(q), batch_size=128)
    ring.set_cost_model(d["m"], tm.COST_PRESIZE)
    ring.submit(q)
    st = ring.wait()
    x = tile_rows(d["pw"], d["off"], d["masks"], q, 128)
    y, _ = mo.forward(d["m"].params(), x, (16, 64, 3))
    h, _, _ = tm.decode_targets(y)
    expect = h[x[:, 2] > 0].sum()
    log(f"pred_hits {st.pred_hits:.1f} oracle {expect:.1f} actual {st.n_hits}")
    assert st.pred_hits == pytest.approx(expect, rel=1e-3)


@pytest.mark.parametrize("mode", [tm.MODE_HITS, tm.MODE_KEYS_ALL, tm.MODE_KEYS_UNIQUE])
def test_presize_avoids_retry_and_keeps_results(trained, mode):
    d = trained
    q = d["qb"]
    cap = 4096
    base = tm.Ring(d["fib"], len(q), mode=mode, hit_capacity=cap)
    base.submit(q)
    sb = base.wait()
    ob, vb = base.results()
    ring = tm.Ring(d["fib"], len(q), mode=mode, hit_capacity=cap)
    ring.set_cost_model(d["m"])                         # pre-size and order
    ring.submit(q)
    sm = ring.wait()
    om, vm = ring.results()
    log(f"mode {mode}: no model overflow={sb.overflow} {sb.device_ms:.2f} ms; "
        f"model overflow={sm.overflow} {sm.device_ms:.2f} ms pred {sm.pred_hits:.0f} actual {sm.n_hits}")
    # a KEYS_UNIQUE ring whose worst case (queries x keys) fits 64M values
    # is sized for it and never retries; the hit-list modes need the model
    if mode == tm.MODE_KEYS_UNIQUE and len(q) * d["fib"].stats()["n_distinct_keys"] <= (1 << 26):
        assert sb.overflow == 0
    else:
        assert sb.overflow == 1
    assert sm.overflow == 0
    assert np.array_equal(ob, om)
    if mode == tm.MODE_KEYS_UNIQUE:
        assert np.array_equal(vb, vm)
    else:
        for i in range(len(q)):
            assert np.array_equal(np.sort(vb[ob[i]:ob[i + 1]]), np.sort(vm[om[i]:om[i + 1]]))


def test_ordered_schedule_matches_oracle(trained):
    """Tile order from the model (enough tiles that ordering is used) gives
    the same results as the identity order, which the oracle checks on a
    sample."""
    d = trained
    _, _, _, q, _ = tm.workload(200000, 262144, seed=302, query_first=2 * 10**6)
    n_check = 3000
    eo, ev = orc.match(d["cw"], d["ck"], d["cv"], q[:n_check], mode=0)
    base = tm.Ring(d["fib"], len(q), batch_size=64)
    base.submit(q)
    base.wait()
    ob, vb = base.results()
    assert np.array_equal(ob[:n_check + 1], eo) and np.array_equal(vb[:int(eo[-1])], ev)
    for graph in (0, 1):
        # one block segment per tile: the schedule the model orders
        ring = tm.Ring(d["fib"], len(q), batch_size=64, use_graph=graph, chunk_blocks=1 << 30)
        ring.set_cost_model(d["m"], tm.COST_ORDER)
        for rep in range(3):
            ring.submit(q)
            st = ring.wait()
            off, vals = ring.results()
            log(f"graph {graph} rep {rep}: {st.match_ms:.2f} ms, model {st.model_ms:.2f} ms")
            assert np.array_equal(off, ob) and np.array_equal(vals, vb), (graph, rep)
        ring.set_cost_model(None)
        ring.submit(q)
        ring.wait()
        off, vals = ring.results()
        assert np.array_equal(off, ob) and np.array_equal(vals, vb)
