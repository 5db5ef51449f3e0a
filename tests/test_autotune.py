"""The autotuner (tm_ring_autotune) picks one of the candidates, applies it,
and the results stay bit-identical to the oracle."""
import numpy as np
import pytest

import oracle as orc
import tagmatch_amd as tm

pytestmark = pytest.mark.gpu


def test_autotune_picks_and_keeps_results(gpu):
    cw, ck, cv, q, _ = tm.workload(50000, 20000, seed=701)
    fib, perm, off, masks = tm.build_fib(cw, ck, cv, 5000)
    ring = tm.Ring(fib, len(q))
    ring.submit(q)
    ring.wait()
    cands = [dict(variant=v, batch_size=b, waves_per_group=w)
             for v in (2, 6) for b in (64, 128, 256) for w in (4, 8)]
    best, times = ring.autotune(len(q), cands, reps=1)
    print(f"best {cands[best]} times {np.round(times, 3)}", flush=True)
    assert 0 <= best < len(cands)
    assert np.all(times > 0) and times[best] == times.min()
    assert ring.cfg.batch_size == cands[best]["batch_size"]
    ring.submit_resident(len(q))
    ring.wait()
    o, v = ring.results()
    eo, ev = orc.match(cw, ck, cv, q, mode=0)
    assert np.array_equal(o, eo) and np.array_equal(v, ev)
