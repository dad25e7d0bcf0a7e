"""GPU test of the serving loop: continuous batching under an HBM budget,
every served result equal to the same request run alone."""
import numpy as np
import pytest

from conftest import gpu_available
from pfamd import capi
from pfamd.serve import instance_bytes, request_stream, serve_batched, serve_sequential

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a gfx950 GPU")]


def test_serving_results_equal_alone():
    reqs = request_stream(n=8, rate=40.0, seed=5, steps_lo=30, steps_hi=120, edge_lo=24,
                          edge_hi=48, max_seeds=4)
    sizes = sorted(instance_bytes(r.problem) for r in reqs)
    budget = sizes[-1] + sizes[-2] + sizes[-3]  # about three at a time
    print(f"[serve] 8 requests, budget {budget} B", flush=True)
    rep = serve_batched(reqs, hbm_budget=budget, tick_steps=16, graph_steps=4, keep=True)
    assert len(rep.completions) == 8
    assert rep.peak_bytes <= budget and rep.peak_instances >= 2  # runs overlapped
    # admitted in arrival order except where a later request that fits
    # overtook one that did not
    by_arrival = sorted(rep.completions, key=lambda c: c.arrival)
    adm = [c.admitted for c in by_arrival]
    assert rep.overtakes > 0 or adm == sorted(adm)
    alone = serve_sequential(reqs, keep=True)
    ref = {c.rid: c for c in alone.completions}
    for c in rep.completions:
        a = ref[c.rid]
        for q, nm in enumerate(("phi", "c", "T")):
            assert np.array_equal(c.fields[q], a.fields[q]), (c.rid, nm)
        assert np.array_equal(c.diag, a.diag), c.rid
    s = rep.summary()
    print(f"  batched: {s}\n  sequential: {alone.summary()}", flush=True)
    assert s["p99_s"] >= s["p50_s"] > 0
