"""Served what-if stream (continuous batching) on the GPU.

A query's answer (joints failed by mode, fragment table, bricks standing, pieces) must be the same whether the query
runs alone or shares its frames with others admitted before, with and after it: bit for bit, since every order in a
frame is scene local. The load-path solve of each query must reach its FP64 residual target.
"""
import numpy as np
import pytest

from shatterforge_amd import serve

pytestmark = pytest.mark.gpu


def _answers(done):
    return {q["id"]: q["result"] for q in done}


def test_whatif_same_alone_and_in_mix():
    print("case: 5 queries of 2k-5k bricks, 90 frames, 3 slots staggered vs each alone", flush=True)
    qs = [serve.make_query(i, 1000 + i, n_min=2000, n_max=5000) for i in range(5)]
    srv = serve.WhatIfServer(slots=3, slot_bricks=6000, slot_bodies=2048, frames=90)
    srv.submit(qs[0], 0.0)
    srv.submit(qs[1], 0.0)
    for _ in range(20):
        srv.tick()
    for q in qs[2:]:
        srv.submit(q, 0.0)
    while srv.busy():
        srv.tick()
    mixed = _answers(srv.done)
    assert len(mixed) == 5
    assert all(q["load_resid"] <= 1e-6 for q in srv.done)
    srv.close()
    for q in qs:
        alone = serve.WhatIfServer(slots=1, slot_bricks=6000, slot_bodies=2048, frames=90)
        alone.submit(q, 0.0)
        while alone.busy():
            alone.tick()
        a = _answers(alone.done)[q["id"]]
        alone.close()
        assert a == mixed[q["id"]], f"query {q['id']}: alone {a} vs in the mix {mixed[q['id']]}"
    assert any(sum(m["failed"]) > 0 or m["pieces"] > 0 for m in mixed.values())


def test_whatif_poisson_stream_completes():
    print("case: 16 queries of 2k-6k bricks at 200 q/s, 300 frames each", flush=True)
    qs = [serve.make_query(i, 2000 + i, n_min=2000, n_max=6000) for i in range(16)]
    arr = serve.poisson_arrivals(200.0, len(qs), seed=3)
    srv = serve.WhatIfServer(slots=16, slot_bricks=6144, slot_bodies=2048, frames=300)
    done = srv.run_stream(qs, arr)
    lat = np.array([q["latency"] for q in done])
    print(f"p50 {np.percentile(lat, 50) * 1e3:.1f} ms p99 {np.percentile(lat, 99) * 1e3:.1f} ms, "
          f"{srv.brick_steps / max(lat.max(), 1e-9):.3e} brick-steps/s", flush=True)
    assert len(done) == len(qs) and np.isfinite(lat).all()
    srv.close()
