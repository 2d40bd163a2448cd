"""Placement serving: continuous batching, admission by HBM budget, and each
result identical alone and in the mix."""
import time

import numpy as np
import pytest

import fluxion as fxn

pytestmark = pytest.mark.gpu


def weight_for(G):
    return fxn.netgen.density_weight_for(G)


def requests(k, seed0, sizes=(1000, 3000, 8000, 20000, 2000, 12000)):
    out = []
    for r in range(k):
        n = sizes[r % len(sizes)]
        G = (64, 128, 0, 256)[r % 4]
        circ, x, y = fxn.netgen.flat(n, n * 11 // 10, seed0 + r)
        cfg = fxn.PlacerConfig(density_grid=G, density_weight=weight_for(G))
        out.append((circ, cfg, x, y))
    return out


def alone(ctx, circ, cfg, x, y, iters, graph):
    job = fxn.FlatJob(ctx, circ, cfg)
    job.set_positions(x, y)
    if graph:
        job.use_graph(True)
    job.run(iters)
    px, py = job.positions()
    h = job.hpwl()
    job.close()
    return px, py, h


def test_results_identical_alone_and_in_mix(ctx):
    print("case: server mix", flush=True)
    reqs = requests(12, fxn.netgen.SERVE_SEED_BASE)
    srv = fxn.PlacementServer(ctx, max_active=6, hbm_budget=8 << 30)
    ids = [srv.submit(c, cfg, x, y, 60) for c, cfg, x, y in reqs]
    srv.drain()
    for rid, (c, cfg, x, y) in zip(ids, reqs):
        r = srv.poll(rid)
        assert r["status"] == 2
        ax, ay, ah = alone(ctx, c, cfg, x, y, 60, graph=True)
        assert np.array_equal(r["x"], ax) and np.array_equal(r["y"], ay), rid
        assert r["hpwl"] == ah
        assert r["t_submit"] <= r["t_start"] <= r["t_finish"]
        srv.release(rid)
    srv.close()


def test_admission_by_budget_and_rejection(ctx):
    print("case: server admission", flush=True)
    reqs = requests(4, fxn.netgen.SERVE_SEED_BASE + 100, sizes=(20000,))
    # a budget that fits one 20k-gate job at a time: requests run one after another
    srv = fxn.PlacementServer(ctx, max_active=8, hbm_budget=12 << 20)
    ids = [srv.submit(c, cfg, x, y, 20) for c, cfg, x, y in reqs]
    srv.drain()
    spans = sorted((srv.poll(i, False)["t_start"], srv.poll(i, False)["t_finish"]) for i in ids)
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0 - 1e-6, "jobs overlapped although the budget admits one"
    big, x, y = fxn.netgen.flat(400000, 440000, 7)
    with pytest.raises(fxn.FluxionError):
        srv.submit(big, fxn.PlacerConfig(density_grid=512, density_weight=weight_for(512)), x, y, 10)
    srv.close()


def test_overlap_speeds_up_small_requests(ctx):
    # many small requests: the served form overlaps them on the device
    print("case: server throughput", flush=True)
    reqs = requests(32, fxn.netgen.SERVE_SEED_BASE + 200, sizes=(1000, 2000))
    t0 = time.perf_counter()
    for c, cfg, x, y in reqs:
        alone(ctx, c, cfg, x, y, 100, graph=False)
    plain = time.perf_counter() - t0
    srv = fxn.PlacementServer(ctx, max_active=16, hbm_budget=8 << 30)
    t0 = time.perf_counter()
    ids = [srv.submit(c, cfg, x, y, 100) for c, cfg, x, y in reqs]
    srv.drain()
    served = time.perf_counter() - t0
    assert all(srv.poll(i, False)["status"] == 2 for i in ids)
    srv.close()
    print(f"32 small requests: plain {plain:.3f}s served {served:.3f}s speedup {plain / served:.2f}x",
          flush=True)
    assert served < plain
