"""Congestion-aware refinement (refine_placement, refiner.rs:34-125) on the GPU
vs the oracle's literal restatement: identical positions, iteration count,
stop reason and FP32 overflow figures."""
import numpy as np
import pytest

import fluxion as fxn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_refine_hand_case(ctx):
    print("case: refine hand case", flush=True)
    nets = [[2 * k, 2 * k + 1] for k in range(6)]
    circ = fxn.Circuit.from_nets(12, nets, 100.0, 100.0)
    for iters, want_x, want_it, want_stop in ((1, 10.0, 1, "max_iterations"),
                                              (3, 10.0, 3, "max_iterations"),
                                              (100, 5.0, 10, "stalled")):
        pl = fxn.Placement(np.full(12, 5.0, np.float32), np.full(12, 5.0, np.float32))
        r = fxn.refine_placement(ctx, pl, circ, fxn.RefineConfig(max_iterations=iters, grid_cols=10,
                                                                 grid_rows=10))
        assert r["iterations"] == want_it and r["stop"] == want_stop
        assert np.all(pl.gate_x == want_x) and np.all(pl.gate_y == 5.0)
        assert r["total_overflow"] == 2.0 and r["max_overflow"] == 2.0


@pytest.mark.parametrize("N,C,flat_iters", [(2000, 32, 30), (20000, 64, 60), (20000, 128, 60)])
def test_refine_matches_oracle(ctx, N, C, flat_iters):
    print(f"case: refine N={N} C={C}", flush=True)
    circ, x0, y0 = fxn.netgen.flat(N, N * 11 // 10, fxn.netgen.SEED_BASE + 700 + N)
    # a converged-ish placement as input (the refiner runs after global placement)
    cfg = fxn.PlacerConfig(density_grid=128, density_weight=fxn.netgen.density_weight_for(128))
    with fxn.FlatJob(ctx, circ, cfg) as job:
        job.set_positions(x0, y0)
        job.run(flat_iters)
        x, y = job.positions()
    rc = fxn.RefineConfig(grid_cols=C, grid_rows=C)
    pl = fxn.Placement(x.copy(), y.copy())
    r = fxn.refine_placement(ctx, pl, circ, rc)
    ox, oy, it, stop, t, m = orc.refine(circ.net_ptr, circ.net_pins, 1000.0, 1000.0, x, y,
                                        cols=C, rows=C)
    print(f"refine N={N} C={C}: {r}, oracle it={it} stop={stop} t={t} m={m}", flush=True)
    assert r["iterations"] == it and r["stop"] == fxn.placer.REFINE_STOP[stop]
    assert r["total_overflow"] == t and r["max_overflow"] == m
    assert np.array_equal(pl.gate_x, ox) and np.array_equal(pl.gate_y, oy)
    assert r["hpwl"] == pytest.approx(orc.hpwl(circ.net_ptr, circ.net_pins, ox, oy), rel=1e-12)
