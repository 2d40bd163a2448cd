#!/usr/bin/env python3
"""Placement iterations/s of the flat global placement loop on one MI355X.

Workload (default): 200,000 unit gates / 220,000 nets on a 1000x1000 die
(net degree 2: 60%, 3-5: 30%, 6-20: 10%; 80% of a net's pins from one
64-gate block), uniform random initial positions, Barnes-Hut repulsion
theta=0.5, 512x512 DCT Poisson density grid (split-BF16 MFMA), FP32 forces,
FP64 HPWL, and a 256x256 int32 congestion grid at the end of the run.

One step = one placement iteration: net springs + Barnes-Hut (keys, sort,
tree, walk) + DCT density + integrator. The timed region runs K steps on the
device-resident job and then the final congestion grid, routing overflow and
FP64 HPWL; positions and netlist are in HBM before it starts.

Prints ONE JSON line (rank 0). --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neon-physics-fluxion_amd"))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_MFMA_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3

WORKLOADS = {
    # name: (gates, nets, seed offset, density grid, congestion grid, iterations)
    "flat_10k": (10000, 11000, 1, 0, 64, 200),
    "flat_200k": (200000, 220000, 2, 512, 256, 500),
    "flat_2m": (2000000, 2200000, 5, 1024, 1024, 50),
}


def phase_roofline(phase, N, E, P, G, stats, avg_ms):
    """Algorithmic work of one launch of a phase -> roofline dict."""
    s = avg_ms / 1e3
    if phase == "bh_walk":
        flops = 17 * stats["interactions"] + 9 * stats["opened"]
        ach = flops / s / 1e12
        return {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_FP32_TFLOPS, "algorithmic": f"17*I+9*V = {flops:.4g} flop"}
    if phase == "dct":
        flops = 8 * G ** 3
        ach = flops / s / 1e12
        return {"bound": "mfma", "achieved": ach, "peak": PEAK_MFMA_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_MFMA_BF16_TFLOPS, "algorithmic": f"8*G^3 = {flops:.4g} flop"}
    bytes_per = {
        "springs": 4 * (E + 1) + 12 * P + 16 * E,
        "bh_keys": 8 * N + 12 * N,
        "bh_sort": 8 * 36 * N,
        "bh_tree": 150 * N,
        "combine": 40 * N + 20 * P + 16 * N,
    }[phase]
    ach = bytes_per / s / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": ach / PEAK_HBM_GBS, "algorithmic": f"{bytes_per:.4g} B"}


def grid_for(n):
    return 128 if n < 20000 else 256 if n < 100000 else 512 if n < 300000 else 1024


def serving(fxn, ctx, num_requests, iters, seed_rng=7):
    """Placement serving stream: log-uniform 1k-500k gates, grid 128-1024 by
    size. Plain = one request at a time, eager launches; served = the
    continuous-batching server (burst for capacity, then Poisson arrivals at
    0.7x capacity for latency)."""
    import numpy as np
    rng = np.random.default_rng(seed_rng)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(500000), num_requests)).astype(int)
    reqs = []
    for r, n in enumerate(sizes):
        circ, x, y = fxn.netgen.flat(int(n), int(n) * 11 // 10, fxn.netgen.SERVE_SEED_BASE + r)
        G = grid_for(n)
        reqs.append((circ, fxn.PlacerConfig(density_grid=G, density_weight=fxn.netgen.density_weight_for(G)),
                     x, y))
    out = {"requests": num_requests, "iterations_per_request": iters,
           "gates_min": int(sizes.min()), "gates_max": int(sizes.max()),
           "gates_mean": float(sizes.mean())}
    t0 = time.perf_counter()
    for circ, cfg, x, y in reqs:
        job = fxn.FlatJob(ctx, circ, cfg)
        job.set_positions(x, y)
        job.run(iters)
        job.hpwl()
        job.close()
    plain = time.perf_counter() - t0
    out["plain_placements_per_s"] = num_requests / plain
    srv = fxn.PlacementServer(ctx, max_active=16, hbm_budget=64 << 30)
    t0 = time.perf_counter()
    ids = [srv.submit(c, cfg, x, y, iters) for c, cfg, x, y in reqs]
    srv.drain()
    burst = time.perf_counter() - t0
    for i in ids:
        srv.release(i)
    cap = num_requests / burst
    out["served_burst_placements_per_s"] = cap
    rate = 0.7 * cap
    gaps = rng.exponential(1.0 / rate, num_requests)
    c0 = srv.clock()
    ids = []
    t_next = time.perf_counter()
    for (c, cfg, x, y), gap in zip(reqs, gaps):
        t_next += gap
        while time.perf_counter() < t_next:
            time.sleep(min(0.001, max(0.0, t_next - time.perf_counter())))
        ids.append(srv.submit(c, cfg, x, y, iters))
    srv.drain()
    lat, fin = [], []
    for i in ids:
        r = srv.poll(i, False)
        lat.append(r["t_finish"] - r["t_submit"])
        fin.append(r["t_finish"])
    srv.close()
    span = max(fin) - c0
    out.update({"poisson_rate_per_s": rate, "served_placements_per_s": num_requests / span,
                "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                "served_over_plain": (num_requests / span) / out["plain_placements_per_s"],
                "burst_over_plain": cap / out["plain_placements_per_s"]})
    return out


def cpu_baseline(circ, x, y, G, dw, seconds):
    """FP32 CPU path of record (the oracle) on a bounded sample of iterations."""
    from oracle import oracle as orc
    n = circ.num_gates
    t0 = time.perf_counter()
    its = 0
    xx, yy, vx, vy = x, y, None, None
    while True:
        xx, yy, vx, vy, *_ = orc.flat_run(n, circ.net_ptr, circ.net_pins, 1000.0, 1000.0, xx, yy,
                                          vx, vy, iters=1, G=G, dens_weight=dw,
                                          nthreads=1)
        its += 1
        el = time.perf_counter() - t0
        if el >= seconds or its >= 50:
            break
    return {"value": its / el, "unit": "placement iters/s", "cores": 1, "kind": "port",
            "sample": f"{its} iterations of the FP32 CPU path of record (oracle/cpu_ref.cpp: "
                      f"springs + Barnes-Hut + DCT G={G} + integrator) at {n} gates, 1 thread, "
                      f"{el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="flat_200k", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--density-weight", type=float, default=None,
                    help="default: fluxion.netgen.density_weight_for(grid)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests in the serving measurement (0 skips it)")
    ap.add_argument("--serve-iters", type=int, default=100)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (HBM counter calibration)")
    ap.add_argument("--fast-math", action="store_true",
                    help="Barnes-Hut walk with rsq/FMA instead of IEEE div/sqrt")
    ap.add_argument("--graph", action="store_true",
                    help="replay a captured hipGraph per step (no per-kernel events)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        a = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy}))
        return

    import numpy as np

    import fluxion as fxn

    N, E, seed_off, G, C, default_steps = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else default_steps
    W = args.warmup
    circ, x0, y0 = fxn.netgen.flat(N, E, fxn.netgen.SEED_BASE + seed_off)
    P = int(circ.net_ptr[-1])
    if args.density_weight is None:
        args.density_weight = fxn.netgen.density_weight_for(G)
    ctx = fxn.GpuContext.new(0)
    cfg = fxn.PlacerConfig(density_grid=G, density_weight=args.density_weight,
                           fast_math=args.fast_math)
    job = fxn.FlatJob(ctx, circ, cfg)
    job.set_positions(x0, y0)

    # Warm-up, then a short fully instrumented pass to find the dominant phase.
    job.run(W)
    job.congestion(C, C)  # allocates the grid buffers outside the timed region
    job.set_timing(fxn.FlatJob.PHASES)
    job.run(max(5, min(20, K)))
    pre = job.read_timing()
    dominant = max(pre, key=lambda k: pre[k][0])
    job.set_timing(())
    if args.graph:
        job.use_graph(True)
        job.run(1)  # captures the iteration graph
    else:
        job.set_timing([dominant])
    # The timed run starts from the generated placement: its final HPWL is the
    # state after exactly K iterations (compared with the oracle's run).
    job.set_positions(x0, y0)
    job.set_velocities(np.zeros(N, np.float32), np.zeros(N, np.float32))
    job.sync()
    t0 = time.perf_counter()
    job.run_async(K)
    counts = job.congestion(C, C)
    tot_of, max_of = job.overflow(C, C)
    hpwl = job.hpwl()
    job.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    dom = job.read_timing().get(dominant) if not args.graph else None
    stats = job.bh_stats()

    value = K / elapsed
    out = {
        "metric": "placement iters/s @200k gates",
        "value": value,
        "unit": "placement iters/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded netlist (no dataset)",
        "config": {"workload": args.workload, "gates": N, "nets": E, "pins": P,
                   "density_grid": G, "density_weight": args.density_weight,
                   "congestion_grid": C, "theta": 0.5,
                   "dct": "split-bf16 MFMA (3-part split, 6 passes), fp32 accumulate",
                   "hpwl": "fp64", "execution": "graph" if args.graph else "eager+events",
                   "bh_walk": "fast (rsq/fma)" if args.fast_math else "exact (ieee div/sqrt)"},
        "final_hpwl": hpwl,
        "routing_overflow": {"total": tot_of, "max": max_of},
        "congestion_cells_nonzero": int((counts > 0).sum()),
        "phases_ms_per_iter": {k: v[0] / max(1, v[1]) for k, v in pre.items()},
        "bh": stats,
    }
    if dom:
        avg = dom[0] / dom[1]
        rf = phase_roofline(dominant, N, E, P, G, stats, avg)
        rf["kernel"] = dominant
        rf["avg_launch_ms"] = avg
        rf["traffic"] = None
        out["roofline"] = rf
    if args.serve_requests > 0:
        out["serving"] = serving(fxn, ctx, args.serve_requests, args.serve_iters)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(circ, x0, y0, G, args.density_weight, args.cpu_seconds)
        out["cpu_baseline"]["ratio_gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    job.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
