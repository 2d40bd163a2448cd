"""Seeded synthetic workloads (generators in csrc/netgen.cpp).

Workloads (name -> sizes) used by the tests, the benchmark and serving:
  flat_10k   : 10,000 gates / 11,000 nets on a 1000x1000 die, 64x64 congestion grid
  flat_200k  : 200,000 gates / 220,000 nets, 512x512 DCT density grid, 256x256 congestion grid
  flat_2m    : 2,000,000 gates / 2,200,000 nets, 1024x1024 DCT grid and congestion grid
  leaves_4096: 4,096 leaf clusters of 16-256 gates (log-uniform), degree 2-8, 100x100 regions
Seeds: workload k uses 0xF1A00000 + k; serving request r uses 0xF1A10000 + r.
"""
from __future__ import annotations

import numpy as np

from ._lib import lib, ptr
from .placer import Circuit

SEED_BASE = 0xF1A00000
SERVE_SEED_BASE = 0xF1A10000


def flat(num_gates: int, num_nets: int, seed: int, die: float = 1000.0):
    """Returns (Circuit, x0, y0)."""
    p = lib.fx_netgen_flat(num_gates, num_nets, seed, die, die, None, None, 0, None, None)
    if p < 0:
        raise ValueError("bad netlist sizes")
    net_ptr = np.zeros(num_nets + 1, np.int32)
    pins = np.zeros(max(1, p), np.int32)
    x = np.zeros(num_gates, np.float32)
    y = np.zeros(num_gates, np.float32)
    q = lib.fx_netgen_flat(num_gates, num_nets, seed, die, die, ptr(net_ptr), ptr(pins),
                           int(pins.shape[0]), ptr(x), ptr(y))
    assert q == p
    return Circuit(num_gates, net_ptr, pins[:p], die, die), x, y


def leaves(num_leaves: int, seed: int, min_n: int = 16, max_n: int = 256, region: float = 100.0):
    """Returns (Circuit, leaf_ptr, leaf_gates, region[L,4]); leaves own consecutive gate ids."""
    counts = np.zeros(3, np.int64)
    lib.fx_netgen_leaves(num_leaves, seed, min_n, max_n, region, None, None, None, None, ptr(counts))
    n, e, p = (int(v) for v in counts)
    leaf_ptr = np.zeros(num_leaves + 1, np.int32)
    reg = np.zeros((num_leaves, 4), np.float32)
    net_ptr = np.zeros(e + 1, np.int32)
    pins = np.zeros(max(1, p), np.int32)
    lib.fx_netgen_leaves(num_leaves, seed, min_n, max_n, region, ptr(leaf_ptr), ptr(reg),
                         ptr(net_ptr), ptr(pins), ptr(counts))
    side = int(np.ceil(np.sqrt(num_leaves))) * region
    circ = Circuit(n, net_ptr, pins[:p], side, side)
    return circ, leaf_ptr, np.arange(n, dtype=np.int32), reg


def density_weight_for(G: int) -> float:
    """Density-force weight of the flat workloads for a DCT grid of side G.

    The reference's eigenvalue (w cos(pi k/n) - w)^2 grows as k^4, so the
    low modes of the density potential are amplified ~1e10 at G=512 and the
    force scales ~G^4. Measured on the 200k-gate workload: weight 1.0 sends
    every gate to a die corner (CPU path too), 1e-4 piles ~90% of the gates
    against a die edge, 3e-6 keeps all gates interior with the density force
    ~13% of the springs. The weight is 3e-6 at G=512, scaled by (512/G)^4.
    """
    return 3e-6 * (512.0 / G) ** 4 if G else 0.0
