"""Recursive spectral bisection (build_hierarchy, partitioner.rs:36-213) on the
GPU vs the oracle's restatement, and the whole pipeline of main.rs:27-121
(BLIF -> hierarchy -> hapr_placement -> refine_placement -> DEF) vs the
oracle's pipeline: identical node tables, leaf orders and final positions."""
import os

import numpy as np
import pytest

import fluxion as fxn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NODE_KEYS = ("start", "count", "depth", "left", "right")


def circuit_for(n):
    if n == 20:
        return fxn.Circuit.from_nets(20, [[i, i + 1] for i in range(19)])
    circ, _, _ = fxn.netgen.flat(n, n * 11 // 10, 900 + n)
    return circ


def check_tree(tree, order, nodes):
    assert tree.num_nodes == len(nodes["start"])
    assert np.array_equal(tree.order, order)
    for k, a in zip(NODE_KEYS, (tree.node_start, tree.node_count, tree.node_depth, tree.node_left,
                                tree.node_right)):
        assert np.array_equal(a, nodes[k]), k


# 20: the reference's chain test; 2000: one workgroup per node from the
# root; 4097: a depth mixing grid-path (2049) and workgroup-path (2048)
# nodes; 20000 / 200000: grid path then workgroup path; 300000: nodes of
# more than 1024 chunks (separate totals kernel in FP64 mode).
@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fp64"])
@pytest.mark.parametrize("n,max_leaf", [(20, 8), (2000, 64), (4097, 64), (20000, 64),
                                        (200000, 64), (300000, 64)])
def test_hierarchy_matches_oracle(ctx, n, max_leaf, exact):
    print(f"case: hierarchy n={n} exact={exact}", flush=True)
    circ = circuit_for(n)
    aptr, adj = circ.build_adjacency()
    cfg = fxn.PartitionConfig(max_leaf_size=max_leaf)
    tree = fxn.build_hierarchy(ctx, circ, cfg, exact=exact)
    order, nodes = orc.build_hierarchy(n, aptr, adj, max_leaf=max_leaf, mode=0 if exact else 2)
    check_tree(tree, order, nodes)
    leaves = tree.leaves()
    assert sum(lf.num_gates() for lf in leaves) == n
    assert np.array_equal(np.sort(tree.order), np.arange(n))
    print(f"n={n} exact={exact}: {tree.num_nodes} nodes, depth {tree.max_depth}, "
          f"{tree.build_ms:.2f} ms on the GPU", flush=True)


def test_hierarchy_edge_cases(ctx):
    print("case: hierarchy edge cases", flush=True)
    # no nets: L = 0, the norm test fails and the stable sort keeps the order
    for n in (1, 2, 100, 3000):
        circ = fxn.Circuit.from_nets(n, [])
        for exact in (True, False):
            tree = fxn.build_hierarchy(ctx, circ, fxn.PartitionConfig(max_leaf_size=4), exact=exact)
            order, nodes = orc.build_hierarchy(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32),
                                               max_leaf=4, mode=0 if exact else 2)
            check_tree(tree, order, nodes)
    # self-loops and repeated pins (a BLIF gate reading its own output)
    nets = [[0, 0, 1], [1, 2, 2, 3], [3, 4], [4, 0], [5, 5]] + [[i, i + 1] for i in range(5, 40)]
    circ = fxn.Circuit.from_nets(41, nets)
    aptr, adj = circ.build_adjacency()
    for exact in (True, False):
        tree = fxn.build_hierarchy(ctx, circ, fxn.PartitionConfig(max_leaf_size=3, min_cluster_size=0),
                                   exact=exact)
        order, nodes = orc.build_hierarchy(41, aptr, adj, max_leaf=3, min_cluster=0,
                                           mode=0 if exact else 2)
        check_tree(tree, order, nodes)
    # zero power iterations: the split follows the initial vector
    circ = circuit_for(2000)
    aptr, adj = circ.build_adjacency()
    tree = fxn.build_hierarchy(ctx, circ, fxn.PartitionConfig(power_iterations=0))
    order, nodes = orc.build_hierarchy(2000, aptr, adj, iters=0)
    check_tree(tree, order, nodes)


def write_blif(path, circ):
    """A BLIF whose nets are circ's: pin 0 of net e drives e<e>, the other
    pins read it."""
    drives = [[] for _ in range(circ.num_gates)]
    reads = [[] for _ in range(circ.num_gates)]
    for e in range(circ.num_nets):
        pins = circ.net_pins[circ.net_ptr[e]:circ.net_ptr[e + 1]]
        drives[pins[0]].append(e)
        for g in pins[1:]:
            reads[g].append(e)
    with open(path, "w") as f:
        f.write(".model synth\n")
        for g in range(circ.num_gates):
            toks = [f"I{k}=e{e}" for k, e in enumerate(reads[g])] + [f"O=e{e}" for e in drives[g]]
            f.write(".gate CELL " + " ".join(toks) + "\n")
        f.write(".end\n")


def oracle_pipeline(circ, eng):
    pc, pl, rc = eng.partition_config(), eng.placer_config(), eng.refine_config()
    n = circ.num_gates
    aptr, adj = orc.build_adjacency(n, circ.net_ptr, circ.net_pins)
    order, nodes = orc.build_hierarchy(n, aptr, adj, pc.max_leaf_size, pc.min_cluster_size,
                                       pc.balance_factor, pc.power_iterations, mode=0)
    reg = orc.hapr_regions(nodes, circ.chip_width, circ.chip_height)
    leaves = np.flatnonzero(nodes["left"] < 0)
    leaf_ptr = np.zeros(len(leaves) + 1, np.int32)
    leaf_ptr[1:] = np.cumsum(nodes["count"][leaves])
    x, y, *_ = orc.place_leaves(leaf_ptr, order, reg[leaves], n, aptr, adj, pl.attract_weight,
                                pl.repel_weight, pl.damping, pl.dt, pl.force_iterations)
    x, y, it, st, tot, mx = orc.refine(circ.net_ptr, circ.net_pins, circ.chip_width,
                                       circ.chip_height, x, y, rc.max_iterations, rc.move_distance,
                                       rc.overflow_threshold, rc.grid_cols, rc.grid_rows)
    return x, y, it


def def_positions(path):
    xs, ys = [], []
    for line in open(path):
        if "+ PLACED" in line:
            t = line.split()
            xs.append(int(t[6]))
            ys.append(int(t[7]))
    return np.array(xs, np.int64), np.array(ys, np.int64)


@pytest.mark.parametrize("which", ["small", "synth2000", "synth20000"])
def test_engine_pipeline_matches_oracle(ctx, tmp_path, which):
    print(f"case: engine pipeline {which}", flush=True)
    if which == "small":
        blif = os.path.join(GOLDEN, "small.blif")
    else:
        n = int(which[5:])
        blif = str(tmp_path / f"{which}.blif")
        write_blif(blif, circuit_for(n))
    eng = fxn.FluxionEngine(os.path.join(GOLDEN, "default.yaml"), ctx=ctx)
    out = str(tmp_path / "placed.def")
    assert eng.place(blif, out) == out
    circ = fxn.parse_blif(blif)
    _, _, placement, summary = eng.run(blif)
    x, y, it = oracle_pipeline(circ, eng)
    assert np.array_equal(placement.gate_x, x) and np.array_equal(placement.gate_y, y)
    assert summary["iterations"] == it
    # DEF: the reference's integration checks (test_integration.py:30-38) and
    # the component coordinates (f32 product truncated toward zero).
    text = open(out).read()
    assert "COMPONENTS" in text and "END COMPONENTS" in text and "+ PLACED" in text
    dx, dy = def_positions(out)
    assert np.array_equal(dx, [int(v) for v in x * np.float32(1000)])
    assert np.array_equal(dy, [int(v) for v in y * np.float32(1000)])
    print(f"{which}: {eng.stats}", flush=True)
