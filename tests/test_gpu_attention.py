"""Attention scoring (compute_all_attention, attention.rs:33-99) on the GPU vs
the oracle's literal restatement: identical FP32 scores and top-k lists on
every level."""
from types import SimpleNamespace

import numpy as np
import pytest

import fluxion as fxn
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def tree_of(start, count, depth, order):
    start, count, depth = (np.asarray(a, np.int32) for a in (start, count, depth))
    return SimpleNamespace(order=np.asarray(order, np.int32), node_start=start, node_count=count,
                           node_depth=depth, num_nodes=len(start))


def check_levels(amap, circ, tree, k):
    depths = tree.node_depth
    want_levels = [d for d in range(int(depths.max()) + 1) if (depths == d).sum() >= 2]
    assert [lv.level_index for lv in amap.levels] == want_levels
    for lv in amap.levels:
        ids = np.flatnonzero(depths == lv.level_index)
        assert np.array_equal(lv.clusters, ids)
        scores, top = orc.attention_level(circ.num_gates, circ.net_ptr, circ.net_pins,
                                          tree.node_start[ids], tree.node_count[ids], tree.order, k)
        assert np.array_equal(lv.dense_scores(), scores), lv.level_index
        assert lv.top_k == [[int(j) for j in r if j >= 0] for r in top], lv.level_index
        sh = lv.shared
        assert (sh != sh.T).nnz == 0


def test_attention_reference_case(ctx):
    print("case: attention reference case", flush=True)
    # attention.rs:122-143 (test_attention_basic)
    circ = fxn.Circuit.from_nets(6, [[0, 1], [0, 3], [3, 4]], 100.0, 100.0)
    tree = tree_of([0, 0, 3], [6, 3, 3], [0, 1, 1], np.arange(6))
    amap = fxn.compute_all_attention(ctx, tree, circ, fxn.AttentionConfig(top_k=5))
    assert len(amap.levels) == 1 and amap.levels[0].level_index == 1
    assert np.array_equal(amap.levels[0].dense_scores(), np.array([[0, 0.5], [0.5, 0]], np.float32))
    assert amap.levels[0].top_k == [[1], [0]]
    check_levels(amap, circ, tree, 5)


def test_attention_kat(ctx):
    print("case: attention KAT", flush=True)
    nets = [[0, 3], [1, 6], [2], [4, 5], [7, 8]]
    circ = fxn.Circuit.from_nets(10, nets)
    tree = tree_of([0, 0, 3, 6, 9], [10, 3, 3, 3, 1], [0, 1, 1, 1, 1], np.arange(10))
    for k in (0, 1, 3, 10):
        amap = fxn.compute_all_attention(ctx, tree, circ, fxn.AttentionConfig(top_k=k))
        check_levels(amap, circ, tree, k)
    lv = amap.levels[0]
    assert lv.totals.tolist() == [3, 2, 2, 0]
    assert lv.top_k == [[1, 2], [0], [0], []]


@pytest.mark.parametrize("n", [2000, 20000])
def test_attention_matches_oracle(ctx, n):
    print(f"case: attention n={n}", flush=True)
    circ, _, _ = fxn.netgen.flat(n, n * 11 // 10, 700 + n)
    tree = fxn.build_hierarchy(ctx, circ, exact=False)
    amap = fxn.compute_all_attention(ctx, tree, circ, fxn.AttentionConfig(top_k=10))
    check_levels(amap, circ, tree, 10)
    print(f"n={n}: {len(amap.levels)} levels, {amap.build_ms:.2f} ms on the GPU", flush=True)


def test_attention_200k_invariants(ctx):
    print("case: attention n=200000 invariants", flush=True)
    circ, _, _ = fxn.netgen.flat(200000, 220000, 902)
    tree = fxn.build_hierarchy(ctx, circ, exact=False)
    amap = fxn.compute_all_attention(ctx, tree, circ)
    pin_net = np.repeat(np.arange(circ.num_nets), np.diff(circ.net_ptr))
    pos = np.empty(circ.num_gates, np.int64)
    pos[tree.order] = np.arange(circ.num_gates)
    for lv in amap.levels[-3:]:
        ids = lv.clusters
        # cluster of every gate at this depth, then distinct (net, cluster) pairs
        edges = np.append(tree.node_start[ids], circ.num_gates)
        pp = pos[circ.net_pins]
        cl = np.searchsorted(edges, pp, side="right") - 1
        ends = tree.node_start[ids] + tree.node_count[ids]
        inside = (cl >= 0) & (cl < len(ids)) & (pp < ends[np.clip(cl, 0, len(ids) - 1)])
        pairs = np.unique(np.stack([pin_net[inside], cl[inside]], 1), axis=0)
        totals = np.bincount(pairs[:, 1], minlength=len(ids))
        assert np.array_equal(lv.totals, totals)
        assert (lv.shared != lv.shared.T).nnz == 0
        rows = lv.scores.tocsr()
        for i in range(0, len(ids), max(1, len(ids) // 64)):
            r = rows.getrow(i)
            order = sorted(zip(-r.data, r.indices))[:10]
            assert lv.top_k[i] == [int(j) for _, j in order]
    print(f"n=200000: {len(amap.levels)} levels, {amap.build_ms:.2f} ms on the GPU", flush=True)
