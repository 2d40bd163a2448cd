"""Attention scoring of HAPR on the MI355X library.

Reference interface (HAPR_fluxion_v4/src/algorithm/attention.rs):
  * AttentionConfig                   attention.rs:27-35
  * AttentionLevel, AttentionMap      attention.rs:13-25
  * compute_all_attention             attention.rs:33-48  -> compute_all_attention
  * compute_level_attention           attention.rs:50-88  (every level in one GPU call)

Scores are kept sparse: AttentionLevel.scores is a scipy CSR matrix of the
non-zero entries of the reference's dense n x n matrix (dense_scores() gives
the dense one); top_k is the reference's list of lists.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib, ptr

__all__ = ["AttentionConfig", "AttentionLevel", "AttentionMap", "compute_all_attention"]


@dataclass
class AttentionConfig:
    """AttentionConfig (attention.rs:27-35; YAML attention.top_k)."""
    top_k: int = 10


@dataclass
class AttentionLevel:
    """AttentionLevel (attention.rs:13-18) plus the integer counts behind the scores."""
    level_index: int
    scores: object  # scipy.sparse.csr_matrix, float32, n x n
    top_k: list
    clusters: np.ndarray  # hierarchy node id of each cluster (pre-order)
    totals: np.ndarray    # nets touching each cluster
    shared: object        # scipy.sparse.csr_matrix, int32, off-diagonal shared-net counts

    def dense_scores(self) -> np.ndarray:
        return self.scores.toarray()


@dataclass
class AttentionMap:
    """AttentionMap (attention.rs:20-25)."""
    levels: list = field(default_factory=list)
    build_ms: float = 0.0


def compute_all_attention(ctx, tree, circuit, config: AttentionConfig | None = None) -> AttentionMap:
    """compute_all_attention(&HierarchyTree, &Circuit, &AttentionConfig) (attention.rs:33-48)."""
    import scipy.sparse as sp
    k = int((config or AttentionConfig()).top_k)
    net = circuit.netlist()
    order = np.ascontiguousarray(tree.order, np.int32)
    start = np.ascontiguousarray(tree.node_start, np.int32)
    count = np.ascontiguousarray(tree.node_count, np.int32)
    depth = np.ascontiguousarray(tree.node_depth, np.int32)
    h = C.c_void_p()
    check(lib.fx_attention_build(ctx.handle, C.byref(net), int(tree.num_nodes), ptr(order),
                                 ptr(start), ptr(count), ptr(depth), k, C.byref(h)))
    out = AttentionMap(build_ms=float(lib.fx_attention_last_ms(ctx.handle)))
    try:
        for lv in range(lib.fx_attention_num_levels(h)):
            li, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
            check(lib.fx_attention_level_info(h, lv, C.byref(li), C.byref(n), C.byref(nnz)))
            n_, z = n.value, nnz.value
            nodes = np.zeros(n_, np.int32)
            totals = np.zeros(n_, np.int32)
            rp = np.zeros(n_ + 1, np.int64)
            col = np.zeros(max(z, 1), np.int32)
            shr = np.zeros(max(z, 1), np.int32)
            sc = np.zeros(max(z, 1), np.float32)
            top = np.full(max(n_ * k, 1), -1, np.int32)
            check(lib.fx_attention_level_get(h, lv, ptr(nodes), ptr(totals), ptr(rp), ptr(col),
                                             ptr(shr), ptr(sc), ptr(top)))
            scores = sp.csr_matrix((sc[:z], col[:z], rp), shape=(n_, n_))
            shared = sp.csr_matrix((shr[:z], col[:z], rp), shape=(n_, n_))
            tops = top[:n_ * k].reshape(n_, k) if k else np.zeros((n_, 0), np.int32)
            out.levels.append(AttentionLevel(li.value, scores, [[int(j) for j in r if j >= 0]
                                                                for r in tops],
                                             nodes, totals, shared))
    finally:
        lib.fx_attention_free(h)
    return out
