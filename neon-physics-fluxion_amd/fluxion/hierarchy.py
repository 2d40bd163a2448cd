"""Hierarchy of HAPR on the MI355X library.

Reference interface (HAPR_fluxion_v4):
  * PartitionConfig                 src/algorithm/partitioner.rs:13-33
  * build_hierarchy                 src/algorithm/partitioner.rs:36-50   -> build_hierarchy
  * ClusterNode                     src/data/hierarchy.rs:9-76
  * HierarchyTree                   src/data/hierarchy.rs:78-191
  * root region + place_level_recursive  src/algorithm/placer.rs:43-64, 111-141
                                    -> HierarchyTree.regions (fx_hierarchy_regions)

The bisection runs on the GPU (fx_build_hierarchy): all nodes of a depth are
bisected together. Node ids, depths and leaf gate orders are the reference's;
an internal node's gate_indices hold its gate set in leaf order (no reference
consumer reads that order: attention uses sets, attention.rs:59-61, and the
placer reads only leaves, placer.rs:66-69).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import FX_HIER_EXACT, FX_HIER_FP64, FxPartitionConfig, check, lib, ptr

__all__ = ["PartitionConfig", "ClusterNode", "HierarchyTree", "build_hierarchy", "hierarchy_size"]


@dataclass
class PartitionConfig:
    """PartitionConfig (partitioner.rs:13-33) / YAML partitioning.* (default.yaml:10-13)."""
    max_leaf_size: int = 64
    min_cluster_size: int = 8
    balance_factor: float = 0.4
    power_iterations: int = 30

    def to_c(self) -> FxPartitionConfig:
        return FxPartitionConfig(int(self.max_leaf_size), int(self.min_cluster_size),
                                 float(self.balance_factor), int(self.power_iterations))


@dataclass
class ClusterNode:
    """ClusterNode (hierarchy.rs:9-24)."""
    id: int
    depth: int
    gate_indices: np.ndarray
    children: list = field(default_factory=list)
    center_x: float = 0.0
    center_y: float = 0.0
    width: float = 0.0
    height: float = 0.0

    def is_leaf(self) -> bool:
        return not self.children

    def num_gates(self) -> int:
        return int(self.gate_indices.shape[0])

    def set_region(self, cx: float, cy: float, w: float, h: float) -> None:
        self.center_x, self.center_y, self.width, self.height = cx, cy, w, h


@dataclass
class HierarchyTree:
    """HierarchyTree (hierarchy.rs:78-191), backed by the flat node table the
    GPU build returns (pre-order ids)."""
    root: ClusterNode
    max_depth: int
    leaf_map: np.ndarray
    num_nodes: int
    order: np.ndarray
    node_start: np.ndarray
    node_count: np.ndarray
    node_depth: np.ndarray
    node_left: np.ndarray
    node_right: np.ndarray
    nodes: list = field(repr=False, default_factory=list)
    build_ms: float = 0.0

    def leaves(self) -> list:
        """Leaves in pre-order (collect_leaves, hierarchy.rs:176-190)."""
        return [n for n in self.nodes if n.is_leaf()]

    def clusters_at_depth(self, depth: int) -> list:
        return [n for n in self.nodes if n.depth == depth]

    def leaf_ids(self) -> np.ndarray:
        return np.flatnonzero(self.node_left < 0).astype(np.int32)

    def regions(self, chip_width: float, chip_height: float) -> np.ndarray:
        """[num_nodes, 4] center_x, center_y, width, height of every node as
        hapr_placement lays them out (placer.rs:43-64, 111-141); also stored
        on the nodes (set_region)."""
        reg = np.zeros((max(self.num_nodes, 1), 4), np.float32)
        check(lib.fx_hierarchy_regions(self.num_nodes, ptr(self.node_left), ptr(self.node_right),
                                       float(chip_width), float(chip_height), ptr(reg)))
        for n in self.nodes:
            n.set_region(*(float(v) for v in reg[n.id]))
        return reg[:self.num_nodes]

    def leaf_batch(self, chip_width: float, chip_height: float):
        """(leaf_ptr, leaf_gates, region) of the leaves in pre-order, the
        batch fx_place_leaves takes (placer.rs:66-69)."""
        reg = self.regions(chip_width, chip_height)
        ids = self.leaf_ids()
        counts = self.node_count[ids]
        leaf_ptr = np.zeros(ids.shape[0] + 1, np.int32)
        leaf_ptr[1:] = np.cumsum(counts)
        # Leaves in pre-order tile `order` from left to right.
        return leaf_ptr, self.order, np.ascontiguousarray(reg[ids])


def hierarchy_size(num_gates: int, config: PartitionConfig | None = None) -> tuple[int, int]:
    """(num_nodes, max_depth) of the hierarchy of num_gates gates."""
    cfg = (config or PartitionConfig()).to_c()
    nn, md = C.c_int32(), C.c_int32()
    check(lib.fx_hierarchy_size(int(num_gates), C.byref(cfg), C.byref(nn), C.byref(md)))
    return nn.value, md.value


def build_hierarchy(ctx, circuit, config: PartitionConfig | None = None,
                    exact: bool = True) -> HierarchyTree:
    """build_hierarchy(&Circuit, &PartitionConfig) (partitioner.rs:36-50) on the GPU.

    exact=True reproduces the reference's FP32 mean/norm arithmetic bit for
    bit; exact=False reduces them in FP64 (faster, a different rounding)."""
    config = config or PartitionConfig()
    cfg = config.to_c()
    n = int(circuit.num_gates)
    nn, md = C.c_int32(), C.c_int32()
    check(lib.fx_hierarchy_size(n, C.byref(cfg), C.byref(nn), C.byref(md)))
    cap = nn.value
    aptr, adj = circuit.build_adjacency()
    adj = adj if adj.size else np.zeros(1, np.int32)
    order = np.zeros(max(n, 1), np.int32)
    arrs = [np.zeros(max(cap, 1), np.int32) for _ in range(5)]
    got = C.c_int32()
    check(lib.fx_build_hierarchy(ctx.handle, n, ptr(aptr), ptr(adj), C.byref(cfg),
                                 FX_HIER_EXACT if exact else FX_HIER_FP64, ptr(order),
                                 *[ptr(a) for a in arrs], cap, C.byref(got)))
    start, count, depth, left, right = (a[:got.value] for a in arrs)
    order = order[:n]
    nodes = [ClusterNode(i, int(depth[i]), order[start[i]:start[i] + count[i]])
             for i in range(got.value)]
    for i in range(got.value):
        if left[i] >= 0:
            nodes[i].children = [nodes[left[i]], nodes[right[i]]]
    leaf_map = np.zeros(n, np.int64)
    for i in np.flatnonzero(left < 0):
        leaf_map[nodes[i].gate_indices] = i
    return HierarchyTree(root=nodes[0], max_depth=md.value, leaf_map=leaf_map,
                         num_nodes=got.value, order=order, node_start=start, node_count=count,
                         node_depth=depth, node_left=left, node_right=right, nodes=nodes,
                         build_ms=float(lib.fx_hierarchy_last_ms(ctx.handle)))
