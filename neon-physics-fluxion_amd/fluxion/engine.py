"""FluxionEngine (HAPR_fluxion_v4/python/fluxion/engine.py:10-48).

The reference's engine shells out to the Rust CLI, which runs main.rs:27-121:
parse -> build_hierarchy -> attention -> hapr_placement -> refine_placement ->
write_def. Here the same pipeline runs in-process on the GPU library:
BLIF and DEF in the library's host code, the bisection, the leaf placement
and the refinement on the GPU. Attention scoring (main.rs:88-93) runs only
when asked for (attention=True): its only consumer is the depth-0 spectral
branch of hapr_placement, which never runs (the root is the only cluster of
depth 0, attention.rs:40-43), so the placement is the same either way.
"""
from __future__ import annotations

import time
from pathlib import Path

import yaml

from .attention import AttentionConfig, compute_all_attention
from .hierarchy import PartitionConfig, build_hierarchy
from .parser import parse_blif, write_def
from .placer import GpuContext, PlacerConfig, RefineConfig, hapr_placement, refine_placement

__all__ = ["FluxionEngine"]


def _section(cfg: dict, name: str) -> dict:
    sec = cfg.get(name) or {}
    if not isinstance(sec, dict):
        raise ValueError(f"config section {name!r} must be a mapping")
    return sec


class FluxionEngine:
    """
    Usage:
        engine = FluxionEngine("config/default.yaml")
        engine.place("circuit.blif", "placed.def")
    """

    def __init__(self, config_path="default.yaml", ctx: GpuContext | None = None,
                 exact: bool = True, attention: bool = False):
        self.config_path = Path(config_path)
        if not self.config_path.exists():
            raise FileNotFoundError(f"Config not found: {self.config_path}")
        with open(self.config_path, "r") as f:
            self.config = yaml.safe_load(f) or {}
        self._ctx = ctx
        self.exact = exact
        self.attention = attention  # main.rs:88-93; no effect on the placement (Q3)
        self.attention_map = None
        self.stats: dict = {}

    # Config keys of config/default.yaml (config.rs:7-72); missing keys take
    # the reference's struct defaults.
    def partition_config(self) -> PartitionConfig:
        p = _section(self.config, "partitioning")
        return PartitionConfig(int(p.get("max_leaf_size", 64)), int(p.get("min_cluster_size", 8)),
                               float(p.get("balance_factor", 0.4)), 30)

    def placer_config(self) -> PlacerConfig:
        p = _section(self.config, "placement")
        return PlacerConfig(force_iterations=int(p.get("force_iterations", 50)),
                            attract_weight=float(p.get("attract_weight", 1.0)),
                            repel_weight=float(p.get("repel_weight", 0.5)),
                            damping=float(p.get("damping", 0.9)), dt=float(p.get("dt", 0.1)))

    def refine_config(self) -> RefineConfig:
        c = _section(self.config, "congestion")
        return RefineConfig(int(c.get("refinement_iterations", 100)),
                            float(c.get("move_distance", 5.0)),
                            float(c.get("overflow_threshold", 1.0)), int(c.get("grid_cols", 256)),
                            int(c.get("grid_rows", 256)))

    def _context(self) -> GpuContext:
        if self._ctx is None:
            self._ctx = GpuContext.new(0)
        return self._ctx

    def run(self, input_file: str):
        """parse -> hierarchy -> placement -> refinement; returns
        (circuit, tree, placement, refine summary)."""
        input_path = Path(input_file)
        if not input_path.exists():
            raise FileNotFoundError(f"Input file not found: {input_path}")
        fmt = str(self.config.get("format", "auto"))
        if input_path.suffix != ".blif" and fmt != "blif":
            raise RuntimeError("Unsupported format (only BLIF supported in this demo)")
        t0 = time.perf_counter()
        circuit = parse_blif(input_path)
        if circuit.num_gates == 0:
            raise RuntimeError("Circuit has 0 gates. Exiting.")
        circuit.build_adjacency()
        t1 = time.perf_counter()
        ctx = self._context()
        tree = build_hierarchy(ctx, circuit, self.partition_config(), exact=self.exact)
        ta = time.perf_counter()
        if self.attention:
            a = _section(self.config, "attention")
            self.attention_map = compute_all_attention(
                ctx, tree, circuit, AttentionConfig(int(a.get("top_k", 10))))
        t2 = time.perf_counter()
        placement = hapr_placement(ctx, circuit, tree, None, self.placer_config(),
                                   fast=not self.exact)
        t3 = time.perf_counter()
        summary = refine_placement(ctx, placement, circuit, self.refine_config())
        placement.update_wirelength(circuit)
        t4 = time.perf_counter()
        self.stats = {"gates": circuit.num_gates, "nets": circuit.num_nets,
                      "leaves": int((tree.node_left < 0).sum()), "max_depth": tree.max_depth,
                      "parse_s": t1 - t0, "hierarchy_s": ta - t1, "attention_s": t2 - ta,
                      "hierarchy_gpu_ms": tree.build_ms, "placement_s": t3 - t2,
                      "refine_s": t4 - t3, "hpwl": placement.total_wirelength,
                      "refine": summary}
        return circuit, tree, placement, summary

    def place(self, input_file: str, output_file: str = "placed.def"):
        """Run the full HAPR pipeline and write the DEF; returns output_file."""
        circuit, _, placement, _ = self.run(input_file)
        t0 = time.perf_counter()
        write_def(output_file, circuit, placement)
        self.stats["def_s"] = time.perf_counter() - t0
        return output_file
