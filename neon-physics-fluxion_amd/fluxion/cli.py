"""Command line of the pipeline (HAPR_fluxion_v4/python/cli.py:5-16)."""
from __future__ import annotations

import argparse

from .engine import FluxionEngine


def main(argv=None):
    parser = argparse.ArgumentParser(description="FLUXION HAPR placement on MI355X")
    parser.add_argument("--config", default="config/default.yaml", help="Config file")
    parser.add_argument("--input", required=True, help="Input circuit (BLIF)")
    parser.add_argument("--output", default="placed.def", help="Output DEF")
    args = parser.parse_args(argv)
    engine = FluxionEngine(args.config)
    engine.place(args.input, args.output)
    print(f"Done! HPWL: {engine.stats['hpwl']:.1f}")


if __name__ == "__main__":
    main()
