#!/usr/bin/env python3
"""Writes a synthetic series of compact-field frames (field00000.bin, ...) for the
frame-loop driver kpt_run, in the format of include/AthenaFieldReader.hpp:274-288.

The field is the island-chain Harris sheet of workloads.island_field: Bx =
tanh(y/w), an island chain in By whose phases drift from frame to frame, Bz,
and an inflow Vy with 1% Gaussian noise. `--grid stretched` puts the y faces on
the tanh-stretched grid (spacing ratio `--ratio` toward y = 0), whose files
kpt_run reads with --allow-stretched-fields. `--distinct K` writes K distinct
frames and links frame f to frame f mod K, for long series on little disk.

    python make_fields.py OUT_DIR --frames 201 --nx 1024 --ny 512 --grid stretched
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import kpt  # noqa: E402
import workloads as W  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("out_dir")
    ap.add_argument("--frames", type=int, required=True, help="number of files (frames 0 .. frames-1)")
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--grid", choices=["stretched", "uniform"], default="stretched")
    ap.add_argument("--ratio", type=float, default=10.0, help="stretched grid: largest over smallest spacing")
    ap.add_argument("--width", type=float, default=0.02, help="current-sheet half width")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--distinct", type=int, default=0, help="write this many frames and link the rest (0: all)")
    args = ap.parse_args(argv)
    if args.frames < 1 or args.distinct < 0:
        ap.error("--frames must be positive and --distinct non-negative")
    os.makedirs(args.out_dir, exist_ok=True)
    y_edges = W.stretched_edges(args.ny, ratio=args.ratio) if args.grid == "stretched" else None
    distinct = args.distinct or args.frames
    for f in range(args.frames):
        path = os.path.join(args.out_dir, f"field{f:05d}.bin")
        if os.path.lexists(path):
            os.remove(path)
        if f < distinct:
            xe, ye, planes = W.island_field(args.nx, args.ny, width=args.width, seed=args.seed, y_edges=y_edges,
                                            frame=f)
            kpt.compact_field_write(path, xe, ye, [np.ascontiguousarray(p) for p in planes])
        else:
            os.symlink(f"field{f % distinct:05d}.bin", path)
    print(f"wrote {min(distinct, args.frames)} frames, {args.frames} files, to {args.out_dir}")


if __name__ == "__main__":
    main()
