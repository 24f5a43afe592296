#!/usr/bin/env python
"""Adaptive band widths (Eq. 5) of trained level stacks, drop-in for the
reference's appendix_experiments/extract_deltas.py.

For every shape with coarse and medium checkpoints under the i3d layout

    <root>/results/<shape>/coarse/best.pth, <root>/results/<shape>/medium/best.pth
    <root>/data/normalized_sphere/input/<shape>.ply   (the input surface points)

it prints max_j |f_1(x_j)|, delta_1 = (1 + eps) max_j |f_1(x_j)|, max_j
|(f_1 + r_1)(x_j)| and delta_2 over the input points x_j (the same line as
extract_deltas.py:70-71) and writes the markdown table of extract_deltas.py:
73-83.  Shapes without both checkpoints are skipped with a note.  The partial
sums are evaluated on the GPU by the HIP library (FP32 by default); there is
no CPU path.
"""
import argparse
import os.path as osp
import sys

import numpy as np
import torch

sys.path.insert(0, osp.dirname(osp.abspath(__file__)))

from mplicits.deltas import partial_sum_maxima  # noqa: E402
from mplicits.ply import read_ply  # noqa: E402
from mplicits.util import from_pth  # noqa: E402

SHAPES = ["normalized_armadillo_gt", "normalized_asian_dragon_gt", "normalized_lucy_gt", "normalized_thai_gt_half"]
EPS = 0.3


def stage_dir(root, shape, stage):
    d = osp.join(root, "results", shape, stage)
    return d if osp.exists(osp.join(d, "best.pth")) else None


def shape_row(root, shape, eps=EPS, precision="fp32", device="cuda"):
    """(shape, n points, max|f1|, delta_1, max|f1 + r1|, delta_2), or None when
    a checkpoint is missing."""
    cd, md = stage_dir(root, shape, "coarse"), stage_dir(root, shape, "medium")
    if not cd or not md:
        return None
    v = read_ply(osp.join(root, "data", "normalized_sphere", "input", f"{shape}.ply"))["vertex"]
    pts = torch.from_numpy(np.stack([v["x"], v["y"], v["z"]], 1).astype(np.float32)).to(device)
    coarse = from_pth(osp.join(cd, "best.pth"), w0=1).eval().to(device)
    medium = from_pth(osp.join(md, "best.pth"), w0=1).eval().to(device)
    m1, m2 = partial_sum_maxima([coarse, medium], pts, precision)  # max |f_1|, max |f_1 + r_1|
    return shape, int(pts.shape[0]), m1, (1 + eps) * m1, m2, (1 + eps) * m2


def write_table(rows, path, eps=EPS):
    with open(path, "w", encoding="utf-8") as f:
        f.write(f"# Adaptive band widths (Eq. 5, eps={eps:g}) from paper checkpoints\n\n")
        f.write("Domain: unit sphere (diameter 2). The max runs over the input surface points x_j, NOT the domain.\n\n")
        f.write("| Shape | #points | max_j abs(f1(x_j)) | delta_1 | max_j abs((f1+r1)(x_j)) | delta_2 |\n")
        f.write("|---|---|---|---|---|---|\n")
        for r in rows:
            f.write(f"| {r[0]} | {r[1]} | {r[2]:.3e} | {r[3]:.3e} | {r[4]:.3e} | {r[5]:.3e} |\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description="Adaptive band widths (Eq. 5) from trained checkpoints.")
    ap.add_argument("--root", default=".", help="i3d root holding results/ and data/")
    ap.add_argument("--shapes", nargs="+", default=SHAPES)
    ap.add_argument("--eps", type=float, default=EPS)
    ap.add_argument("--out", default="deltas_table.md", help="markdown table to write")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16", "bf16x3", "bf16x3f"])
    ap.add_argument("--device", default="cuda")
    args = ap.parse_args(argv)
    rows = []
    for shape in args.shapes:
        r = shape_row(args.root, shape, args.eps, args.precision, args.device)
        if r is None:
            print(f"[skip] {shape}: missing checkpoints")
            continue
        rows.append(r)
        print(f"{shape}: n={r[1]}  max|f1|={r[2]:.3e} d1={r[3]:.3e}  max|f1+r1|={r[4]:.3e} d2={r[5]:.3e}")
    write_table(rows, args.out, args.eps)
    print(f"table -> {args.out}")
    return rows


if __name__ == "__main__":
    main()
