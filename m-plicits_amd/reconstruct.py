#!/usr/bin/env python
"""Mesh extraction CLI, drop-in for the reference's reconstruct.py.

Same arguments, prints (model summaries, SAMPLING_TIME_S:<s>, Done) and
outputs as reconstruct.py:36-133; the models are evaluated and meshed on the
GPU by the HIP library.  Extra options: --precision (fp32 | bf16 | fp16 | bf16x3 | bf16x3f) and
--grid (reference | lattice).
"""
import argparse
import os
import os.path as osp
import sys

import numpy as np
import torch

sys.path.insert(0, osp.dirname(osp.abspath(__file__)))

from mplicits.meshing import create_mesh_multistage  # noqa: E402
from mplicits.util import from_pth  # noqa: E402


def adaptive_deltas(models, input_ply, device, eps=0.3, batch=2 ** 18):
    """delta_i = (1 + eps) * max_j |F_i(x_j)| over the training points, F_i the
    partial sum up to level i (reconstruct.py:18-33, Eq. 5 of the paper).
    ``input_ply``: path of a PLY point cloud (x, y, z vertex properties)."""
    from mplicits.deltas import partial_sum_maxima
    from mplicits.ply import read_ply
    v = read_ply(input_ply)["vertex"]
    pts = torch.from_numpy(np.stack([v["x"], v["y"], v["z"]], 1).astype(np.float32)).to(device)
    maxima = partial_sum_maxima(models, pts, getattr(models[0], "precision", "fp32"), batch, count=len(models) - 1)
    return [(1.0 + eps) * m for m in maxima]


def main(argv=None):
    parser = argparse.ArgumentParser(description="Run marching cubes using a trained model.")
    parser.add_argument("model_path", help="Path to the PyTorch weights file")
    parser.add_argument("--coarse_path", default=None,
                        help="Path to the PyTorch weights file for the coarse network")
    parser.add_argument("--medium_path", default=None, help="Path to the medium resolution model. "
                        "Only use this when reconstructing a fine-level model.")
    parser.add_argument("output_path", help="Path to the output mesh file")
    parser.add_argument("--w0", type=int, default=1, help="Value for \\omega_0. Default is 1.")
    parser.add_argument("--resolution", "-r", default=128, type=int,
                        help="Resolution to use on marching cubes. Default is 128.")
    parser.add_argument("--device", default="cpu", help="Device to run the inference.")
    parser.add_argument("--multistage", action="store_true",
                        help="Use the multiscale optimization (finer levels only evaluated near the surface).")
    parser.add_argument("--input", default=None, help="Training point cloud (.ply) for the adaptive deltas.")
    parser.add_argument("--deltas", type=float, nargs="+", default=None,
                        help="Explicit band widths (one per residual level).")
    parser.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16", "bf16x3", "bf16x3f"],
                        help="Arithmetic of the hidden-layer GEMMs.")
    parser.add_argument("--grid", default="reference", choices=["reference", "lattice"],
                        help="Sample coordinates: the reference's float32 grid, or the regular lattice.")
    args = parser.parse_args(argv)
    out_dir = osp.split(args.output_path)[0]
    if out_dir and not osp.exists(out_dir):
        os.makedirs(out_dir)

    if not torch.cuda.is_available():
        raise RuntimeError("this implementation runs on the GPU only (HIP); no GPU is visible")
    device = args.device if "cuda" in args.device else "cuda"
    if device != args.device:
        print(f"[WARNING] Selected \"{args.device}\" for inference; the HIP path runs on \"{device}\".")
    device = torch.device(device)

    model = from_pth(args.model_path, w0=args.w0).eval().to(device)
    model.precision = args.precision
    print(model)
    model_list = [model]
    print(f"Running marching cubes running with resolution {args.resolution}")
    if args.medium_path is not None:
        print("Using residual approach for fine level")
        model_medium = from_pth(args.medium_path, w0=1).eval().to(device)
        model_list.append(model_medium)
        print("========== Medium model ==========")
        print(model_medium)
    if args.coarse_path is not None:
        print("Using the residual approach")
        model_coarse = from_pth(args.coarse_path, w0=1).eval().to(device)
        model_list.append(model_coarse)
        print("========== Coarse model ==========")
        print(model_coarse)

    if args.multistage:
        if args.deltas is not None:
            deltas = list(args.deltas)
        elif args.input is not None:
            deltas = adaptive_deltas(model_list[::-1], args.input, device)
            print("adaptive band widths (Eq. 5):", ", ".join(f"{d:.3e}" for d in deltas))
        else:
            deltas = [0.1, 0.06][:len(model_list) - 1]
            print("[WARNING] --multistage without --input: using fixed band widths"
                  f" {deltas}; pass the training point cloud for the paper's adaptive bands.")
    else:
        deltas = []

    _, _, _, _, sampling_time = create_mesh_multistage(
        model_list[::-1], deltas=deltas, filename=args.output_path, N=args.resolution,
        max_batch=32 ** 3, device=device, precision=args.precision, grid=args.grid)
    print(f"SAMPLING_TIME_S:{sampling_time}")
    print("Done")


if __name__ == "__main__":
    main()
