#!/usr/bin/env python
"""Headless counterpart of the renderer executable's model loading and
-benchmark protocol (renderer/cuda/renderer.cu:568-595, state.cu:662-712):
load an .exp descriptor and its .bin checkpoints, sphere-trace the three-level
stack on the GPU, write the last frame as a PPM and print the BENCHMARK line.

    python m-plicits_amd/render.py -experiment_file=data/released/model.exp \\
        -data_root=. -resolution=512x512 -benchmark=100 -delta=0.02

Tracing runs until every ray hits (|F| < -eps) or leaves the bounding sphere
(at most -max_steps steps) instead of the renderer's fixed per-level
iteration counts; the BENCHMARK line reports that as iters=adaptive:<max>.
"""
import argparse
import os.path as osp
import sys
import time

import numpy as np
import torch

sys.path.insert(0, osp.dirname(osp.abspath(__file__)))
from mplicits.renderer_io import load_experiment  # noqa: E402
from mplicits.stack import LevelStack  # noqa: E402
from mplicits.synthetic import look_at, perspective  # noqa: E402
from mplicits.tracer import camera_block, trace  # noqa: E402


def _args(argv):
    # the renderer takes -key=value flags; accept --key value too
    norm = []
    for a in argv:
        if a.startswith("-") and not a.startswith("--"):
            a = "-" + a
        norm.append(a)
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--experiment_file", required=True)
    ap.add_argument("--data_root", default=None)
    ap.add_argument("--resolution", default="512x512")
    ap.add_argument("--benchmark", type=int, default=0, help="frames to time (0: render one frame)")
    ap.add_argument("--delta", type=float, default=0.02, help="band width of the coarse level")
    ap.add_argument("--delta1", type=float, default=None, help="band width of the medium level (default: delta)")
    ap.add_argument("--max_steps", type=int, default=128)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--out", default="img_render.ppm")
    return ap.parse_args(norm)


def main(argv=None):
    args = _args(sys.argv[1:] if argv is None else argv)
    if not torch.cuda.is_available():
        raise SystemExit("render.py needs a GPU (the tracer runs only through the HIP library)")
    width, height = (int(v) for v in args.resolution.lower().split("x"))
    levels, desc = load_experiment(args.experiment_file, data_root=args.data_root, device="cuda")
    stack = LevelStack(levels, args.precision)
    deltas = [args.delta, args.delta if args.delta1 is None else args.delta1]
    eye = np.array([0.0, 0.0, 2.5])
    cams = camera_block([(np.linalg.inv(look_at(eye)), np.linalg.inv(perspective(45.0, width / height)))])
    kw = dict(deltas=deltas, max_steps=args.max_steps, hit_eps=args.eps, flip_y=desc.get("flip_y", False))
    out = trace(stack, cams, width, height, **kw)
    if args.benchmark > 0:
        for _ in range(3):
            trace(stack, cams, width, height, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.benchmark):
            out = trace(stack, cams, width, height, **kw)
        torch.cuda.synchronize()
        fps = args.benchmark / (time.perf_counter() - t0)
        print(f"BENCHMARK resolution={width}x{height} iters=adaptive:{args.max_steps} delta={args.delta:.4f} "
              f"normal_lod=2 frames={args.benchmark} avg_fps={fps:.1f}")
    rgba = out["rgba"][0].cpu().numpy().astype(np.uint32)
    rgb = np.stack([(rgba >> s) & 255 for s in (0, 8, 16)], -1).astype(np.uint8)
    with open(args.out, "wb") as f:
        f.write(f"P6 {width} {height} 255\n".encode())
        f.write(rgb[::-1].tobytes() if not desc.get("flip_y", False) else rgb.tobytes())
    print(f"hits {out['hits']}, evaluations per level {out['evals']}, image {args.out}")


if __name__ == "__main__":
    main()
