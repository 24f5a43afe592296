#!/usr/bin/env python
"""Headless counterpart of the renderer executable (renderer/cuda/renderer.cu):
its model loading (.exp descriptor and .bin checkpoints, state.cu:662-712),
its tracing options (-iters, -delta, -threshold, -normal_lod, -shading,
-skip_lod0, -no_residual; renderer.cu:754-790) and its -benchmark protocol
(renderer.cu:568-595: 30 warm-up frames, then N timed frames, the BENCHMARK
line, a line appended to benchmark.csv and the last frame written as a PPM).

    python m-plicits_amd/render.py -experiment_file=data/released/model.exp \\
        -data_root=. -benchmark=500 -iters=20,5,5 -delta=0.02 -normal_lod=2

The default camera is the renderer's initial arcball view (eye at (0, 0, 3.5)
looking down -z, 45 degree field of view).  Frames are the renderer's
fixed-iteration frames (mp_render), replayed from a captured graph as the
renderer replays its CUDA graph.  -adaptive traces with the persistent tracer
instead (mp_trace: per-ray hit test, at most -max_steps steps); its BENCHMARK
line reports iters=adaptive:<max_steps>.
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
from mplicits.tracer import FrameRenderer, camera_block, renderer_camera, trace  # noqa: E402

WARMUP_FRAMES = 30  # g_bench_warmup, renderer.cu:82


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
    ap.add_argument("--benchmark", type=int, default=None, help="frames to time (<= 0: 500; absent: one frame)")
    ap.add_argument("--iters", default="20,0,0", help="sphere-tracing steps of the level 0 / 1 / 2 phases")
    ap.add_argument("--delta", default="0", help="band of level 0 (and of level 1: d0,d1)")
    ap.add_argument("--threshold", type=float, default=0.05)
    ap.add_argument("--normal_lod", type=int, default=0)
    ap.add_argument("--shading", default="phong", choices=["phong", "normals"])
    ap.add_argument("--skip_lod0", action="store_true")
    ap.add_argument("--no_residual", action="store_true")
    ap.add_argument("--lod2_reads_lod0", action="store_true",
                    help="level 2's distances read level 0's output, as the renderer binds them")
    ap.add_argument("--adaptive", action="store_true", help="persistent tracer with a per-ray hit test")
    ap.add_argument("--max_steps", type=int, default=128)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--precision", default="fp16", choices=["fp32", "bf16", "fp16"],
                    help="the renderer computes in FP16")
    ap.add_argument("--csv", default="benchmark.csv")
    ap.add_argument("--out", default="img_render.ppm")
    return ap.parse_args(norm)


def write_ppm(path, rgba, flip_y):
    v = rgba.cpu().numpy().astype(np.int64) & 0xFFFFFFFF
    rgb = np.stack([(v >> s) & 255 for s in (0, 8, 16)], -1).astype(np.uint8)
    h, w = rgb.shape[:2]
    with open(path, "wb") as f:
        f.write(f"P6 {w} {h} 255\n".encode())
        f.write(rgb[::-1].tobytes() if not flip_y else rgb.tobytes())


def main(argv=None):
    args = _args(sys.argv[1:] if argv is None else argv)
    if not torch.cuda.is_available():
        raise SystemExit("render.py needs a GPU (frames are drawn only through the HIP library)")
    width, height = (int(v) for v in args.resolution.lower().split("x"))
    iters = [int(v) for v in args.iters.split(",")] + [0, 0]
    iters = iters[:3]
    deltas = [float(v) for v in args.delta.split(",")]
    d0, d1 = deltas[0], (deltas[1] if len(deltas) > 1 else deltas[0])
    levels, desc = load_experiment(args.experiment_file, data_root=args.data_root, device="cuda")
    flip_y = desc.get("flip_y", False)
    stack = LevelStack(levels, args.precision)
    cams = camera_block([renderer_camera(width, height)])
    frames = None if args.benchmark is None else (args.benchmark if args.benchmark > 0 else 500)

    if args.adaptive:
        kw = dict(deltas=[d0, d1][:stack.n_levels - 1], max_steps=args.max_steps, hit_eps=args.eps, flip_y=flip_y,
                  pixel_center=False)

        def draw():
            return trace(stack, cams, width, height, **kw)
        out = draw()
        rgba = out["rgba"][0]
        iters_txt = f"adaptive:{args.max_steps}"
        info = f"hits {out['hits']}, evaluations per level {out['evals']}"
    else:
        r = FrameRenderer(stack, cams, width, height, iters=iters, delta=d0, threshold=args.threshold,
                          normal_lod=args.normal_lod, shading=args.shading, residual=not args.no_residual,
                          skip_lod0=args.skip_lod0, flip_y=flip_y, lod2_reads_lod0=args.lod2_reads_lod0)
        evals, shaded = r.stats()

        def draw():
            r.frame(graph=True)
        rgba = r.rgba[0]
        iters_txt = "/".join(str(i) for i in iters)
        info = f"shaded {shaded}, evaluations per level {evals}"
    if frames:
        for _ in range(WARMUP_FRAMES):
            draw()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            out = draw()
        torch.cuda.synchronize()
        fps = frames / (time.perf_counter() - t0)
        if args.adaptive:
            rgba = out["rgba"][0]
        print(f"BENCHMARK resolution={width}x{height} iters={iters_txt} delta={d0:.4f} "
              f"normal_lod={args.normal_lod} frames={frames} avg_fps={fps:.1f}")
        if args.csv:
            with open(args.csv, "a", encoding="utf-8") as f:
                f.write(f"{width}x{height},{iters_txt},{d0:.4f},{args.normal_lod},{frames},{fps:.2f}\n")
    write_ppm(args.out, rgba, flip_y)
    print(f"{info}, image {args.out}")


if __name__ == "__main__":
    main()
