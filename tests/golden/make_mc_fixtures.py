"""Writes tests/golden/mc_lewiner_fixtures.npz: marching-cubes golden outputs.

The reference meshes with skimage.measure.marching_cubes(vol, 0.0,
spacing=[h]*3) (i3d/meshing.py:296-298; scikit-image 0.25.2 pinned in
requirements.txt:7).  That version is not installed here; scikit-image 0.18.3
is (under /opt/conda/bin/python3.9), with the same Lewiner implementation and
default arguments (gradient_direction='descent', allow_degenerate=True).  This
script runs it on small volumes and stores inputs and outputs:

  cubeNNN  : every one of the 256 corner sign patterns of a single 2x2x2 cube,
             8 random magnitude draws each (exercises every case, sub-case and
             ambiguity test of the tables)
  rand*    : random-valued volumes (many ambiguous faces and interiors)
  zeros*   : volumes with exact zeros at grid points
  sphere24, bumpy32 : smooth SDFs

Usage: /opt/conda/bin/python3.9 tests/golden/make_mc_fixtures.py
"""
import os.path as osp

import numpy as np
from skimage import __version__ as skv
from skimage.measure import marching_cubes

OUT = osp.join(osp.dirname(osp.abspath(__file__)), "mc_lewiner_fixtures.npz")


def run(vol):
    try:
        v, f, n, val = marching_cubes(vol, 0.0)
    except (ValueError, RuntimeError):
        v, f, n, val = (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32),
                        np.zeros((0, 3), np.float32), np.zeros(0, np.float32))
    return (np.asarray(v, np.float32), np.asarray(f, np.int32), np.asarray(n, np.float32),
            np.asarray(val, np.float32))


def main():
    rng = np.random.default_rng(20240611)
    vols = {}
    # corner p of the cube at (x, y, z) = ((p ^ (p >> 1)) & 1, (p >> 1) & 1, (p >> 2) & 1),
    # volume index [z, y, x] (x is the fastest axis)
    for pat in range(256):
        for d in range(8):
            vol = np.empty((2, 2, 2), np.float32)
            for p in range(8):
                x, y, z = (p ^ (p >> 1)) & 1, (p >> 1) & 1, (p >> 2) & 1
                mag = rng.uniform(0.05, 1.0)
                vol[z, y, x] = mag if (pat >> p) & 1 else -mag
            vols[f"cube{pat:03d}_{d}"] = vol
    for i, shape in enumerate([(3, 3, 3), (4, 5, 6), (6, 7, 8), (9, 9, 9), (12, 10, 11)]):
        for k in range(4):
            vols[f"rand{i}_{k}"] = rng.normal(size=shape).astype(np.float32)
    for k in range(4):
        v = rng.normal(size=(7, 8, 9)).astype(np.float32)
        v[rng.random(v.shape) < 0.2] = 0.0
        vols[f"zeros{k}"] = v
    g = np.linspace(-1, 1, 24, dtype=np.float32)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    vols["sphere24"] = (np.sqrt(x * x + y * y + z * z) - 0.6).astype(np.float32)
    g = np.linspace(-1, 1, 32, dtype=np.float32)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    vols["bumpy32"] = (r - 0.5 - 0.08 * np.exp(-((x - 0.5) ** 2 + y ** 2 + z ** 2) / 0.02)
                       - 0.05 * np.sin(7 * x) * np.sin(5 * y) * np.sin(3 * z)).astype(np.float32)
    out = {"skimage_version": np.array(skv)}
    for name, vol in vols.items():
        v, f, n, val = run(vol)
        out[name + "/vol"] = vol
        out[name + "/verts"] = v
        out[name + "/faces"] = f
        out[name + "/normals"] = n
        out[name + "/values"] = val
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, "with", len(vols), "volumes; scikit-image", skv)


if __name__ == "__main__":
    main()
