"""Writes tests/golden/bumpy_surface_2k.ply: a training point cloud for the
adaptive band widths of reconstruct.py (--input): 2048 points on the seeded
bumpy sphere (the workload's target shape), each moved by N(0, 0.01) noise, as
a binary PLY with float x, y, z vertices and no faces.

Usage: python tests/golden/make_point_cloud.py
"""
import os.path as osp
import sys

import numpy as np

ROOT = osp.dirname(osp.dirname(osp.dirname(osp.abspath(__file__))))
sys.path.insert(0, osp.join(ROOT, "m-plicits_amd"))

from mplicits.ply import save_ply  # noqa: E402
from mplicits.synthetic import MASTER_SEED, BumpySphere  # noqa: E402

OUT = osp.join(osp.dirname(osp.abspath(__file__)), "bumpy_surface_2k.ply")


def main():
    shape = BumpySphere()
    pts = shape.surface_points(2048, MASTER_SEED + 70).numpy()
    rng = np.random.Generator(np.random.PCG64(MASTER_SEED + 71))
    pts = pts + rng.normal(0.0, 0.01, size=pts.shape)
    save_ply(pts.astype(np.float32), np.zeros((0, 3), np.int32), OUT)
    print(OUT, len(pts))


if __name__ == "__main__":
    main()
