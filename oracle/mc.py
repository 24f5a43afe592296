"""ctypes wrapper of the C Lewiner marching-cubes restatement (oracle/mc_lewiner.c)."""
import ctypes
import os.path as osp

import numpy as np

from . import build as obuild

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not osp.exists(obuild.LIB):
            obuild.build()
        lib = ctypes.CDLL(obuild.LIB)
        lib.mc_oracle_run.restype = ctypes.c_int64
        lib.mc_oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        _lib = lib
    return _lib


def triangles(vol, level=0.0):
    """(positions [T, 3, 3] float32 in voxel units, axis0..2 order; keys [T, 3])."""
    lib = _load()
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    nz, ny, nx = vol.shape
    cap = max(1024, 4 * vol.size)
    while True:
        pos = np.empty((cap, 3, 3), np.float32)
        key = np.empty((cap, 3), np.int64)
        n = lib.mc_oracle_run(vol.ctypes.data, nz, ny, nx, level, pos.ctypes.data, key.ctypes.data, cap)
        if n <= cap:
            return pos[:n], key[:n]
        cap = int(n)


def canonical(tri_pos):
    """Order-free form of a triangle list: each triangle rotated (orientation
    kept) to start at its lexicographically smallest vertex, then the list
    sorted.  Returns [T, 9] float32."""
    t = np.asarray(tri_pos, dtype=np.float32).reshape(-1, 3, 3)
    if len(t) == 0:
        return t.reshape(0, 9)
    # index of the smallest vertex (lexicographic over x, y, z)
    order = np.lexsort((t[:, :, 2], t[:, :, 1], t[:, :, 0]), axis=-1)[:, 0]
    rot = np.stack([t[np.arange(len(t)), (order + k) % 3] for k in range(3)], axis=1)
    flat = rot.reshape(-1, 9)
    idx = np.lexsort(flat.T[::-1])
    return flat[idx]


def mesh_triangles(verts, faces):
    """[T, 3, 3] positions of a (verts, faces) mesh."""
    verts = np.asarray(verts, np.float32)
    faces = np.asarray(faces, np.int64)
    return verts[faces] if len(faces) else np.zeros((0, 3, 3), np.float32)


_OX = np.array([0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0])
_OY = np.array([0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1])
_OZ = np.array([0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0])
_AX = np.array([0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2])


def global_vertex_ids(key, shape):
    """Map (cell * 13 + local edge) keys (cell = linear index of the cell's
    lowest grid point) to ids shared between cells: a grid
    edge is (lower point, axis) -> point * 4 + axis; a cube centre -> cell
    point * 4 + 3."""
    nz, ny, nx = shape
    key = np.asarray(key, np.int64)
    cell, e = key // 13, key % 13
    cx, cy, cz = cell % nx, (cell // nx) % ny, cell // (nx * ny)
    centre = e == 12
    ee = np.where(centre, 0, e)
    px = cx + np.where(centre, 0, _OX[ee])
    py = cy + np.where(centre, 0, _OY[ee])
    pz = cz + np.where(centre, 0, _OZ[ee])
    ax = np.where(centre, 3, _AX[ee])
    return ((pz * ny + py) * nx + px) * 4 + ax
