"""Training-data generation on the GPU: drop-ins for the sampling helpers of
experiment_scripts/utils.py and for the part of open3d's RaycastingScene the
trainers use (add_triangles, compute_distance;
experiment_scripts/train_sdf_on_neighborhood_fine.py:145-146, utils.py:77-78).

The random draws are the reference's own torch calls in the same order
(torch.randperm on ``device``, then torch.rand on the CPU generator), so a
seeded run samples the same points; the arithmetic runs in one HIP kernel per
call (mp_neighborhood_samples), the point-to-mesh distances on a GPU
bounding-volume hierarchy (mp_mesh_*), and calc_deltas' search loop on the
device (mp_calc_deltas).
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _host(a, dtype):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


class RaycastingScene:
    """Unsigned point-to-mesh distances (open3d.t.geometry.RaycastingScene's
    add_triangles / compute_distance).  One mesh per scene."""

    def __init__(self):
        self._lib = _lib.load()
        self._handle = None
        self._keep = None

    def add_triangles(self, vertices, faces=None):
        """vertices [V, 3] and faces [F, 3] (arrays or tensors), or an object
        with ``vertex["positions"]`` / ``triangle["indices"]`` (the open3d
        tensor mesh layout).  Builds the hierarchy; returns the geometry id 0."""
        if faces is None:
            vertices, faces = vertices.vertex["positions"], vertices.triangle["indices"]
            vertices = vertices.numpy() if hasattr(vertices, "numpy") else vertices
            faces = faces.numpy() if hasattr(faces, "numpy") else faces
        if self._handle is not None:
            raise NotImplementedError("this scene holds one mesh")
        v = _host(vertices, np.float32).reshape(-1, 3)
        f = _host(faces, np.int32).reshape(-1, 3)
        h = ctypes.c_void_p()
        _lib.check(self._lib.mp_mesh_create(v.ctypes.data_as(ctypes.c_void_p), len(v),
                                            f.ctypes.data_as(ctypes.c_void_p), len(f), _lib.stream_handle(),
                                            ctypes.byref(h)))
        self._handle = h
        return 0

    def info(self):
        """(triangles, hierarchy nodes, hierarchy depth)."""
        t, n, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        _lib.check(self._lib.mp_mesh_info(self._handle, ctypes.byref(t), ctypes.byref(n), ctypes.byref(d)))
        return t.value, n.value, d.value

    def compute_distance(self, points):
        """Unsigned distance [...] of points [..., 3] (a GPU tensor) to the mesh."""
        if self._handle is None:
            raise ValueError("add_triangles first")
        _lib.require_gpu_tensor(points, "points")
        p = points.reshape(-1, 3).to(torch.float32).contiguous()
        out = torch.empty(p.shape[0], device=p.device, dtype=torch.float32)
        _lib.check(self._lib.mp_mesh_distance(self._handle, _lib.ptr(p), p.shape[0], _lib.ptr(out),
                                              _lib.stream_handle()))
        return out.reshape(points.shape[:-1])

    def close(self):
        if self._handle is not None:
            self._lib.mp_mesh_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sample_on_surface(vertices, n_points, device):
    """utils.py:10-43: n_points random rows of ``vertices`` (all of them, in
    order, when n_points equals the row count) and their indices."""
    if n_points == vertices.shape[0]:
        return vertices, torch.arange(start=0, end=n_points, device=device)
    idx = torch.randperm(vertices.shape[0], device=device)[:n_points]
    return vertices[idx, ...], idx


def calc_deltas(vertices, scene, max_delta=0.05, n_iters=100, device=torch.device("cpu")):
    """utils.py:46-89: [V, 7] = (x, y, z, unit normal, delta), delta the
    offset along the normal whose distance to the mesh equals it."""
    _lib.require_gpu_tensor(vertices, "vertices")
    v = vertices.detach().to(torch.float32).contiguous()
    out = torch.empty(v.shape[0], 7, device=v.device, dtype=torch.float32)
    done = ctypes.c_int()
    lib = _lib.load()
    _lib.check(lib.mp_calc_deltas(scene._handle, _lib.ptr(v), v.shape[0], v.shape[1], float(max_delta), int(n_iters),
                                  _lib.ptr(out), ctypes.byref(done), _lib.stream_handle()))
    calc_deltas.iterations = done.value
    return out


def create_training_data_on_neighborhood(vertices, n_on_surf, device=torch.device("cpu"), delta=0.02,
                                         use_neigh_delta=False, dithering_sampling=True, delta_eps=0.0005):
    """utils.py:92-169: {"on_surf": [points, normals, sdfs]} with the sampled
    vertices (sdf 0), their offsets along the normal (sdf = offset) and, with
    dithering, band points around p + delta / 2 n labelled -1."""
    _lib.require_gpu_tensor(vertices, "vertices")
    v = vertices.detach().to(torch.float32).contiguous()
    V = v.shape[0]
    if n_on_surf == V:
        idx = None
    else:
        idx = torch.randperm(V, device=device)[:n_on_surf].to(v.device)
    N = min(n_on_surf, V)
    rnd = torch.rand((N, 3)).to(v.device)
    m = (3 if dithering_sampling else 2) * N
    pts = torch.empty(m, 3, device=v.device, dtype=torch.float32)
    nrm = torch.empty(m, 3, device=v.device, dtype=torch.float32)
    sdf = torch.empty(m, device=v.device, dtype=torch.float32)
    lib = _lib.load()
    _lib.check(lib.mp_neighborhood_samples(_lib.ptr(v), V, v.shape[1], _lib.ptr(idx), _lib.ptr(rnd), N, float(delta),
                                           float(delta_eps), int(bool(use_neigh_delta)), int(bool(dithering_sampling)),
                                           _lib.ptr(pts), _lib.ptr(nrm), _lib.ptr(sdf), _lib.stream_handle()))
    return {"on_surf": [pts, nrm, sdf]}
