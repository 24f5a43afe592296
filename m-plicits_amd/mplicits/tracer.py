"""Multiscale sphere tracing on the GPU (mp_trace).

The renderer-side counterpart of the reference: its CUDA renderer traces a
level stack with fixed iteration counts per level on every pixel
(renderer/cuda/state.cu:848-881) and shades with GEMM normals
(renderer/cuda/gemms.cu:263-337, kernels.cu:122-275); the Python version is
renderer/python/siren-mips.py:126-248.  Here one persistent kernel traces all
cameras with a hit test and ray compaction, then the hit points get analytic
normals and Lambert shading.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .stack import LevelStack


def camera_block(cameras, device="cuda"):
    """[n_cams, 32] float32: per camera inv_view then inv_proj, column-major
    (glm layout).  ``cameras``: iterable of (inv_view, inv_proj) 4x4 matrices in
    the usual row-major math layout."""
    rows = [np.concatenate([np.asarray(iv, np.float64).T.reshape(-1), np.asarray(ip, np.float64).T.reshape(-1)])
            for iv, ip in cameras]
    return torch.tensor(np.stack(rows), dtype=torch.float32, device=device)


def trace(stack, cameras, width, height, deltas=(), max_steps=128, hit_eps=1e-4, bound_radius=1.2,
          band_scale=1.4, pixel_center=True, flip_y=False, normals=True):
    """Trace every pixel of every camera through ``stack`` (a LevelStack or a
    list of SIREN levels, coarse first).

    Returns a dict of GPU tensors: depth [n, H, W] (-1 = miss), steps [n, H, W],
    rgba [n, H, W] int32 (RGBA8), normals [n, H, W, 3] (if requested), and
    host-side stats (evaluations per level, hits).  Synchronises the stream.
    """
    if not isinstance(stack, LevelStack):
        stack = LevelStack(stack, precision=getattr(stack[0], "precision", "fp32"))
    stack.refresh()
    lib = _lib.load()
    cams = cameras if isinstance(cameras, torch.Tensor) else camera_block(cameras, stack.device)
    if cams.dtype != torch.float32 or cams.dim() != 2 or cams.shape[1] != 32:
        raise ValueError("cameras must be a [n, 32] float32 tensor or a list of (inv_view, inv_proj)")
    cams = cams.to(stack.device).contiguous()
    n = cams.shape[0]
    deltas = [float(d) for d in deltas]
    if len(deltas) != stack.n_levels - 1:
        raise ValueError(f"need {stack.n_levels - 1} deltas, got {len(deltas)}")
    prm = _lib.mp_trace_params()
    prm.width, prm.height, prm.max_steps = int(width), int(height), int(max_steps)
    prm.hit_eps, prm.bound_radius, prm.band_scale = float(hit_eps), float(bound_radius), float(band_scale)
    prm.pixel_center, prm.flip_y, prm.n_deltas = int(bool(pixel_center)), int(bool(flip_y)), len(deltas)
    for i, d in enumerate(deltas):
        prm.deltas[i] = d
    dev = stack.device
    depth = torch.empty(n, height, width, dtype=torch.float32, device=dev)
    steps = torch.empty(n, height, width, dtype=torch.int32, device=dev)
    rgba = torch.empty(n, height, width, dtype=torch.int32, device=dev)
    nrm = torch.empty(n, height, width, 3, dtype=torch.float32, device=dev) if normals else None
    stats = (ctypes.c_int64 * (_lib.MP_MAX_LEVELS + 1))()
    with torch.cuda.device(dev):
        _lib.check(lib.mp_trace(stack._handle, _lib.ptr(cams), n, ctypes.byref(prm), _lib.ptr(depth),
                                _lib.ptr(nrm), _lib.ptr(rgba), _lib.ptr(steps),
                                ctypes.cast(stats, _lib.c_int64_p), _lib.stream_handle()))
    out = {"depth": depth, "steps": steps, "rgba": rgba,
           "evals": [int(stats[l]) for l in range(stack.n_levels)], "hits": int(stats[_lib.MP_MAX_LEVELS])}
    if normals:
        out["normals"] = nrm
    return out
