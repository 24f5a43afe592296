"""Multiscale sphere tracing on the GPU (mp_trace) and the renderer's
fixed-iteration frame (mp_render).

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


def renderer_camera(width, height):
    """(inv_view, inv_proj) of the renderer's initial arcball view:
    translate(0, 0, -3.5), 45 degree field of view, near 1e-5, far 1e4
    (renderer/cuda/camera.cu:9-17,77,131)."""
    from .synthetic import look_at, perspective
    return (np.linalg.inv(look_at((0.0, 0.0, 3.5))),
            np.linalg.inv(perspective(45.0, width / height, near=1e-5, far=1e4)))


def _params(width, height, deltas, max_steps, hit_eps, bound_radius, band_scale, pixel_center, flip_y):
    prm = _lib.mp_trace_params()
    prm.width, prm.height, prm.max_steps = int(width), int(height), int(max_steps)
    prm.hit_eps, prm.bound_radius, prm.band_scale = float(hit_eps), float(bound_radius), float(band_scale)
    prm.pixel_center, prm.flip_y, prm.n_deltas = int(bool(pixel_center)), int(bool(flip_y)), len(deltas)
    for i, d in enumerate(deltas):
        prm.deltas[i] = float(d)
    return prm


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
    prm = _params(width, height, deltas, max_steps, hit_eps, bound_radius, band_scale, pixel_center, flip_y)
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


def _render_params(width, height, iters, delta, threshold, normal_lod, shading, residual, skip_lod0, flip_y,
                   lod2_reads_lod0, bound_radius, band_scale):
    if shading not in ("phong", "normals"):
        raise ValueError(f"shading must be 'phong' or 'normals', got {shading!r}")
    iters = tuple(int(i) for i in iters) + (0,) * (3 - len(iters))
    prm = _lib.mp_render_params()
    prm.width, prm.height = int(width), int(height)
    for i in range(3):
        prm.iters[i] = iters[i]
    prm.delta, prm.threshold, prm.normal_lod = float(delta), float(threshold), int(normal_lod)
    prm.shading = _lib.MP_SHADE_PHONG if shading == "phong" else _lib.MP_SHADE_NORMALS
    prm.residual, prm.skip_lod0, prm.flip_y = int(bool(residual)), int(bool(skip_lod0)), int(bool(flip_y))
    prm.lod2_reads_lod0 = int(bool(lod2_reads_lod0))
    prm.bound_radius, prm.band_scale = float(bound_radius), float(band_scale)
    return prm


class FrameRenderer:
    """The renderer's fixed-iteration frame (mp_render): ``iters`` steps per
    level phase (level 0; levels 0 + 1; levels 0 (+ 1) + 2), Phong or normal
    shading with the normal of levels 0..normal_lod, white background --
    MIPplicit::SdfInference with the renderer's -iters / -delta / -threshold /
    -normal_lod / -shading / -skip_lod0 / -no_residual options
    (renderer/cuda/state.cu:816-1055, renderer/cuda/renderer.cu:754-790).

    Outputs are preallocated GPU tensors: rgba [n, H, W] int32 (RGBA8), pos
    [n, H, W, 3], dist [n, H, W], normals [n, H, W, 3].  ``frame(graph=True)``
    runs the first frame eagerly, captures the next into a CUDA/HIP graph and
    replays it afterwards (the reference captures its tracing loop the same
    way, state.cu:850)."""

    def __init__(self, stack, cameras, width, height, iters=(20, 0, 0), delta=0.0, threshold=0.05, normal_lod=0,
                 shading="phong", residual=True, skip_lod0=False, flip_y=False, lod2_reads_lod0=False,
                 bound_radius=1.2, band_scale=1.4):
        if not isinstance(stack, LevelStack):
            stack = LevelStack(stack, precision=getattr(stack[0], "precision", "fp32"))
        stack.refresh()
        self.stack = stack
        self.lib = _lib.load()
        cams = cameras if isinstance(cameras, torch.Tensor) else camera_block(cameras, stack.device)
        if cams.dtype != torch.float32 or cams.dim() != 2 or cams.shape[1] != 32:
            raise ValueError("cameras must be a [n, 32] float32 tensor or a list of (inv_view, inv_proj)")
        self.cams = cams.to(stack.device).contiguous()
        self.n = self.cams.shape[0]
        self.prm = _render_params(width, height, iters, delta, threshold, normal_lod, shading, residual, skip_lod0,
                                  flip_y, lod2_reads_lod0, bound_radius, band_scale)
        dev = stack.device
        self.rgba = torch.empty(self.n, height, width, dtype=torch.int32, device=dev)
        self.pos = torch.empty(self.n, height, width, 3, dtype=torch.float32, device=dev)
        self.dist = torch.empty(self.n, height, width, dtype=torch.float32, device=dev)
        self.normals = torch.empty(self.n, height, width, 3, dtype=torch.float32, device=dev)
        self._graph = None
        self._ran = False

    def _launch(self, stats=None):
        _lib.check(self.lib.mp_render(self.stack._handle, _lib.ptr(self.cams), self.n, ctypes.byref(self.prm),
                                      _lib.ptr(self.rgba), _lib.ptr(self.pos), _lib.ptr(self.dist),
                                      _lib.ptr(self.normals), stats, _lib.stream_handle()))

    def frame(self, graph=False):
        """Render one frame into the output tensors (asynchronous)."""
        with torch.cuda.device(self.stack.device):
            if not graph or not self._ran:
                self._launch()
                self._ran = True
                return
            if self._graph is None:
                torch.cuda.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._launch()
                self._graph = g
            self._graph.replay()

    def stats(self):
        """Render one frame eagerly and return (level evaluations [3], shaded
        pixels) (host-sync point)."""
        st = (ctypes.c_int64 * 4)()
        with torch.cuda.device(self.stack.device):
            self._launch(ctypes.cast(st, _lib.c_int64_p))
        return [int(st[i]) for i in range(3)], int(st[3])


def render(stack, cameras, width, height, **kw):
    """One fixed-iteration frame (see FrameRenderer).  Returns a dict of GPU
    tensors rgba, pos, dist, normals and the host-side stats evals (level
    evaluations performed) and shaded.  Synchronises the stream."""
    r = FrameRenderer(stack, cameras, width, height, **kw)
    evals, shaded = r.stats()
    return {"rgba": r.rgba, "pos": r.pos, "dist": r.dist, "normals": r.normals, "evals": evals, "shaded": shaded}


class BatchedTrace:
    """Many stacks, one camera each, traced by one persistent launch
    (mp_trace_batch_*); the run is captured once in a CUDA/HIP graph and
    replayed.

    jobs: sequence of (stack, camera, width, height, deltas) with ``stack`` a
    LevelStack (all in one precision) and ``camera`` an (inv_view, inv_proj)
    pair or a [32] float32 tensor.  Outputs per job: ``depth[j]`` [H, W],
    ``steps[j]``, ``rgba[j]`` and ``normals[j]`` [H, W, 3].
    """

    def __init__(self, jobs, max_steps=128, hit_eps=1e-4, bound_radius=1.2, band_scale=1.4,
                 pixel_center=True, flip_y=False, normals=True):
        lib = _lib.load()
        self._lib = lib
        self.stacks, self.cams = [], []
        self.depth, self.steps, self.rgba, self.normals = [], [], [], []
        cj = (_lib.mp_trace_job * len(jobs))()
        for j, (stack, cam, width, height, deltas) in enumerate(jobs):
            stack.refresh()
            dev = stack.device
            c = cam if isinstance(cam, torch.Tensor) else camera_block([cam], dev)[0]
            c = c.to(dev, torch.float32).reshape(32).contiguous()
            self.stacks.append(stack)
            self.cams.append(c)
            self.depth.append(torch.empty(height, width, dtype=torch.float32, device=dev))
            self.steps.append(torch.empty(height, width, dtype=torch.int32, device=dev))
            self.rgba.append(torch.empty(height, width, dtype=torch.int32, device=dev))
            self.normals.append(torch.empty(height, width, 3, dtype=torch.float32, device=dev) if normals else None)
            cj[j].stack = stack._handle.value
            cj[j].cam = c.data_ptr()
            cj[j].params = _params(width, height, list(deltas), max_steps, hit_eps, bound_radius, band_scale,
                                   pixel_center, flip_y)
            cj[j].depth = self.depth[-1].data_ptr()
            cj[j].normals = self.normals[-1].data_ptr() if normals else None
            cj[j].rgba = self.rgba[-1].data_ptr()
            cj[j].steps = self.steps[-1].data_ptr()
        handle = ctypes.c_void_p()
        _lib.check(lib.mp_trace_batch_plan(cj, len(jobs), _lib.stream_handle(), ctypes.byref(handle)))
        self._handle = handle
        self.graph = None

    def run(self, graph=True):
        if not graph:
            _lib.check(self._lib.mp_trace_batch_run(self._handle, _lib.stream_handle()))
            return
        if self.graph is None:
            self.run(graph=False)  # kernel attributes and code objects are set up eagerly
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                _lib.check(self._lib.mp_trace_batch_run(self._handle, _lib.stream_handle()))
            self.graph = g
        self.graph.replay()

    def stats(self):
        """Per job: evaluations per level and hits of the last run (synchronises)."""
        n = len(self.stacks)
        k = _lib.MP_MAX_LEVELS + 1
        buf = (ctypes.c_int64 * (n * k))()
        _lib.check(self._lib.mp_trace_batch_stats(self._handle, ctypes.cast(buf, _lib.c_int64_p),
                                                   _lib.stream_handle()))
        return [{"evals": [buf[j * k + l] for l in range(st.n_levels)], "hits": buf[j * k + k - 1]}
                for j, st in enumerate(self.stacks)]

    def close(self):
        self.graph = None
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self._lib.mp_trace_batch_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
