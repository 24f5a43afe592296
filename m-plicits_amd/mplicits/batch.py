"""Batched extraction of many level stacks (uneven widths, levels and
resolutions) through mp_extract_batch_*: one persistent launch per stage, the
whole run captured once in a CUDA/HIP graph and replayed.

The reference extracts one model per ``create_mesh_multistage`` call
(i3d/meshing.py:135-264); every model here follows the same per-chunk culling
semantics as ``LevelStack.extract_volume``.
"""
import ctypes

import torch

from . import _lib
from .stack import LevelStack


class BatchedExtraction:
    """Plan of a batched band-culled extraction.

    Parameters
    ----------
    stacks: LevelStack objects (all in one precision), or lists of SIREN levels
        (coarse first) together with ``precision``.
    resolutions: grid resolution N per model.
    deltas: per model, the band widths (len = levels - 1; empty = no culling).
    grid: "reference" (the warped grid of i3d/meshing.py:42-63) or "lattice".
    chunk: culling chunk (max_batch of the reference loop).
    """

    def __init__(self, stacks, resolutions, deltas, grid="reference", chunk=32 ** 3, precision="bf16"):
        lib = _lib.load()
        self._lib = lib
        self.stacks = [s if isinstance(s, LevelStack) else LevelStack(s, precision) for s in stacks]
        if not (len(self.stacks) == len(resolutions) == len(deltas)) or not self.stacks:
            raise ValueError("stacks, resolutions and deltas must have the same non-zero length")
        self.resolutions = [int(N) for N in resolutions]
        self.deltas = [list(map(float, d)) for d in deltas]
        dev = self.stacks[0].device
        self.volumes = [torch.empty(N ** 3, device=dev, dtype=torch.float32) for N in self.resolutions]
        for st in self.stacks:
            st.refresh()
        jobs = (_lib.mp_extract_job * len(self.stacks))()
        for j, (st, N, d, vol) in enumerate(zip(self.stacks, self.resolutions, self.deltas, self.volumes)):
            if d and len(d) != st.n_levels - 1:
                raise ValueError(f"model {j}: need {st.n_levels - 1} deltas, got {len(d)}")
            jobs[j].stack = st._handle.value
            jobs[j].N = N
            jobs[j].grid_mode = _lib.GRID_MODES[grid]
            jobs[j].chunk = int(chunk)
            jobs[j].n_deltas = len(d)
            for i, v in enumerate(d):
                jobs[j].deltas[i] = v
            jobs[j].vol = vol.data_ptr()
        handle = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(lib.mp_extract_batch_plan(jobs, len(self.stacks), _lib.stream_handle(),
                                                 ctypes.byref(handle)))
        self._handle = handle
        self.graph = None

    def run(self, graph=True):
        """Run the batch on torch's current stream.  With ``graph`` the launches
        are captured once (after one eager run) and replayed afterwards."""
        if not graph:
            _lib.check(self._lib.mp_extract_batch_run(self._handle, _lib.stream_handle()))
            return
        if self.graph is None:
            self.run(graph=False)  # kernel attributes and code objects are set up eagerly
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                _lib.check(self._lib.mp_extract_batch_run(self._handle, _lib.stream_handle()))
            self.graph = g
        self.graph.replay()

    def counts(self):
        """Samples evaluated per model and level by the last run (synchronises)."""
        n = len(self.stacks)
        buf = (ctypes.c_int64 * (n * _lib.MP_MAX_LEVELS))()
        _lib.check(self._lib.mp_extract_batch_counts(self._handle, ctypes.cast(buf, _lib.c_int64_p),
                                                      _lib.stream_handle()))
        return [[buf[j * _lib.MP_MAX_LEVELS + l] for l in range(st.n_levels)]
                for j, st in enumerate(self.stacks)]

    def volume(self, j):
        N = self.resolutions[j]
        return self.volumes[j].view(N, N, N)

    def close(self):
        self.graph = None
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self._lib.mp_extract_batch_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
