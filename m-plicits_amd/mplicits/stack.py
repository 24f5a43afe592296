"""Host handle of a SIREN level stack on the GPU (mp_stack in the C ABI).

A ``LevelStack`` borrows the float32 weights of one or more SIREN levels (coarse
first), packs them into MFMA fragment order in the requested precision and runs
the HIP entry points on torch's current stream.  Parameter changes (optimizer
steps, ``load_state_dict``) are picked up through the tensors' version counters
and trigger a re-pack on the device.
"""
import ctypes

import torch

from . import _lib


def _linear_layers(model):
    """(weight, bias, omega) per layer of an i3d-style SIREN, in order.

    Accepts this package's SIREN and any module with the reference's
    ``net.{k}.0.weight/bias`` layout (e.g. the reference SIREN).
    """
    net = getattr(model, "net", None)
    if net is None:
        raise TypeError(f"{type(model).__name__} has no .net: not an i3d SIREN")
    layers = []
    n = len(net)
    for k, seq in enumerate(net):
        lin = seq[0]
        if not isinstance(lin, torch.nn.Linear):
            raise TypeError(f"layer {k} of {type(model).__name__} is not a Linear")
        omega = float(seq[1].w0) if (k < n - 1 and len(seq) > 1) else 1.0
        layers.append((lin.weight, lin.bias, omega))
    return layers


class LevelStack:
    """Levels (coarse first) packed for the HIP kernels.

    Parameters
    ----------
    models: sequence of SIREN modules (3 -> ... -> 1), one per level.
    precision: "fp32", "bf16", "fp16", "bf16x3" or "bf16x3f": arithmetic of
        the hidden GEMMs.  "bf16x3" holds every operand as a BF16 pair hi + lo
        and forms hi.hi + hi.lo + lo.hi (levels wider than 256 run FP16); it
        serves evaluation (gradients in forward mode), extraction, tracing and
        rendering, not batched extraction or training.  "bf16x3f" is the same
        with the last of a three-hidden-layer level's two products on the FP16
        matrix cores (FP16 weights against the activations as an FP16 pair:
        two MFMAs per step instead of three, level error near 5e-5); it serves
        evaluation (gradients in forward mode), extraction and tracing, not
        rendering, batched extraction or training.
    """

    def __init__(self, models, precision="fp32"):
        lib = _lib.load()
        self.models = list(models)
        if not 1 <= len(self.models) <= _lib.MP_MAX_LEVELS:
            raise ValueError(f"a stack holds 1..{_lib.MP_MAX_LEVELS} levels, got {len(self.models)}")
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}, got {precision!r}")
        self.precision = precision
        self._layers = [_linear_layers(m) for m in self.models]
        dev = self._layers[0][0][0].device
        for layers in self._layers:
            for w, b, _ in layers:
                _lib.require_gpu_tensor(w, "SIREN weight")
                if w.device != dev:
                    raise ValueError("all levels of a stack must live on the same device")
        self.device = dev
        self._flat_w = []
        self._flat_b = []
        self._descs = (_lib.mp_level_desc * len(self.models))()
        self._hidden_arrays = []
        for i, layers in enumerate(self._layers):
            n_w = sum(w.numel() for w, _, _ in layers)
            n_b = sum(b.numel() for _, b, _ in layers)
            self._flat_w.append(torch.empty(n_w, device=dev, dtype=torch.float32))
            self._flat_b.append(torch.empty(n_b, device=dev, dtype=torch.float32))
            hidden = [w.shape[0] for w, _, _ in layers[:-1]]
            if layers[0][0].shape[1] != 3 or layers[-1][0].shape[0] != 1:
                raise NotImplementedError("the GPU kernels evaluate 3 -> ... -> 1 SIRENs")
            arr = (ctypes.c_int * len(hidden))(*hidden)
            self._hidden_arrays.append(arr)
            d = self._descs[i]
            d.n_in, d.n_out, d.n_hidden = 3, 1, len(hidden)
            d.hidden = arr
            d.weights = self._flat_w[i].data_ptr()
            d.biases = self._flat_b[i].data_ptr()
        self._versions = None
        self._copy_weights()
        handle = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(lib.mp_stack_create(self._descs, len(self.models), _lib.PRECISIONS[precision],
                                           _lib.stream_handle(), ctypes.byref(handle)))
        self._handle = handle
        self._lib = lib

    @property
    def n_levels(self):
        return len(self.models)

    def set_tiles(self, column_blocks, family="extract"):
        """Tile shape of one kernel family (the autotuner's knobs):
        "extract": 1 (one column block per wave, eight-wave workgroups) or 2
        (default); "batch": 1 (each width's default column blocks) or 2 (half
        of them); "trace": 0 (follow "extract"), 1 or 2 column blocks."""
        _lib.check(self._lib.mp_stack_set_kernel_tiles(self._handle, _lib.TILE_FAMILIES[family], int(column_blocks)))

    def _current_versions(self):
        # the frequencies are read from the modules each time: update_omegas()
        # bakes them into the weights and resets them to 1 in place
        return tuple((w._version, b._version, w.data_ptr(), b.data_ptr(), omega)
                     for m in self.models for w, b, omega in _linear_layers(m))

    @torch.no_grad()
    def _copy_weights(self):
        self._layers = [_linear_layers(m) for m in self.models]
        for i, layers in enumerate(self._layers):
            ow = ob = 0
            for w, b, omega in layers:
                nw, nb = w.numel(), b.numel()
                self._flat_w[i][ow:ow + nw].copy_(w.detach().reshape(-1))
                self._flat_b[i][ob:ob + nb].copy_(b.detach().reshape(-1))
                if omega != 1.0:
                    self._flat_w[i][ow:ow + nw].mul_(omega)
                    self._flat_b[i][ob:ob + nb].mul_(omega)
                ow += nw
                ob += nb
        self._versions = self._current_versions()

    def refresh(self):
        """Re-pack if any parameter changed since the last pack."""
        if self._current_versions() != self._versions:
            self._copy_weights()
            _lib.check(self._lib.mp_stack_update(self._handle, _lib.stream_handle(self.device)))

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self._lib.mp_stack_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- entry points -----------------------------------------------------------
    def eval(self, xyz, lvl0=0, lvl1=None, grad=False):
        """Sum of levels [lvl0, lvl1) at points xyz [n, 3]; optionally the input gradient."""
        self.refresh()
        lvl1 = self.n_levels if lvl1 is None else lvl1
        _lib.require_gpu_tensor(xyz, "xyz")
        if xyz.dtype != torch.float32 or xyz.shape[-1] != 3:
            raise ValueError(f"xyz must be float32 [..., 3], got {xyz.dtype} {tuple(xyz.shape)}")
        pts = xyz.reshape(-1, 3).contiguous()
        n = pts.shape[0]
        sdf = torch.empty(n, device=pts.device, dtype=torch.float32)
        g = torch.empty(n, 3, device=pts.device, dtype=torch.float32) if grad else None
        with torch.cuda.device(pts.device):
            _lib.check(self._lib.mp_eval(self._handle, _lib.ptr(pts), n, lvl0, lvl1, _lib.ptr(sdf),
                                         _lib.ptr(g), _lib.stream_handle()))
        sdf = sdf.reshape(xyz.shape[:-1])
        if grad:
            return sdf, g.reshape(xyz.shape)
        return sdf

    def extract_volume(self, N, deltas=(), chunk=32 ** 3, grid="reference", out=None):
        """Band-culled N^3 volume (create_mesh_multistage sampling semantics).

        Returns (vol [N, N, N] float32 on the device, per-level sample counts).
        Synchronises the stream (the band counts are read back).
        """
        self.refresh()
        if grid not in _lib.GRID_MODES:
            raise ValueError(f"grid must be one of {sorted(_lib.GRID_MODES)}, got {grid!r}")
        deltas = [float(d) for d in deltas]
        if deltas and len(deltas) != self.n_levels - 1:
            raise ValueError(f"need 0 or {self.n_levels - 1} deltas, got {len(deltas)}")
        vol = out if out is not None else torch.empty(N, N, N, device=self.device, dtype=torch.float32)
        if vol.numel() != N ** 3 or vol.dtype != torch.float32 or not vol.is_contiguous():
            raise ValueError("out must be a contiguous float32 tensor with N^3 elements")
        d_arr = (ctypes.c_float * max(1, len(deltas)))(*deltas)
        counts = (ctypes.c_int64 * self.n_levels)()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.mp_extract_volume(
                self._handle, int(N), _lib.GRID_MODES[grid], int(chunk),
                ctypes.cast(d_arr, _lib.c_float_p), len(deltas), _lib.ptr(vol),
                ctypes.cast(counts, _lib.c_int64_p), _lib.stream_handle()))
        return vol, [int(c) for c in counts]


def grid_coords(N, grid="reference", first=0, n=None, device="cuda"):
    """Sample coordinates of grid indices [first, first + n) (gen_mc_coordinate_grid)."""
    lib = _lib.load()
    n = N ** 3 - first if n is None else n
    xyz = torch.empty(n, 3, device=device, dtype=torch.float32)
    with torch.cuda.device(xyz.device):
        _lib.check(lib.mp_grid_coords(int(N), _lib.GRID_MODES[grid], int(first), int(n),
                                      _lib.ptr(xyz), _lib.stream_handle()))
    return xyz
