"""ctypes binding of libmplicits_hip.so (the C ABI declared in include/mplicits.h).

The library is loaded from this package directory.  There is no fallback: if it
is missing or cannot be loaded, every GPU entry point raises.
"""
import ctypes
import os.path as osp
import threading

import torch

LIB_PATH = osp.join(osp.dirname(osp.abspath(__file__)), "libmplicits_hip.so")

MP_OK = 0
MP_EINVAL = -1
MP_ENOMEM = -2
MP_EHIP = -3
MP_EUNSUPPORTED = -4

MP_FP32 = 0
MP_BF16 = 1
MP_FP16 = 2
MP_BF16X3 = 3  # three-term BF16 (hi.hi + hi.lo + lo.hi); no batched extraction or training
MP_BF16X3F = 4  # three-term BF16 with the last hidden product in FP16 (split activations); no render/batch/train
PRECISIONS = {"fp32": MP_FP32, "bf16": MP_BF16, "fp16": MP_FP16, "bf16x3": MP_BF16X3, "bf16x3f": MP_BF16X3F}

MP_GRID_REFERENCE = 0
MP_GRID_LATTICE = 1
GRID_MODES = {"reference": MP_GRID_REFERENCE, "lattice": MP_GRID_LATTICE}

MP_MAX_LEVELS = 4

MP_TILES_EXTRACT = 0
MP_TILES_BATCH = 1
MP_TILES_TRACE = 2
TILE_FAMILIES = {"extract": MP_TILES_EXTRACT, "batch": MP_TILES_BATCH, "trace": MP_TILES_TRACE}

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int64_p = ctypes.POINTER(ctypes.c_int64)


class mp_level_desc(ctypes.Structure):
    _fields_ = [
        ("n_in", ctypes.c_int),
        ("n_out", ctypes.c_int),
        ("n_hidden", ctypes.c_int),
        ("hidden", ctypes.POINTER(ctypes.c_int)),
        ("weights", ctypes.c_void_p),
        ("biases", ctypes.c_void_p),
    ]


class mp_trace_params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("max_steps", ctypes.c_int),
        ("hit_eps", ctypes.c_float),
        ("bound_radius", ctypes.c_float),
        ("band_scale", ctypes.c_float),
        ("pixel_center", ctypes.c_int),
        ("flip_y", ctypes.c_int),
        ("n_deltas", ctypes.c_int),
        ("deltas", ctypes.c_float * MP_MAX_LEVELS),
    ]


class mp_extract_job(ctypes.Structure):
    _fields_ = [
        ("stack", ctypes.c_void_p),
        ("N", ctypes.c_int),
        ("grid_mode", ctypes.c_int),
        ("chunk", ctypes.c_int64),
        ("n_deltas", ctypes.c_int),
        ("deltas", ctypes.c_float * (MP_MAX_LEVELS - 1)),
        ("vol", ctypes.c_void_p),
    ]


class mp_render_params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("iters", ctypes.c_int * 3),
        ("delta", ctypes.c_float),
        ("threshold", ctypes.c_float),
        ("normal_lod", ctypes.c_int),
        ("shading", ctypes.c_int),
        ("residual", ctypes.c_int),
        ("skip_lod0", ctypes.c_int),
        ("flip_y", ctypes.c_int),
        ("lod2_reads_lod0", ctypes.c_int),
        ("bound_radius", ctypes.c_float),
        ("band_scale", ctypes.c_float),
    ]


MP_SHADE_NORMALS, MP_SHADE_PHONG = 0, 1


class mp_trace_job(ctypes.Structure):
    _fields_ = [
        ("stack", ctypes.c_void_p),
        ("cam", ctypes.c_void_p),
        ("params", mp_trace_params),
        ("depth", ctypes.c_void_p),
        ("normals", ctypes.c_void_p),
        ("rgba", ctypes.c_void_p),
        ("steps", ctypes.c_void_p),
    ]


# Exported symbols: name -> (restype, argtypes).  The CPU tests check that the
# library exports every one of them.
SIGNATURES = {
    "mp_abi_version": (ctypes.c_int, []),
    "mp_last_error": (ctypes.c_char_p, []),
    "mp_stack_create": (ctypes.c_int, [ctypes.POINTER(mp_level_desc), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "mp_stack_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mp_stack_destroy": (None, [ctypes.c_void_p]),
    "mp_stack_levels": (ctypes.c_int, [ctypes.c_void_p]),
    "mp_stack_set_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mp_autotune": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]),
    "mp_stack_set_kernel_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "mp_autotune_kernel": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]),
    "mp_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                               ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mp_extract_volume": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                         c_float_p, ctypes.c_int, ctypes.c_void_p, c_int64_p,
                                         ctypes.c_void_p]),
    "mp_extract_volume_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                                 c_float_p, ctypes.c_int, ctypes.c_void_p, c_int64_p,
                                                 ctypes.c_void_p]),
    "mp_grid_coords": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_void_p, ctypes.c_void_p]),
    "mp_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                c_int64_p, ctypes.c_void_p]),
    "mp_render": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 c_int64_p, ctypes.c_void_p]),
    "mp_trace_batch_plan": (ctypes.c_int, [ctypes.POINTER(mp_trace_job), ctypes.c_int, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "mp_trace_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mp_trace_batch_stats": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, ctypes.c_void_p]),
    "mp_trace_batch_destroy": (None, [ctypes.c_void_p]),
    "mp_train_param_count": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int]),
    "mp_train_fwd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "mp_train_bwd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mp_mesh_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "mp_mesh_destroy": (None, [ctypes.c_void_p]),
    "mp_mesh_info": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, c_int64_p, ctypes.POINTER(ctypes.c_int)]),
    "mp_mesh_distance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "mp_calc_deltas": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]),
    "mp_neighborhood_samples": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_int64, ctypes.c_float, ctypes.c_float,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]),
    "mp_extract_batch_plan": (ctypes.c_int, [ctypes.POINTER(mp_extract_job), ctypes.c_int, ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "mp_extract_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mp_extract_batch_counts": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, ctypes.c_void_p]),
    "mp_extract_batch_destroy": (None, [ctypes.c_void_p]),
    "mp_mc_workspace_size": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mp_mc_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, c_int64_p, c_int64_p, ctypes.c_void_p,
                                   ctypes.c_void_p]),
    "mp_mc_emit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}

_lib = None
_lock = threading.Lock()


class MPError(RuntimeError):
    """A nonzero status from the HIP library."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} (status {code})")
        self.code = code


def load():
    """Load the library (once).  Raises if it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not osp.exists(LIB_PATH):
                raise RuntimeError(
                    f"libmplicits_hip.so not found at {LIB_PATH}; build it with "
                    "`python m-plicits_amd/build_lib.py` (or __graft_entry__.build())")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(code):
    """Map a status code to the exception types of the reference."""
    if code == MP_OK:
        return
    msg = load().mp_last_error().decode(errors="replace")
    if code == MP_EINVAL:
        raise ValueError(msg)
    if code == MP_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if code == MP_ENOMEM:
        raise MemoryError(msg)
    raise MPError(code, msg)


def stream_handle(device=None):
    """hipStream_t of torch's current stream, as a void pointer."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def require_gpu_tensor(t, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise RuntimeError(
            f"{name} must be a tensor on the GPU: the mplicits hot path runs only through "
            "the HIP library (no CPU fallback)")
