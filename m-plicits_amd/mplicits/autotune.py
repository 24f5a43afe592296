"""Tile-shape autotuning of the level kernels (mp_autotune_kernel).

Each kernel family has a tile knob (``LevelStack.set_tiles(..., family)``):
``"extract"`` (the level kernels of evaluation and extraction: one column block
per wave in eight-wave workgroups, or the default tiles), ``"batch"`` (the
batched extraction kernel: each width's default column blocks, or half of
them) and ``"trace"`` (the tracers: one or two column blocks per wave).
``best_tiles(width, n_hidden, precision, family=...)`` returns the measured
best choice for that level shape, cached in a text file
(``$MPLICITS_AUTOTUNE_CACHE`` or ``~/.cache/mplicits/autotune.txt``);
``tune_stack(stack)`` measures (or looks up) and applies the choice of the
stack's widest level for every family.
"""
import ctypes
import os
import os.path as osp

from . import _lib

FAMILIES = ("extract", "batch", "trace")


def default_cache():
    return os.environ.get("MPLICITS_AUTOTUNE_CACHE",
                          osp.join(osp.expanduser("~"), ".cache", "mplicits", "autotune.txt"))


def best_tiles(width, n_hidden, precision="bf16", cache=None, family="extract"):
    lib = _lib.load()
    path = default_cache() if cache is None else cache
    if path:
        os.makedirs(osp.dirname(osp.abspath(path)), exist_ok=True)
    out = ctypes.c_int(0)
    _lib.check(lib.mp_autotune_kernel(path.encode() if path else None, _lib.TILE_FAMILIES[family], int(width),
                                      int(n_hidden), _lib.PRECISIONS[precision], ctypes.byref(out)))
    return out.value


def tune_stack(stack, cache=None, families=("extract",)):
    """Measure (or look up) and apply the tile shape of ``stack``'s widest
    level for each of ``families``; returns {family: tiles} (one value when a
    single family is asked for)."""
    from .stack import _linear_layers
    best = None
    for model in stack.models:
        layers = _linear_layers(model)
        width = max(w.shape[0] for w, _, _ in layers[:-1])
        key = (width, len(layers) - 1)
        if best is None or key > best:
            best = key
    if isinstance(families, str):
        families = (families,)
    out = {}
    for fam in families:
        out[fam] = best_tiles(best[0], best[1], stack.precision, cache, family=fam)
        stack.set_tiles(out[fam], family=fam)
    return out[families[0]] if len(families) == 1 else out
