"""Tile-shape autotuning of the level kernels (mp_autotune).

``best_tiles(width, n_hidden, precision)`` returns the measured best number of
point-column blocks per wave for that level shape, cached in a text file
(``$MPLICITS_AUTOTUNE_CACHE`` or ``~/.cache/mplicits/autotune.txt``);
``tune_stack(stack)`` applies the choice of the stack's widest level.
"""
import ctypes
import os
import os.path as osp

from . import _lib


def default_cache():
    return os.environ.get("MPLICITS_AUTOTUNE_CACHE",
                          osp.join(osp.expanduser("~"), ".cache", "mplicits", "autotune.txt"))


def best_tiles(width, n_hidden, precision="bf16", cache=None):
    lib = _lib.load()
    path = default_cache() if cache is None else cache
    if path:
        os.makedirs(osp.dirname(osp.abspath(path)), exist_ok=True)
    out = ctypes.c_int(0)
    _lib.check(lib.mp_autotune(path.encode() if path else None, int(width), int(n_hidden),
                               _lib.PRECISIONS[precision], ctypes.byref(out)))
    return out.value


def tune_stack(stack, cache=None):
    """Measure (or look up) and apply the tile shape of ``stack``'s widest level."""
    from .stack import _linear_layers
    best = None
    for model in stack.models:
        layers = _linear_layers(model)
        width = max(w.shape[0] for w, _, _ in layers[:-1])
        key = (width, len(layers) - 1)
        if best is None or key > best:
            best = key
    cols = best_tiles(best[0], best[1], stack.precision, cache)
    stack.set_tiles(cols)
    return cols
