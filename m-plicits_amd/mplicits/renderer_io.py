"""The renderer's on-disk model formats: flat float32 ``.bin`` pairs and the
``.exp`` experiment descriptor.

Writer semantics follow renderer/scripts/export_experiment.py:28-82 (weights
of every layer row-major [out, in] and concatenated in layer order, biases
concatenated in a second file, both headerless little-endian float32; the
descriptor lists ``lod{0,1,2}_{layers,hidden,weights,biases}`` with
``layers`` = the number of hidden width x width GEMMs).  Reader semantics
follow LoadExperimentFile and ReadBinaryFile (renderer/cuda/state.cu:662-712,
renderer/cuda/tools.cu:68-79): ``key = value`` lines, ``#`` comments, unknown
keys reported and ignored, file paths relative to a data root.
"""
import os
import os.path as osp
import sys

import numpy as np
import torch

from .model import SIREN

STAGES = ("coarse", "medium", "fine")

# descriptor keys and their parsers (state.cu:679-700)
_INT_KEYS = {"lod0_layers", "lod0_hidden", "lod1_layers", "lod1_hidden", "lod2_layers", "lod2_hidden",
             "textures_layers", "textures_hidden", "surface_w0", "textures_w0"}
_BOOL_KEYS = {"is3d", "swap_y_and_z", "invert_z", "flip_y"}
_PATH_KEYS = {"lod0_weights", "lod0_biases", "lod1_weights", "lod1_biases", "lod2_weights", "lod2_biases",
              "textures_weights", "textures_biases"}


def flatten_state_dict(sd):
    """(weights, biases, width, hidden GEMMs) of a SIREN state dict, in the
    renderer's order: tensors alternate weight / bias layer by layer."""
    tensors = list(sd.values())
    if len(tensors) % 2:
        raise ValueError("a SIREN state dict alternates weight and bias tensors")
    weights = [tensors[i].detach().cpu().float().numpy().reshape(-1) for i in range(0, len(tensors), 2)]
    biases = [tensors[i + 1].detach().cpu().float().numpy().reshape(-1) for i in range(0, len(tensors), 2)]
    width = int(tensors[0].shape[0])
    n_gemms = sum(1 for t in tensors[2:-2:2] if tuple(t.shape) == (width, width))
    return weights, biases, width, n_gemms


def write_bin(path, arrays):
    np.concatenate([np.asarray(a, dtype="<f4").reshape(-1) for a in arrays]).tofile(path)


def read_bin(path):
    return np.fromfile(path, dtype="<f4")


def export_experiment(levels, out_dir, name, data_root="data", w0=1, flip_y=False, swap_y_and_z=False,
                      invert_z=False):
    """Write ``{name}_{stage}_{weights,biases}.bin`` for the three levels
    (coarse, medium, fine; frequency baked, as checkpoints are saved) and the
    descriptor ``{name}.exp``.  Returns the descriptor path."""
    if len(levels) != 3:
        raise ValueError("the renderer's descriptor holds exactly three levels (coarse, medium, fine)")
    os.makedirs(out_dir, exist_ok=True)
    rel_dir = osp.basename(osp.normpath(out_dir))
    lines = [f"# exported by mplicits.renderer_io from {name}"]
    for lod, (stage, model) in enumerate(zip(STAGES, levels)):
        sd = model.state_dict() if hasattr(model, "state_dict") else model
        w, b, width, n_gemms = flatten_state_dict(sd)
        stem = f"{name}_{stage}"
        write_bin(osp.join(out_dir, stem + "_weights.bin"), w)
        write_bin(osp.join(out_dir, stem + "_biases.bin"), b)
        lines += [f"lod{lod}_layers = {n_gemms}", f"lod{lod}_hidden = {width}",
                  f"lod{lod}_weights = {data_root}/{rel_dir}/{stem}_weights.bin",
                  f"lod{lod}_biases = {data_root}/{rel_dir}/{stem}_biases.bin"]
    lines += ["textures_layers = 0", "is3d = 1", f"surface_w0 = {int(w0)}", "textures_w0 = 0",
              f"swap_y_and_z = {int(bool(swap_y_and_z))}", f"invert_z = {int(bool(invert_z))}",
              f"flip_y = {int(bool(flip_y))}"]
    desc = osp.join(out_dir, name + ".exp")
    with open(desc, "w", encoding="utf-8") as f:
        f.write("\n".join(lines) + "\n")
    return desc


def parse_experiment(path):
    """Descriptor keys -> values (ints, bools, path strings).  Raises
    FileNotFoundError for a missing file and ValueError when lod0..2 weights
    are not all defined."""
    if not osp.exists(path):
        raise FileNotFoundError(f"cannot open experiment file '{path}'")
    e = {"is3d": True, "surface_w0": 1, "textures_w0": 0}
    with open(path, encoding="utf-8") as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("#") or "=" not in line:
                continue
            k, v = (s.strip() for s in line.split("=", 1))
            if k in _INT_KEYS:
                e[k] = int(v)
            elif k in _BOOL_KEYS:
                e[k] = int(v) != 0
            elif k in _PATH_KEYS:
                e[k] = v
            else:
                print(f"experiment file: unknown key '{k}' ignored", file=sys.stderr)
    if not all(e.get(f"lod{i}_weights") for i in range(3)):
        raise ValueError(f"experiment file '{path}' must define lod0, lod1 and lod2 (coarse, medium, fine)")
    return e


def _level_from_bins(weights, biases, width, n_gemms, w0, device):
    """SIREN 3 -> width x (n_gemms + 1) -> 1 filled from the flat arrays."""
    model = SIREN(3, 1, [width] * (n_gemms + 1), w0=w0, ww=w0, delay_init=True)
    wo = bo = 0
    with torch.no_grad():
        for seq in model.net:
            lin = seq[0]
            nw, nb = lin.weight.numel(), lin.bias.numel()
            if wo + nw > weights.size or bo + nb > biases.size:
                raise ValueError(f"checkpoint arrays too short for width {width} with {n_gemms} hidden GEMMs")
            lin.weight.copy_(torch.from_numpy(weights[wo:wo + nw].copy()).view_as(lin.weight))
            lin.bias.copy_(torch.from_numpy(biases[bo:bo + nb].copy()))
            wo += nw
            bo += nb
    if wo != weights.size or bo != biases.size:
        raise ValueError(f"checkpoint arrays hold {weights.size} / {biases.size} values, expected {wo} / {bo}")
    return model.to(device)


def load_experiment(path, data_root=None, device="cpu"):
    """(levels [coarse, medium, fine], descriptor dict) from an ``.exp`` file.
    File paths in the descriptor are joined to ``data_root`` when given (the
    renderer's -data_root), else used as written."""
    e = parse_experiment(path)
    levels = []
    for lod in range(3):
        wf, bf = e[f"lod{lod}_weights"], e[f"lod{lod}_biases"]
        if data_root:
            wf, bf = osp.join(data_root, wf), osp.join(data_root, bf)
        for f in (wf, bf):
            if not osp.exists(f):
                raise FileNotFoundError(f"cannot open checkpoint file '{f}' (use data_root to point at the data folder)")
        levels.append(_level_from_bins(read_bin(wf), read_bin(bf), e[f"lod{lod}_hidden"], e[f"lod{lod}_layers"],
                                       e["surface_w0"], device))
    return levels, e
