"""Checkpoint loading as the reference's scripts import it (i3d/util.py:79-142)."""
import os.path as osp

import torch

from .model import SIREN, _widths_from_state_dict, siren_v1_to_v2


def from_pth(path, device="cpu", w0=1, ww=None):
    """SIREN from a weights file; widths inferred from the bias shapes.

    Raises FileNotFoundError if ``path`` does not exist.  A version-1 file is
    converted; the converted weights are written to
    ``path.split(".")[0] + "_v2.pth"`` (i3d/util.py:133-140).
    """
    if not osp.exists(path):
        raise FileNotFoundError(f"Weights file not found at \"{path}\"")
    weights = torch.load(path, map_location=torch.device(device), weights_only=True)
    n_in, n_out, hidden = _widths_from_state_dict(weights)
    model = SIREN(n_in, n_out, hidden, w0=w0, ww=ww, delay_init=True)
    try:
        model.load_state_dict(weights)
    except RuntimeError:
        print("Found weights from old version of SIREN. Converting to v2.")
        new_weights, _ = siren_v1_to_v2(weights, True)
        torch.save(new_weights, path.split(".")[0] + "_v2.pth")
        model.load_state_dict(new_weights)
    return model
