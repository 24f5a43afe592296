"""SIREN model, initialisation, frequency baking and checkpoint loading.

Drop-in for the reference's ``i3d.model`` (i3d/model.py).  The module keeps the
reference's parameter layout (``net.{k}.0.weight`` [out, in] and
``net.{k}.0.bias``), its initialisation and RNG call order, and its checkpoint
semantics, so checkpoints move between the two unchanged.  ``forward`` runs on
the HIP library: tensors on the GPU go through the fused level kernels (or the
hand-written forward/backward when gradients are required); tensors on the CPU
raise, because there is no CPU path in this package.
"""
from collections import OrderedDict
import math
import os.path as osp

import torch
from torch import nn


@torch.no_grad()
def sine_init(m, w0):
    """U(-sqrt(6/fan_in)/w0, +sqrt(6/fan_in)/w0) for a Linear weight (i3d/model.py:10-15)."""
    if hasattr(m, "weight"):
        fan_in = m.weight.size(-1)
        bound = math.sqrt(6 / fan_in) / w0
        m.weight.uniform_(-bound, bound)


@torch.no_grad()
def first_layer_sine_init(m):
    """U(-1/fan_in, 1/fan_in) for the first Linear weight (i3d/model.py:18-22)."""
    if hasattr(m, "weight"):
        fan_in = m.weight.size(-1)
        m.weight.uniform_(-1 / fan_in, 1 / fan_in)


class SineLayer(nn.Module):
    """x -> sin(w0 * x) (i3d/model.py:25-36)."""

    def __init__(self, w0=30):
        super().__init__()
        self.w0 = w0

    def forward(self, x):
        raise RuntimeError("SineLayer is evaluated inside the fused SIREN kernels only")

    def __repr__(self):
        return f"SineLayer(w0={self.w0})"


class SIREN(nn.Module):
    """SIREN with ``len(hidden_layer_config)`` sine layers (i3d/model.py:39-180).

    Parameters follow the reference: ``n_in_features``, ``n_out_features``,
    ``hidden_layer_config``, ``w0`` (first-layer frequency), ``ww`` (hidden
    frequency, defaults to w0) and ``delay_init``.  The frequencies are applied
    by the kernels as ``sin(w * (W x + b))``; checkpoints saved after
    ``update_omegas(1)`` carry them baked into the weights.

    Extra attribute ``precision`` ("fp32", "bf16", "fp16", "bf16x3" or "bf16x3f")
    selects the arithmetic of the hidden-layer GEMMs on the GPU.
    """

    def __init__(self, n_in_features, n_out_features, hidden_layer_config=[], w0=30, ww=None,
                 delay_init=False):
        super().__init__()
        if not hidden_layer_config:
            raise ValueError("hidden_layer_config must name at least one hidden layer")
        self.in_features = n_in_features
        self.out_features = n_out_features
        self.w0 = w0
        self.ww = w0 if ww is None else ww
        self.precision = "fp32"
        # The Linear layers are constructed in order, so their default
        # (kaiming weight, uniform bias) draws consume the RNG exactly as the
        # reference constructor does.
        widths = [n_in_features] + list(hidden_layer_config)
        layers = []
        for k in range(len(hidden_layer_config)):
            layers.append(nn.Sequential(nn.Linear(widths[k], widths[k + 1]),
                                        SineLayer(self.w0 if k == 0 else self.ww)))
        layers.append(nn.Sequential(nn.Linear(widths[-1], n_out_features)))
        self.net = nn.Sequential(*layers)
        if not delay_init:
            self.reset_weights()

    # -- reference API ------------------------------------------------------
    def forward(self, x, preserve_grad=False):
        """Returns ``{"model_in": coords, "model_out": y}`` (i3d/model.py:107-129).

        ``coords`` is ``x`` itself when ``preserve_grad`` is true, else a detached
        clone that requires grad (so that ``gradient(model_out, model_in)``
        works as in the reference).
        """
        coords = x if preserve_grad else x.clone().detach().requires_grad_(True)
        from . import fused  # deferred: needs the HIP library
        y = fused.siren_forward(self, coords)
        return {"model_in": coords, "model_out": y}

    def reset_weights(self):
        """Sitzmann et al. initialisation (i3d/model.py:131-141)."""
        self.net[0].apply(first_layer_sine_init)
        self.net[1:].apply(lambda m: sine_init(m, self.ww))
        return self

    def update_omegas(self, w0=1, ww=None):
        """Bake the frequency into the weights (i3d/model.py:143-180).

        Every state_dict tensor except the last two is multiplied by
        ``self.w0 / w0`` (the reference uses w0 for the hidden layers too, which
        is exact when w0 == ww, as in every shipped configuration).
        """
        if ww is None:
            ww = w0
        sd = self.state_dict()
        keys = list(sd.keys())
        factor = self.w0 / w0
        for k in keys[:-2]:
            sd[k] = sd[k] * factor
        self.load_state_dict(sd)
        self.w0 = w0
        self.ww = ww
        self.net[0][1].w0 = w0
        for i in range(1, len(self.net) - 1):
            self.net[i][1].w0 = ww
        return self

    # -- helpers used by the kernels ------------------------------------------
    def hidden_widths(self):
        return [seq[0].out_features for seq in self.net[:-1]]

    def omegas(self):
        """Per sine layer frequency (w0 for layer 0, ww for the rest)."""
        return [seq[1].w0 for seq in self.net[:-1]]

    def __repr__(self):
        return (f"SIREN(in={self.in_features}, out={self.out_features}, "
                f"hidden={self.hidden_widths()}, w0={self.w0}, ww={self.ww})")


def siren_v1_to_v2(model_in, check_equals=False):
    """Strip the 4-character key prefix of version-1 checkpoints (i3d/model.py:282-319)."""
    model_out = OrderedDict((k[4:], v) for k, v in model_in.items())
    divergences = []
    if check_equals:
        for k in model_in.keys():
            if not torch.equal(model_in[k], model_out[k[4:]]):
                divergences.append((k, k[4:]))
    return model_out, divergences


def _widths_from_state_dict(weights):
    """Topology from the bias shapes, as i3d/model.py:349-361 infers it."""
    n_layers = len(weights) // 2
    keys = list(weights.keys())
    bias_keys = [k for k in keys if "bias" in k]
    hidden = [weights[bias_keys[i]].shape[0] for i in range(n_layers - 1)]
    n_in = weights[keys[0]].shape[1]
    n_out = weights[keys[-1]].shape[0]
    return n_in, n_out, hidden


def from_state_dict(weights, device="cpu", w0=1, ww=None):
    """SIREN mirroring ``weights`` (i3d/model.py:322-380).  Returns (model, upgraded_to_v2)."""
    n_in, n_out, hidden = _widths_from_state_dict(weights)
    model = SIREN(n_in, n_out, hidden, w0=w0, ww=ww, delay_init=True)
    upgraded = False
    try:
        model.load_state_dict(weights)
    except RuntimeError:
        print("Found weights from old version of SIREN. Converting to v2.")
        new_weights, _ = siren_v1_to_v2(weights, True)
        model.load_state_dict(new_weights)
        upgraded = True
    return model, upgraded


def from_pth(path, device="cpu", w0=1, ww=None):
    """SIREN from a weights file (i3d/model.py:383-424).

    Raises FileNotFoundError if ``path`` does not exist.  A version-1 file is
    converted and the converted weights are saved next to it as
    ``path.split(".")[0] + "_v2.pth"``.
    """
    if not osp.exists(path):
        raise FileNotFoundError(f"Weights file not found at \"{path}\"")
    weights = torch.load(path, map_location=torch.device(device), weights_only=True)
    model, upgraded = from_state_dict(weights, device=device, w0=w0, ww=ww)
    if upgraded:
        torch.save(model.state_dict(), path.split(".")[0] + "_v2.pth")
    return model.to(device=device)
