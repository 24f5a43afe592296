"""SIREN forward on the HIP library: the entry behind ``SIREN.forward``.

Without gradients the model is evaluated by the fused level kernel.  With
gradients enabled the call goes through the hand-written forward/backward
behind ``torch.autograd.Function`` (module ``mplicits.training``).
"""
import torch

from . import _lib
from .stack import LevelStack


def model_stack(model, precision=None):
    """The single-level LevelStack of ``model`` (cached on the module)."""
    precision = precision or getattr(model, "precision", "fp32")
    dev = next(model.parameters()).device
    cache = model.__dict__.setdefault("_mp_stacks", {})
    key = (precision, str(dev))
    st = cache.get(key)
    if st is None:
        st = LevelStack([model], precision=precision)
        cache[key] = st
    return st


def siren_forward(model, coords):
    """model_out [..., 1] of ``model`` at ``coords`` [..., 3] on the GPU."""
    _lib.require_gpu_tensor(coords, "SIREN input")
    if coords.shape[-1] != 3:
        raise ValueError(f"SIREN input must have 3 features, got shape {tuple(coords.shape)}")
    needs_graph = torch.is_grad_enabled() and (
        coords.requires_grad or any(p.requires_grad for p in model.parameters()))
    if needs_graph:
        from . import training
        return training.siren_apply(model, coords)
    x = coords.detach()
    if x.dtype != torch.float32:
        x = x.float()
    sdf = model_stack(model).eval(x)
    return sdf.unsqueeze(-1)
