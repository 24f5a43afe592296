"""Residual-training loss and parameter gradients in FP64 torch autograd
(oracle; test infrastructure only).

restates loss_add_detail_fine (i3d/loss_functions.py:327-394) with the
gradient of i3d/diff_operators.py:259-265 and the double backward of the
trainers (experiment_scripts/train_sdf_on_neighborhood_fine.py:220-252), on
SIRENs given as FP64 layer lists (oracle.siren.torch_forward)."""
import torch

from . import siren as osiren


def level_value_grad(layers, x, create_graph):
    xr = x.clone().requires_grad_(True)
    f = osiren.torch_forward(layers, xr)
    (g,) = torch.autograd.grad(f.sum(), xr, create_graph=create_graph)
    return f, g


def loss_add_detail_fine(fine, medium, coarse, x, gt_sdf, gt_normals, weights=(3e3, 1e2, 5e1)):
    """Loss dict; fine is a list of (W, b) FP64 tensors requiring grad."""
    f, gf = level_value_grad(fine, x, True)
    r, gr = level_value_grad(medium, x, False)
    c, gc = level_value_grad(coarse, x, False)
    keep = gt_sdf != -1
    sdf = torch.where(keep, (gt_sdf - f - r.detach() - c.detach()) ** 2, torch.zeros_like(f))
    tg = gf + gr.detach() + gc.detach()
    cos = torch.nn.functional.cosine_similarity(gt_normals, tg, dim=-1)[..., None]
    normal = torch.where(keep, 1 - cos, torch.zeros_like(gt_sdf))
    eik = ((tg.norm(dim=-1) - 1.0) ** 2).unsqueeze(-1)
    return {"sdf_constraint": sdf.mean() * weights[0], "normal_constraint": normal.mean() * weights[1],
            "grad_constraint": eik.mean() * weights[2]}


def param_grads(layers, x, loss_of):
    """d loss / d (W, b) for every layer by FP64 autograd."""
    params = [p for wb in layers for p in wb]
    loss = loss_of(layers, x)
    return torch.autograd.grad(loss, params)
