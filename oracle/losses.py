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


def config_loss(fine, medium, coarse, x, f_t, n_t):
    """The benchmark's training loss on the level sum F = f + r + c (only the
    fine level trainable; r, c and their gradients detached, as
    loss_add_detail_fine treats them):
        mean|F - f_t| + mean||grad F| - 1| + mean(1 - cos(grad F, n_t)).
    f_t [n], n_t [n, 3]; returns a scalar tensor."""
    f, gf = level_value_grad(fine, x, True)
    r, gr = level_value_grad(medium, x, False)
    c, gc = level_value_grad(coarse, x, False)
    F = f + r.detach() + c.detach()
    G = gf + gr.detach() + gc.detach()
    gn = G.norm(dim=-1)
    cos = torch.nn.functional.cosine_similarity(G, n_t, dim=-1)
    return (F - f_t).abs().mean() + (gn - 1).abs().mean() + (1 - cos).mean()


def true_sdf(layers, x, gt_sdf, gt_normals):
    """The coarse trainer's loss (true_sdf, i3d/loss_functions.py:123-163) on one
    level given as FP64 (W, b) tensors: 3e3 mean[f^2 | gt = 0] + 2e2
    mean[(gt - f)^2 | gt != 0] + 1e2 mean[1 - cos(grad f, n) | gt = 0] + 5e1
    mean[(|grad f| - 1)^2] (every mean over all points, zeros elsewhere)."""
    f, g = level_value_grad(layers, x, True)
    on = gt_sdf == 0
    zero = torch.zeros_like(f)
    sdf_on = torch.where(on, f ** 2, zero)
    sdf_off = torch.where(~on, (gt_sdf - f) ** 2, zero)
    cos = torch.nn.functional.cosine_similarity(g, gt_normals, dim=-1)[..., None]
    normal = torch.where(on, 1 - cos, torch.zeros_like(gt_sdf))
    eik = ((g.norm(dim=-1) - 1.0) ** 2).unsqueeze(-1)
    return {"sdf_on_surf": sdf_on.mean() * 3e3, "sdf_off_surf": sdf_off.mean() * 2e2,
            "normal_constraint": normal.mean() * 1e2, "grad_constraint": eik.mean() * 5e1}
