"""SDF training losses: drop-in for the parts of i3d/loss_functions.py used by
the residual trainers.  Same names, arguments, weights and returned dicts; the
gradients come from mplicits.diff_operators.gradient, i.e. from the
hand-written kernels when the models are this package's SIRENs."""
from copy import deepcopy

import torch
import torch.nn.functional as F

from . import diff_operators


def sdf_constraint_on_surf(gt_sdf, pred_sdf):
    """pred^2 where gt == 0 (i3d/loss_functions.py:9-10)."""
    return torch.where(gt_sdf == 0, pred_sdf ** 2, torch.zeros_like(pred_sdf))


def sdf_constraint_off_surf(gt_sdf, pred_sdf):
    """(gt - pred)^2 where gt != 0 (i3d/loss_functions.py:13-16)."""
    return torch.where(gt_sdf != 0, (gt_sdf - pred_sdf) ** 2, torch.zeros_like(pred_sdf))


def vector_aligment_on_surf(gt_sdf, gt_vectors, pred_vectors):
    """1 - cos(pred, gt) where gt == 0 (i3d/loss_functions.py:25-30)."""
    return torch.where(gt_sdf == 0,
                       1 - F.cosine_similarity(pred_vectors, gt_vectors, dim=-1)[..., None],
                       torch.zeros_like(gt_sdf))


def eikonal_constraint(gradient):
    """(|grad| - 1)^2 (i3d/loss_functions.py:41-42)."""
    return (gradient.norm(dim=-1) - 1.0) ** 2


def on_surface_normal_constraint(gt_sdf, gt_normals, grad):
    """1 - cos(grad, n) where gt == 0 (i3d/loss_functions.py:57-66)."""
    return torch.where(gt_sdf == 0, 1 - F.cosine_similarity(grad, gt_normals, dim=-1)[..., None],
                       torch.zeros_like(grad[..., :1]))


def true_sdf(X, gt):
    """Loss of the coarse trainer (i3d/loss_functions.py:123-163)."""
    gt_sdf, gt_normals = gt["sdf"], gt["normals"]
    coords, pred_sdf = X["model_in"], X["model_out"]
    grad = diff_operators.gradient(pred_sdf, coords)
    return {
        "sdf_on_surf": sdf_constraint_on_surf(gt_sdf, pred_sdf).mean() * 3e3,
        "sdf_off_surf": sdf_constraint_off_surf(gt_sdf, pred_sdf).mean() * 2e2,
        "normal_constraint": vector_aligment_on_surf(gt_sdf, gt_normals, grad).mean() * 1e2,
        "grad_constraint": eikonal_constraint(grad).unsqueeze(-1).mean() * 5e1,
    }


class loss_add_detail_fine(torch.nn.Module):
    """Loss of the fine residual level on top of two frozen levels
    (i3d/loss_functions.py:327-394): with S = f + r + c the level sum,
    3e3 mean[(gt - S)^2 | gt != -1] + 1e2 mean[1 - cos(n_gt, grad S) | gt != -1]
    + 5e1 mean[(|grad S| - 1)^2]; grad r and grad c are detached."""

    DEFAULT_WEIGHTS = {"sdf_constraint": 3e3, "normal_constraint": 1e2, "grad_constraint": 5e1}

    def __init__(self, model_coarse, res_model, constraint_weights: dict = dict()):
        super().__init__()
        self.model_coarse = model_coarse.eval()
        self.res_model = res_model.eval()
        if not constraint_weights:
            self.constraint_weights = deepcopy(loss_add_detail_fine.DEFAULT_WEIGHTS)
        else:
            self.constraint_weights = {k: float(v) for k, v in constraint_weights.items()}

    def forward(self, X, gt):
        gt_sdf, gt_normals = gt["sdf"], gt["normals"]
        coords, pred_sdf = X["model_in"], X["model_out"]
        grad = diff_operators.gradient(pred_sdf, coords)
        coarse = self.model_coarse(coords)
        coarse_sdf = coarse["model_out"]
        coarse_grad = diff_operators.gradient(coarse_sdf, coarse["model_in"]).detach()
        res = self.res_model(coords)
        res_sdf = res["model_out"]
        res_grad = diff_operators.gradient(res_sdf, res["model_in"]).detach()
        keep = gt_sdf != -1
        sdf = torch.where(keep, (gt_sdf - pred_sdf - res_sdf - coarse_sdf) ** 2, torch.zeros_like(pred_sdf))
        total_grad = grad + res_grad + coarse_grad
        normal = torch.where(keep, 1 - F.cosine_similarity(gt_normals, total_grad, dim=-1)[..., None],
                             torch.zeros_like(gt_sdf))
        eik = eikonal_constraint(total_grad).unsqueeze(-1)
        w = self.constraint_weights
        return {"sdf_constraint": sdf.mean() * w["sdf_constraint"],
                "normal_constraint": normal.mean() * w["normal_constraint"],
                "grad_constraint": eik.mean() * w["grad_constraint"]}
