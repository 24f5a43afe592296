"""Adaptive band widths (Eq. 5 of the paper): delta_i = (1 + eps) max_j |F_i(x_j)|
over the input surface points x_j, F_i the partial sum of the levels below
level i (reconstruct.py:18-33, appendix_experiments/extract_deltas.py:33-53).

The partial sums are evaluated by the library's level kernels (one stack, the
levels summed on the device); the maximum is a device reduction per batch.
"""
import torch

from .stack import LevelStack


def partial_sum_maxima(models, points, precision="fp32", batch=1 << 21, count=None):
    """[max_j |F_1(x_j)|, ..., max_j |F_k(x_j)|] for the levels ``models``
    (coarse first) at ``points`` [n, 3] (a CUDA tensor), F_i = f_0 + ... +
    f_{i-1} the sum of the first i levels (the sum that decides level i's
    band); k = ``count`` (default: every level)."""
    pts = points.float().contiguous()  # the level kernels refuse host tensors
    stack = LevelStack(models, precision=precision)
    out = []
    with torch.no_grad():
        for level in range(1, (len(models) if count is None else count) + 1):
            worst = torch.zeros((), device=pts.device)
            for i in range(0, pts.shape[0], batch):
                worst = torch.maximum(worst, stack.eval(pts[i:i + batch], 0, level).abs().max())
            out.append(float(worst))
    return out


def band_widths(models, points, eps=0.3, precision="fp32"):
    """delta_i = (1 + eps) max_j |F_i(x_j)|, one per residual level of
    ``models`` (the last level needs no band of its own sum)."""
    return [(1.0 + eps) * m for m in partial_sum_maxima(models, points, precision, count=len(models) - 1)]
