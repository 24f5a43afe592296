"""Pins the FP64 loss oracle (oracle/losses.py): its parameter gradients,
obtained by double backward through autograd, against central finite
differences of the loss itself."""
import numpy as np
import torch

from oracle import losses as olosses


def _layers(widths, seed):
    g = torch.Generator().manual_seed(seed)
    out = []
    for a, b in zip(widths[:-1], widths[1:]):
        w = (torch.rand(b, a, generator=g, dtype=torch.float64) * 2 - 1) * (3.0 / a ** 0.5)
        bias = (torch.rand(b, generator=g, dtype=torch.float64) * 2 - 1) * 0.5
        out.append((w, bias))
    return out


def test_double_backward_matches_finite_differences():
    fine = [(w.requires_grad_(True), b.requires_grad_(True)) for w, b in _layers([3, 16, 16, 1], 1)]
    med = _layers([3, 12, 12, 1], 2)
    coarse = _layers([3, 12, 12, 1], 3)
    g = torch.Generator().manual_seed(4)
    x = torch.rand(64, 3, generator=g, dtype=torch.float64) * 2 - 1
    gt = torch.rand(64, 1, generator=g, dtype=torch.float64) * 0.2 - 0.1
    gt[::4] = -1
    n = torch.nn.functional.normalize(torch.randn(64, 3, generator=g, dtype=torch.float64), dim=-1)

    def total(layers):
        d = olosses.loss_add_detail_fine(layers, med, coarse, x, gt, n)
        return sum(d.values())

    grads = torch.autograd.grad(total(fine), [p for wb in fine for p in wb])
    h = 1e-6
    rng = np.random.default_rng(0)
    for k in range(len(grads)):
        p = [q for wb in fine for q in wb][k]
        for idx in rng.choice(p.numel(), min(3, p.numel()), replace=False):
            flat = p.detach().view(-1)
            old = flat[idx].item()
            flat[idx] = old + h
            lp = total(fine).item()
            flat[idx] = old - h
            lm = total(fine).item()
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            an = grads[k].reshape(-1)[idx].item()
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (k, idx, fd, an)
