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


def test_config_loss_double_backward_matches_finite_differences():
    # L1 / eikonal / cosine loss on the level sum: parameter gradients through
    # the create_graph input gradient, against central differences (points
    # away from the |F - f_t| and ||grad F| - 1| kinks by construction)
    fine = [(w.requires_grad_(True), b.requires_grad_(True)) for w, b in _layers([3, 16, 16, 1], 5)]
    med = _layers([3, 12, 12, 1], 6)
    coarse = _layers([3, 12, 12, 1], 7)
    g = torch.Generator().manual_seed(8)
    x = torch.rand(64, 3, generator=g, dtype=torch.float64) * 2 - 1
    F0 = sum(olosses.level_value_grad(lv, x, False)[0] for lv in (fine, med, coarse)).detach()
    f_t = (F0.squeeze(-1) + torch.where(torch.rand(64, generator=g, dtype=torch.float64) < 0.5, 0.3, -0.3))
    n = torch.nn.functional.normalize(torch.randn(64, 3, generator=g, dtype=torch.float64), dim=-1)

    def total(layers):
        return olosses.config_loss(layers, med, coarse, x, f_t[:, None], n)

    grads = torch.autograd.grad(total(fine), [p for wb in fine for p in wb])
    h = 1e-6
    rng = np.random.default_rng(1)
    params = [q for wb in fine for q in wb]
    for k, p in enumerate(params):
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


def test_config_loss_closed_form_terms():
    # one level f(x) = a . x + b (a single Linear): grad F = a everywhere, so the
    # loss is |F - f_t| + ||a| - 1| + (1 - cos(a, n)) in closed form
    a = torch.tensor([[0.6, 0.0, 0.8]], dtype=torch.float64) * 2.0  # |a| = 2
    lin = [(a.clone().requires_grad_(True), torch.tensor([0.1], dtype=torch.float64, requires_grad=True))]
    zero = [(torch.zeros(1, 3, dtype=torch.float64), torch.zeros(1, dtype=torch.float64))]
    x = torch.tensor([[0.5, 0.0, 0.0], [0.0, 0.0, 1.0]], dtype=torch.float64)
    f_t = torch.tensor([[0.0], [0.0]], dtype=torch.float64)
    n = torch.tensor([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]], dtype=torch.float64)
    loss = olosses.config_loss(lin, zero, zero, x, f_t, n).item()
    F = [0.6 * 2 * 0.5 + 0.1, 0.8 * 2 + 0.1]
    want = (abs(F[0]) + abs(F[1])) / 2 + 1.0 + ((1 - 0.8) + (1 - 0.6)) / 2
    assert abs(loss - want) < 1e-14


def test_true_sdf_closed_form_terms():
    # f(x) = a . x + b: grad f = a at every point; two on-surface points
    # (gt = 0) and two off-surface points
    a = torch.tensor([[0.0, 0.6, 0.8]], dtype=torch.float64) * 1.5  # |a| = 1.5
    lin = [(a, torch.tensor([0.2], dtype=torch.float64))]
    x = torch.tensor([[0, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, 1]], dtype=torch.float64)
    f = [0.2, 1.1, 0.2, 1.4]
    gt = torch.tensor([[0.0], [0.0], [0.5], [1.0]], dtype=torch.float64)
    n = torch.tensor([[0, 0, 1], [0, 1, 0], [1, 0, 0], [1, 0, 0]], dtype=torch.float64)
    d = olosses.true_sdf(lin, x, gt, n)
    assert abs(d["sdf_on_surf"].item() - 3e3 * (f[0] ** 2 + f[1] ** 2) / 4) < 1e-10
    assert abs(d["sdf_off_surf"].item() - 2e2 * ((0.5 - f[2]) ** 2 + (1.0 - f[3]) ** 2) / 4) < 1e-10
    assert abs(d["normal_constraint"].item() - 1e2 * ((1 - 0.8) + (1 - 0.6)) / 4) < 1e-12
    assert abs(d["grad_constraint"].item() - 5e1 * 0.25) < 1e-12


def test_true_sdf_double_backward_matches_finite_differences():
    lay = [(w.requires_grad_(True), b.requires_grad_(True)) for w, b in _layers([3, 16, 16, 1], 9)]
    g = torch.Generator().manual_seed(10)
    x = torch.rand(48, 3, generator=g, dtype=torch.float64) * 2 - 1
    gt = torch.rand(48, 1, generator=g, dtype=torch.float64) * 0.2 - 0.1
    gt[::3] = 0
    n = torch.nn.functional.normalize(torch.randn(48, 3, generator=g, dtype=torch.float64), dim=-1)
    total = lambda L: sum(olosses.true_sdf(L, x, gt, n).values())  # noqa: E731
    params = [q for wb in lay for q in wb]
    grads = torch.autograd.grad(total(lay), params)
    h = 1e-6
    rng = np.random.default_rng(2)
    for k, p in enumerate(params):
        for idx in rng.choice(p.numel(), min(3, p.numel()), replace=False):
            flat = p.detach().view(-1)
            old = flat[idx].item()
            flat[idx] = old + h
            lp = total(lay).item()
            flat[idx] = old - h
            lm = total(lay).item()
            flat[idx] = old
            an = grads[k].reshape(-1)[idx].item()
            assert abs((lp - lm) / (2 * h) - an) <= 1e-5 * max(1.0, abs(an)), (k, idx)
