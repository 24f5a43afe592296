"""Hand-written forward / input-gradient / backward (mplicits.training) against
FP64 autograd of the restated loss (oracle/losses.py)."""
import numpy as np
import pytest
import torch

from mplicits import loss_functions, training
from mplicits.stack import LevelStack
from mplicits.synthetic import seeded_siren
from oracle import losses as olosses
from oracle import siren as osiren

pytestmark = pytest.mark.gpu


def _torch_layers(model, requires_grad=False):
    return [(torch.from_numpy(w).requires_grad_(requires_grad), torch.from_numpy(b).requires_grad_(requires_grad))
            for w, b in osiren.layers_from_model(model)]


def _batch(n, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(n, 3, generator=g) * 2 - 1
    gt = torch.rand(n, 1, generator=g) * 0.1 - 0.05
    gt[::3] = -1
    nrm = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=-1)
    return x, gt, nrm


def _models(fine_hidden=(64, 64, 64)):
    coarse = seeded_siren([64] * 3, 30.0, seed=11, last_scale=5.0).cuda()
    med = seeded_siren([64] * 3, 60.0, seed=12, last_scale=0.1).cuda()
    torch.manual_seed(13)
    from mplicits.model import SIREN
    fine = SIREN(3, 1, list(fine_hidden), w0=90, ww=90).cuda()  # unbaked: exercises the omega chain rule
    with torch.no_grad():
        fine.net[-1][0].weight.mul_(0.1)
    return coarse, med, fine


def _oracle_grads(coarse, med, fine, x, gt, nrm):
    # the oracle's layers carry omega baked in; gradients w.r.t. the module's
    # (unbaked) parameters follow by the chain rule d/dW = omega d/dW_eff
    fl = _torch_layers(fine, True)
    loss = sum(olosses.loss_add_detail_fine(fl, _torch_layers(med), _torch_layers(coarse), x.double(), gt.double(),
                                            nrm.double()).values())
    grads = torch.autograd.grad(loss, [p for wb in fl for p in wb])
    oms = [float(seq[1].w0) if k < len(fine.net) - 1 else 1.0 for k, seq in enumerate(fine.net)]
    return loss.item(), [g * oms[i // 2] for i, g in enumerate(grads)]


@pytest.mark.parametrize("precision,tol,fine_hidden", [("fp32", 2e-3, (64, 64, 64)), ("bf16", 6e-2, (64, 64, 64)),
                                                       ("fp32", 2e-3, (50, 30, 50))])
def test_dropin_residual_loss_gradients(precision, tol, fine_hidden):
    # (50, 30, 50): widths that are no multiple of 4 take the scalar kernels
    coarse, med, fine = _models(fine_hidden)
    for m in (coarse, med, fine):
        m.precision = precision
    x, gt, nrm = _batch(2048, 1)
    ref_loss, ref = _oracle_grads(coarse, med, fine, x, gt, nrm)
    loss_fn = loss_functions.loss_add_detail_fine(coarse, med)
    out = fine(x.cuda())
    losses = loss_fn(out, {"sdf": gt.cuda(), "normals": nrm.cuda()})
    loss = sum(losses.values())
    fine.zero_grad()
    loss.backward()
    assert abs(loss.item() - ref_loss) <= 1e-3 * abs(ref_loss) + (1e-2 if precision != "fp32" else 0)
    for p, r in zip(fine.parameters(), ref):
        g = p.grad.double().cpu()
        err = (g - r).norm() / r.norm()
        assert err < tol, (tuple(p.shape), float(err))


def test_fused_step_equals_dropin_path():
    coarse, med, fine = _models((48, 48, 48))
    x, gt, nrm = _batch(1000, 2)
    x, gt, nrm = x.cuda(), gt.cuda().squeeze(-1), nrm.cuda()
    frozen = LevelStack([coarse, med], "fp32")

    def loss_fn(F, G):
        return training.config_loss(F, G, gt, nrm)

    fine.zero_grad()
    loss_a, F, G = training.train_step(fine, x, loss_fn, frozen=frozen)
    ga = [p.grad.clone() for p in fine.parameters()]
    # the same loss through autograd of the drop-in modules
    fine.zero_grad()
    xr = x.clone().requires_grad_(True)
    f = fine(xr, preserve_grad=True)["model_out"].squeeze(-1)
    gf = torch.autograd.grad(f.sum(), xr, create_graph=True)[0]
    fc, gc = frozen.eval(x, grad=True)
    loss_b = training.config_loss(f + fc, gf + gc, gt, nrm)
    loss_b.backward()
    assert abs(loss_a.item() - loss_b.item()) < 1e-5
    for a, p in zip(ga, fine.parameters()):
        assert torch.allclose(a, p.grad, rtol=1e-3, atol=1e-6 * a.abs().max().item())


def test_forward_tangents_are_the_input_gradient():
    _, _, fine = _models((64, 32, 64))
    x = (torch.rand(777, 3) * 2 - 1).cuda()
    layers = training.baked_layers(fine)
    f, fd, _ = training.forward_tangents(layers, x, training.unit_tangents(777, x.device))
    fr, gr = osiren.forward_grad(osiren.layers_from_model(fine), x.double().cpu().numpy())
    assert np.abs(f.double().cpu().numpy() - fr).max() < 1e-5 * max(1, np.abs(fr).max())
    assert np.abs(fd.double().cpu().numpy() - gr).max() < 1e-4 * np.abs(gr).max()


def _blocked(t):
    """[rows, W] -> K-blocked copy: element (r, u) at ((r / 8) * W + u) * 8 + r % 8."""
    rows, w = t.shape
    return t.view(rows // 8, 8, w).transpose(1, 2).contiguous().view(-1)


@pytest.mark.parametrize("dtype,code", [(torch.bfloat16, 1), (torch.float16, 2)])
@pytest.mark.parametrize("rows,wout,win", [(64, 256, 256), (12224, 512, 256),
                                           (1 << 16, 256, 512)])
def test_split_k_weight_gradient_matches_fp32_gemm(dtype, code, rows, wout, win):
    from mplicits import _lib
    lib = _lib.load()
    g = torch.Generator(device="cuda").manual_seed(rows + wout)
    A = torch.randn(rows, wout, device="cuda", generator=g).to(dtype)
    X = torch.randn(rows, win, device="cuda", generator=g).to(dtype)
    chunks = int(lib.mp_train_dw_chunks(rows, wout, win))
    part = torch.full((chunks, wout, win), float("nan"), device="cuda")
    Ab, Xb = _blocked(A), _blocked(X)  # held until the kernel has run
    _lib.check(lib.mp_train_dw(_lib.ptr(Ab), _lib.ptr(Xb), rows, wout, win, code,
                               _lib.ptr(part), _lib.stream_handle()))
    torch.cuda.synchronize()
    got = part.sum(0)
    ref = A.float().t() @ X.float()
    # same products, FP32 accumulation in a different order
    bound = 1e-4 * (A.float().abs().t() @ X.float().abs()) + 1e-6
    assert torch.all((got - ref).abs() <= bound), float(((got - ref).abs() / bound).max())


def test_split_k_weight_gradient_rejects_ragged_rows():
    from mplicits import _lib
    lib = _lib.load()
    A = torch.zeros(96, 256, device="cuda", dtype=torch.bfloat16)
    part = torch.empty(int(lib.mp_train_dw_chunks(96, 256, 256)), 256, 256, device="cuda")
    with pytest.raises(ValueError, match="multiple of 64"):
        _lib.check(lib.mp_train_dw(_lib.ptr(A), _lib.ptr(A), 96, 256, 256, 1, _lib.ptr(part), _lib.stream_handle()))


def test_wide_level_gradients_take_the_blocked_path():
    coarse, med, fine = _models((256, 256, 256))
    x, gt, nrm = _batch(2048, 3)
    assert training._blocked_dw("bf16", 2048, 3, 256, 256)
    ref_loss, ref = _oracle_grads(coarse, med, fine, x, gt, nrm)
    fine.precision = "bf16"
    for m in (coarse, med):
        m.precision = "bf16"
    loss_fn = loss_functions.loss_add_detail_fine(coarse, med)
    losses = loss_fn(fine(x.cuda()), {"sdf": gt.cuda(), "normals": nrm.cuda()})
    loss = sum(losses.values())
    fine.zero_grad()
    loss.backward()
    for p, r in zip(fine.parameters(), ref):
        err = (p.grad.double().cpu() - r).norm() / r.norm()
        assert err < 6e-2, (tuple(p.shape), float(err))
