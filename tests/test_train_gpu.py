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


@pytest.mark.parametrize("precision,tol,fine_hidden", [("fp32", 1e-4, (64, 64, 64)), ("bf16", 6e-2, (64, 64, 64)),
                                                       ("fp32", 1e-4, (50, 30, 50))])
def test_dropin_residual_loss_gradients(precision, tol, fine_hidden):
    # (50, 30, 50): ragged widths (padded to whole GEMM tiles)
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


def test_level_forward_is_the_input_gradient():
    _, _, fine = _models((64, 32, 64))
    x = (torch.rand(777, 3) * 2 - 1).cuda()
    f, fd = training.level_forward(fine, x)
    fr, gr = osiren.forward_grad(osiren.layers_from_model(fine), x.double().cpu().numpy())
    assert np.abs(f.double().cpu().numpy() - fr).max() < 1e-5 * max(1, np.abs(fr).max())
    assert np.abs(fd.double().cpu().numpy() - gr).max() < 1e-4 * np.abs(gr).max()


def _abi_train(st, x, s, g, keep):
    """mp_train_fwd / mp_train_bwd called through the C ABI."""
    from mplicits import _lib
    lib = _lib.load()
    n = x.shape[0]
    f = torch.empty(n, device="cuda")
    fd = torch.empty(n, 3, device="cuda")
    print(f"mp_train_fwd/bwd: n={n} keep={keep}", flush=True)
    _lib.check(lib.mp_train_fwd(st._handle, 0, _lib.ptr(x), n, _lib.ptr(f), _lib.ptr(fd), keep, _lib.stream_handle()))
    dp = torch.full((int(lib.mp_train_param_count(st._handle, 0)),), float("nan"), device="cuda")
    _lib.check(lib.mp_train_bwd(st._handle, 0, _lib.ptr(x), n, _lib.ptr(s), _lib.ptr(g), _lib.ptr(dp),
                                _lib.stream_handle()))
    torch.cuda.synchronize()
    return f, fd, dp


@pytest.mark.parametrize("hidden,precision,tol", [((512, 512, 512), "fp32", 1e-4), ((100, 200, 50), "fp32", 1e-4),
                                                  ((256, 256), "fp32", 1e-4), ((64, 64, 64, 64), "fp32", 1e-4),
                                                  ((512, 512, 512), "bf16", 3e-2), ((256, 256, 256), "fp16", 1e-2)])
def test_train_entry_points_against_fp64_autograd(hidden, precision, tol):
    """f, grad f and the parameter gradients of sum [s f + g . grad f] (the
    seeds of any loss on f and grad f) from the C ABI against FP64 autograd
    through the oracle's SIREN, with the create_graph gradient: max error per
    parameter tensor relative to its max."""
    m = seeded_siren(list(hidden), 90.0, seed=70 + len(hidden), last_scale=0.1).cuda()
    st = LevelStack([m], precision)
    st.refresh()
    gen = torch.Generator().manual_seed(5)
    n = 3001  # not a multiple of the point padding
    x = (torch.rand(n, 3, generator=gen) * 2 - 1).cuda()
    s = torch.randn(n, generator=gen).cuda()
    g = torch.randn(n, 3, generator=gen).cuda()
    f, fd, dp = _abi_train(st, x, s, g, keep=1)
    lay = [(torch.from_numpy(w).requires_grad_(True), torch.from_numpy(b).requires_grad_(True))
           for w, b in osiren.layers_from_model(m)]
    fr, gr = olosses.level_value_grad(lay, x.double().cpu(), True)
    obj = (s.double().cpu()[:, None] * fr).sum() + (g.double().cpu() * gr).sum()
    ref = torch.autograd.grad(obj, [w for w, _ in lay] + [b for _, b in lay])
    ftol = 1e-5 if precision == "fp32" else 3e-3
    assert (f.double().cpu() - fr[:, 0].detach()).abs().max() <= ftol * max(1.0, fr.abs().max().item())
    assert (fd.double().cpu() - gr.detach()).abs().max() <= 10 * ftol * gr.abs().max().item()
    assert not torch.isnan(dp).any()
    off = 0
    for r in ref:
        got = dp[off:off + r.numel()].double().cpu().view_as(r)
        off += r.numel()
        err = ((got - r).abs().max() / r.abs().max()).item()
        assert err <= tol, (tuple(r.shape), err)
    assert off == dp.numel()
    # recomputed activations (no keep) give the same gradients, bit for bit
    _, _, dp2 = _abi_train(st, x, s, g, keep=0)
    assert torch.equal(dp, dp2)


@pytest.mark.parametrize("precision,tol", [("fp16", 1.5e-3), ("bf16", 2e-2), ("fp32", 1e-5)])
def test_train_gradients_with_mean_loss_seeds(precision, tol):
    """Seeds of a mean loss over many points are O(1/n) (here 2^-16 .. 2^-14):
    FP16 adjoints would flush to zero or keep few bits without the backward's
    seed scaling.  The parameter gradients must stay within the precision's
    bound of FP64 autograd, and in FP16 (where the seeds are normalised by a
    power of two) seeds x 2^5 give exactly 2^5 x the gradients."""
    m = seeded_siren([256] * 3, 90.0, seed=71, last_scale=0.1).cuda()
    st = LevelStack([m], precision)
    st.refresh()
    gen = torch.Generator().manual_seed(6)
    n = 5003
    x = (torch.rand(n, 3, generator=gen) * 2 - 1).cuda()
    s = (torch.randn(n, generator=gen) * 2.0 ** -15).cuda()
    g = (torch.randn(n, 3, generator=gen) * 2.0 ** -16).cuda()
    _, _, dp = _abi_train(st, x, s, g, keep=1)
    lay = [(torch.from_numpy(w).requires_grad_(True), torch.from_numpy(b).requires_grad_(True))
           for w, b in osiren.layers_from_model(m)]
    fr, gr = olosses.level_value_grad(lay, x.double().cpu(), True)
    obj = (s.double().cpu()[:, None] * fr).sum() + (g.double().cpu() * gr).sum()
    ref = torch.autograd.grad(obj, [w for w, _ in lay] + [b for _, b in lay])
    assert torch.isfinite(dp).all()
    off = 0
    for r in ref:
        got = dp[off:off + r.numel()].double().cpu().view_as(r)
        off += r.numel()
        err = ((got - r).abs().max() / r.abs().max()).item()
        print(f"{precision} {tuple(r.shape)}: max rel err {err:.2e}")
        assert err <= tol, (tuple(r.shape), err)
    if precision == "fp16":
        _, _, dp32 = _abi_train(st, x, s * 32, g * 32, keep=1)
        assert torch.equal(dp32, dp * 32)


@pytest.mark.parametrize("last_scale", [1.0, 5.0])
def test_fp16_seed_scaling_headroom(last_scale):
    """The FP16 backward brings max |seed| into [2^10, 2^11); the adjoints are
    the seeds times the output weights, so large output weights (x1, x5 the
    initialisation's) and a seed far above the rest must still leave every
    FP16 adjoint finite: gradients finite and within 1e-2 of FP64 autograd
    (max error per tensor relative to its max) on a 512x3 level."""
    m = seeded_siren([512] * 3, 90.0, seed=72, last_scale=last_scale).cuda()
    st = LevelStack([m], "fp16")
    st.refresh()
    gen = torch.Generator().manual_seed(7)
    n = 4099
    x = (torch.rand(n, 3, generator=gen) * 2 - 1).cuda()
    s = torch.randn(n, generator=gen) * 2.0 ** -12
    g = torch.randn(n, 3, generator=gen) * 2.0 ** -12
    s[17] = 2.0 ** -4  # one seed 2^8 x the others' scale
    s, g = s.cuda(), g.cuda()
    _, _, dp = _abi_train(st, x, s, g, keep=1)
    assert torch.isfinite(dp).all()
    lay = [(torch.from_numpy(w).requires_grad_(True), torch.from_numpy(b).requires_grad_(True))
           for w, b in osiren.layers_from_model(m)]
    fr, gr = olosses.level_value_grad(lay, x.double().cpu(), True)
    obj = (s.double().cpu()[:, None] * fr).sum() + (g.double().cpu() * gr).sum()
    ref = torch.autograd.grad(obj, [w for w, _ in lay] + [b for _, b in lay])
    off = 0
    for r in ref:
        got = dp[off:off + r.numel()].double().cpu().view_as(r)
        off += r.numel()
        err = ((got - r).abs().max() / r.abs().max()).item()
        print(f"fp16 x{last_scale} {tuple(r.shape)}: max rel err {err:.2e}")
        assert err <= 1e-2, (tuple(r.shape), err)


def test_train_entry_points_check_their_arguments():
    from mplicits import _lib
    lib = _lib.load()
    m = seeded_siren([64], 30.0, seed=3).cuda()  # one sine layer: no hidden GEMM to train through
    st = LevelStack([m], "fp32")
    x = torch.zeros(8, 3, device="cuda")
    rc = lib.mp_train_fwd(st._handle, 0, _lib.ptr(x), 8, None, None, 0, _lib.stream_handle())
    assert rc == _lib.MP_EUNSUPPORTED and b"sine layers" in lib.mp_last_error()
    m2 = seeded_siren([64, 64], 30.0, seed=4).cuda()
    st2 = LevelStack([m2], "bf16x3")
    rc = lib.mp_train_fwd(st2._handle, 0, _lib.ptr(x), 8, None, None, 0, _lib.stream_handle())
    assert rc == _lib.MP_EUNSUPPORTED and b"three-term" in lib.mp_last_error()
    rc = lib.mp_train_fwd(st2._handle, 1, _lib.ptr(x), 8, None, None, 0, _lib.stream_handle())
    assert rc == _lib.MP_EINVAL


@pytest.mark.parametrize("precision,tol", [("bf16", 3e-2), ("fp32", 1e-4)])
def test_train_entry_points_at_many_tiles(precision, tol):
    """40000 points (160000 activation rows: many row tiles per GEMM and many
    split-K chunks per weight gradient), 512-wide layers, against FP64
    autograd of the oracle's SIREN run on the GPU."""
    m = seeded_siren([512, 512, 512], 90.0, seed=91, last_scale=0.1).cuda()
    st = LevelStack([m], precision)
    st.refresh()
    gen = torch.Generator().manual_seed(9)
    n = 40000
    x = (torch.rand(n, 3, generator=gen) * 2 - 1).cuda()
    s = torch.randn(n, generator=gen).cuda()
    g = torch.randn(n, 3, generator=gen).cuda()
    f, fd, dp = _abi_train(st, x, s, g, keep=1)
    lay = [(torch.from_numpy(w).cuda().requires_grad_(True), torch.from_numpy(b).cuda().requires_grad_(True))
           for w, b in osiren.layers_from_model(m)]
    fr, gr = olosses.level_value_grad(lay, x.double(), True)
    obj = (s.double()[:, None] * fr).sum() + (g.double() * gr).sum()
    ref = torch.autograd.grad(obj, [w for w, _ in lay] + [b for _, b in lay])
    ftol = 1e-5 if precision == "fp32" else 3e-3
    assert (f.double() - fr[:, 0].detach()).abs().max() <= ftol * max(1.0, fr.abs().max().item())
    assert (fd.double() - gr.detach()).abs().max() <= 10 * ftol * gr.abs().max().item()
    off = 0
    for r in ref:
        got = dp[off:off + r.numel()].double().view_as(r)
        off += r.numel()
        err = ((got - r).abs().max() / r.abs().max()).item()
        assert err <= tol, (tuple(r.shape), err)


def test_wide_level_gradients_bf16():
    coarse, med, fine = _models((256, 256, 256))
    x, gt, nrm = _batch(2048, 3)
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
