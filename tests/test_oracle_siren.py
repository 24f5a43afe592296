"""Pins the oracle SIREN (oracle/siren.py) and the init / checkpoint semantics
of mplicits.model against independent computations."""
import numpy as np
import pytest
import torch

from mplicits.model import SIREN, from_pth, from_state_dict, siren_v1_to_v2
from mplicits import util
from oracle import siren as osiren


def _layers(seed=0, hidden=(16, 16, 16), omega=30.0):
    torch.manual_seed(seed)
    m = SIREN(3, 1, list(hidden), w0=omega, ww=omega)
    return m, osiren.layers_from_model(m)


def test_forward_matches_torch_fp64():
    m, layers = _layers()
    x = torch.rand(257, 3, dtype=torch.float64) * 2 - 1
    tl = [(torch.from_numpy(w), torch.from_numpy(b)) for w, b in layers]
    ref = osiren.torch_forward(tl, x).squeeze(-1).numpy()
    np.testing.assert_allclose(osiren.forward(layers, x.numpy()), ref, rtol=0, atol=1e-12)


def test_gradient_against_autograd_and_finite_differences():
    m, layers = _layers(1, (32, 24, 32))
    x = torch.rand(64, 3, dtype=torch.float64) * 2 - 1
    f, g = osiren.forward_grad(layers, x.numpy())
    tl = [(torch.from_numpy(w), torch.from_numpy(b)) for w, b in layers]
    xr = x.clone().requires_grad_(True)
    y = osiren.torch_forward(tl, xr).sum()
    (ga,) = torch.autograd.grad(y, xr)
    np.testing.assert_allclose(g, ga.numpy(), rtol=1e-10, atol=1e-10)
    h = 1e-6
    fd = np.stack([(osiren.forward(layers, x.numpy() + h * e) - osiren.forward(layers, x.numpy() - h * e)) / (2 * h)
                   for e in np.eye(3)], 1)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(g).max())


def test_torch_forward_grad_matches_autograd():
    # the FP64 device evaluation the benchmark's checker runs at traced hits:
    # pinned to torch autograd of the plain forward, and a dropped cosine
    # factor (plain W . J) fails it
    m, layers = _layers(2, (32, 24, 32))
    x = torch.rand(97, 3, dtype=torch.float64) * 2 - 1
    tl = [(torch.from_numpy(w), torch.from_numpy(b)) for w, b in layers]
    f, g = osiren.torch_forward_grad(tl, x)
    xr = x.clone().requires_grad_(True)
    y = osiren.torch_forward(tl, xr)
    (ga,) = torch.autograd.grad(y.sum(), xr)
    np.testing.assert_allclose(f.numpy(), y.detach()[:, 0].numpy(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(g.numpy(), ga.numpy(), rtol=1e-10, atol=1e-10)
    # contrast: the chain without the hidden layers' cosine factors
    h, jac = x, None
    for k, (w, b) in enumerate(tl[:-1]):
        a = h @ w.T + b
        jac = torch.cos(a)[:, :, None] * w[None] if k == 0 else torch.matmul(w, jac)
        h = torch.sin(a)
    g_bad = torch.matmul(tl[-1][0], jac)[:, 0, :]
    assert (g_bad - ga).abs().max().item() > 1e-3 * ga.abs().max().item()


def test_init_follows_reference_rng_order_and_bounds():
    torch.manual_seed(123)
    m = SIREN(3, 1, [64, 64, 64], w0=30, ww=30)
    ref = osiren.reference_init_draws([64, 64, 64], 30, 30, 123)
    for (w, b), seq in zip(ref, m.net):
        assert torch.equal(seq[0].weight, w) and torch.equal(seq[0].bias, b)
    assert m.net[0][0].weight.abs().max() <= 1 / 3
    assert m.net[1][0].weight.abs().max() <= np.sqrt(6 / 64) / 30
    assert m.net[3][0].weight.abs().max() <= np.sqrt(6 / 64) / 30
    assert m.net[1][0].bias.abs().max() <= 1 / np.sqrt(64)
    # overwriting the weights before constructing every layer would draw differently
    torch.manual_seed(123)
    lin0 = torch.nn.Linear(3, 64)
    with torch.no_grad():
        lin0.weight.uniform_(-1 / 3, 1 / 3)
    assert not torch.equal(lin0.weight, m.net[0][0].weight)


def test_update_omegas_bakes_all_but_last_layer():
    m, _ = _layers(2, (16, 16))
    x = torch.rand(50, 3, dtype=torch.float64) * 2 - 1
    before = osiren.forward(osiren.layers_from_model(m), x.numpy())
    sd = {k: v.clone() for k, v in m.state_dict().items()}
    m.update_omegas(1)
    keys = list(sd)
    for k in keys[:-2]:
        assert torch.allclose(m.state_dict()[k], sd[k] * 30)
    for k in keys[-2:]:
        assert torch.equal(m.state_dict()[k], sd[k])
    assert m.w0 == 1 and m.net[1][1].w0 == 1
    after = osiren.forward(osiren.layers_from_model(m), x.numpy())
    np.testing.assert_allclose(after, before, atol=1e-5)


def test_checkpoint_loading(tmp_path):
    m, _ = _layers(3, (16, 32, 8))
    p = tmp_path / "best.pth"
    torch.save(m.state_dict(), p)
    for loader in (from_pth, util.from_pth):
        m2 = loader(str(p), w0=30)
        assert m2.hidden_widths() == [16, 32, 8]
        for (k, a), (_, b) in zip(m.state_dict().items(), m2.state_dict().items()):
            assert torch.equal(a, b)
    with pytest.raises(FileNotFoundError):
        from_pth(str(tmp_path / "missing.pth"))
    # version-1 keys (4-character prefix) are converted and written next to the file
    v1 = {"net." + k: v for k, v in m.state_dict().items()}
    p1 = tmp_path / "old.pth"
    torch.save(v1, p1)
    m3 = util.from_pth(str(p1))
    assert torch.equal(m3.net[0][0].weight, m.net[0][0].weight)
    assert (tmp_path / "old_v2.pth").exists()
    m4, upgraded = from_state_dict(v1)
    assert upgraded
    out, div = siren_v1_to_v2(v1, check_equals=True)
    assert not div and list(out) == list(m.state_dict())
