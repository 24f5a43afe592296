"""Fused level kernel (mp_eval) against the FP64 oracle (restated i3d SIREN forward)."""
import numpy as np
import pytest
import torch

from mplicits.stack import LevelStack, grid_coords
from mplicits.synthetic import seeded_siren
from oracle import grid as ogrid
from oracle import siren as osiren

pytestmark = pytest.mark.gpu

# (hidden widths, omega, last-layer scale): padded widths 64..512, ragged widths,
# one sine layer (no hidden GEMM) and a deeper stack (ping-pong path).
CASES = [
    ([64, 64], 30.0, 1.0),
    ([128, 128, 128], 30.0, 1.0),
    ([256, 256, 256], 30.0, 1.0),
    ([512, 512, 512], 90.0, 0.1),
    ([100, 200, 50], 30.0, 1.0),
    ([256], 30.0, 1.0),
    ([256, 256, 256, 256, 256], 30.0, 1.0),
]
# Max |f_gpu - f_fp64| relative to max |f_fp64| over the sample.
REL_TOL = {"fp32": 2e-5, "fp16": 2e-3, "bf16": 1e-2, "bf16x3": 2e-5}


def _points(n, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(n, 3, generator=g, dtype=torch.float64) * 2 - 1


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "bf16x3"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_forward_matches_fp64_oracle(case, precision):
    hidden, omega, last = case
    m = seeded_siren(hidden, omega, seed=1234 + len(hidden) + hidden[0], last_scale=last).cuda()
    x64 = _points(10007, 7)  # ragged: not a multiple of any tile
    ref = osiren.forward(osiren.layers_from_model(m), x64.numpy())
    st = LevelStack([m], precision)
    f = st.eval(x64.float().cuda()).double().cpu().numpy()
    err = np.abs(f - ref).max()
    scale = np.abs(ref).max()
    print(f"{precision} {hidden}: max err {err:.3e}, max |f| {scale:.3e}")
    # three-term BF16 runs levels wider than 256 in FP16
    tol = REL_TOL["fp16"] if precision == "bf16x3" and max(hidden) > 256 else REL_TOL[precision]
    assert err <= tol * scale


def test_fp32_absolute_error_on_unit_outputs():
    # A baked w=30 SIREN whose outputs are O(1): FP32 path within 1e-5 absolute.
    m = seeded_siren([256, 256, 256], 30.0, seed=99, last_scale=30.0).cuda()
    x64 = _points(20000, 3)
    ref = osiren.forward(osiren.layers_from_model(m), x64.numpy())
    f = LevelStack([m], "fp32").eval(x64.float().cuda()).double().cpu().numpy()
    assert np.abs(ref).max() > 0.3
    assert np.abs(f - ref).max() <= 1e-5


def test_bf16x3_absolute_error_on_unit_outputs():
    # the model of the FP32 test above (outputs up to ~3.5): three-term BF16
    # within 5e-5 absolute, plain BF16 off by more than 1e-3
    m = seeded_siren([256, 256, 256], 30.0, seed=99, last_scale=30.0).cuda()
    x64 = _points(20000, 3)
    ref = osiren.forward(osiren.layers_from_model(m), x64.numpy())
    f3 = LevelStack([m], "bf16x3").eval(x64.float().cuda()).double().cpu().numpy()
    f1 = LevelStack([m], "bf16").eval(x64.float().cuda()).double().cpu().numpy()
    print(f"bf16x3 max err {np.abs(f3 - ref).max():.3e}, bf16 {np.abs(f1 - ref).max():.3e}")
    assert np.abs(f3 - ref).max() <= 5e-5
    assert np.abs(f1 - ref).max() > 1e-3


# Three-term BF16 with the last hidden product in FP16 (split activations):
# levels with three hidden layers up to width 256 take it; every other level
# runs plain three-term BF16 (or FP16 above 256), bit for bit.
@pytest.mark.parametrize("case", CASES + [([64, 64, 64], 30.0, 1.0)], ids=lambda c: "x".join(map(str, c[0])))
def test_bf16x3f_forward(case):
    hidden, omega, last = case
    m = seeded_siren(hidden, omega, seed=1234 + len(hidden) + hidden[0], last_scale=last).cuda()
    x64 = _points(10007, 7)
    x = x64.float().cuda()
    f = LevelStack([m], "bf16x3f").eval(x)
    if len(hidden) != 3 or max(hidden) > 256:
        assert torch.equal(f, LevelStack([m], "bf16x3").eval(x))
        return
    ref = osiren.forward(osiren.layers_from_model(m), x64.numpy())
    err = np.abs(f.double().cpu().numpy() - ref).max()
    scale = np.abs(ref).max()
    e16 = np.abs(LevelStack([m], "fp16").eval(x).double().cpu().numpy() - ref).max()
    print(f"bf16x3f {hidden}: max err {err:.3e} (fp16 {e16:.3e}), max |f| {scale:.3e}")
    # the FP16 last product rounds only its weights (and the lo activations):
    # below the error of the all-FP16 path, within 3e-4 of the output scale
    assert err <= 3e-4 * scale
    assert err < 0.8 * e16


def test_bf16x3f_error_scales_with_the_output_weights():
    # outputs up to ~3.5 (output weights x 30): the error of the FP16 last
    # product grows with them, staying within 3e-4 of the output scale and
    # below the all-FP16 path's (the three-term path: 5e-5 absolute here)
    m = seeded_siren([256, 256, 256], 30.0, seed=99, last_scale=30.0).cuda()
    x64 = _points(20000, 3)
    ref = osiren.forward(osiren.layers_from_model(m), x64.numpy())
    x = x64.float().cuda()
    err = np.abs(LevelStack([m], "bf16x3f").eval(x).double().cpu().numpy() - ref).max()
    e16 = np.abs(LevelStack([m], "fp16").eval(x).double().cpu().numpy() - ref).max()
    print(f"bf16x3f max err {err:.3e} (fp16 {e16:.3e}), max |f| {np.abs(ref).max():.3e}")
    assert err <= 3e-4 * np.abs(ref).max()
    assert err < 0.8 * e16


def test_bf16x3f_refusals():
    # rendering, the batched extraction and training refuse the mixed path;
    # tracing refuses a stack with a level of another depth up to width 256
    from mplicits import training
    from mplicits.batch import BatchedExtraction
    from mplicits.tracer import FrameRenderer, trace, renderer_camera
    m = seeded_siren([256, 256, 256], 30.0, seed=5).cuda()
    st = LevelStack([m], "bf16x3f")
    x = _points(1000, 1).float().cuda()
    with pytest.raises(NotImplementedError):
        FrameRenderer(st, [renderer_camera(64, 64)], 64, 64).frame()
    with pytest.raises(NotImplementedError):
        BatchedExtraction([st], [32], [[]])
    with pytest.raises(NotImplementedError):
        training.level_forward(m, x, stack=st)
    two = LevelStack([m, seeded_siren([128, 128], 60.0, seed=6, last_scale=0.1).cuda()], "bf16x3f")
    with pytest.raises(NotImplementedError):
        trace(two, [renderer_camera(64, 64)], 64, 64, deltas=[0.05])


def test_bf16x3_levels_mixed_with_fp16():
    levels = [seeded_siren([256, 256, 256], 30.0, seed=7).cuda(),
              seeded_siren([512, 512, 512], 90.0, seed=8, last_scale=0.1).cuda()]
    x = _points(3001, 4).float().cuda()
    st = LevelStack(levels, "bf16x3")
    f, g = st.eval(x, grad=True)
    # level 1 (512 wide) runs FP16: the sum (and its gradient) equals level 0
    # in three-term BF16 plus level 1 in FP16, evaluated apart
    f0, g0 = LevelStack(levels[:1], "bf16x3").eval(x, grad=True)
    f1, g1 = LevelStack(levels[1:], "fp16").eval(x, grad=True)
    assert (f - (f0 + f1)).abs().max().item() <= 1e-6
    assert (g - (g0 + g1)).abs().max().item() <= 1e-4 * g.abs().max().item()
    assert torch.equal(st.eval(x), f)


def test_level_sum_and_empty_input():
    levels = [seeded_siren([128, 128], 30.0, seed=s).cuda() for s in (1, 2, 3)]
    st = LevelStack(levels, "fp32")
    x64 = _points(513, 11)
    ref = sum(osiren.forward(osiren.layers_from_model(m), x64.numpy()) for m in levels[1:])
    f = st.eval(x64.float().cuda(), lvl0=1, lvl1=3).double().cpu().numpy()
    assert np.abs(f - ref).max() <= 1e-5
    empty = st.eval(torch.zeros(0, 3, device="cuda"))
    assert empty.shape == (0,)
    with pytest.raises(ValueError):
        st.eval(x64.float().cuda(), lvl0=2, lvl1=1)


def test_weights_update_repacks():
    m = seeded_siren([64, 64], 30.0, seed=5).cuda()
    st = LevelStack([m], "fp32")
    x = _points(300, 2).float().cuda()
    f0 = st.eval(x).clone()
    with torch.no_grad():
        m.net[-1][0].bias.add_(0.25)
    f1 = st.eval(x)
    assert torch.allclose(f1, f0 + 0.25, atol=1e-6)


@pytest.mark.parametrize("mode", ["reference", "lattice"])
@pytest.mark.parametrize("N", [4, 33, 512])
def test_grid_coords_bit_exact(N, mode):
    total = N ** 3
    first = max(0, total - 70000) if N == 512 else 0
    xyz = grid_coords(N, mode, first=first, n=total - first).cpu().numpy()
    np.testing.assert_array_equal(xyz, ogrid.coords(N, np.arange(first, total), mode))


@pytest.mark.parametrize("mode", ["reference", "lattice"])
@pytest.mark.parametrize("N", [2, 8, 64, 256, 512, 1024])
def test_grid_coords_power_of_two_ranges_bit_exact(N, mode):
    # power-of-two N takes the exact-scaling path: ranges at the start, around
    # the float32 integer limits 2^24 .. 2^29 (where float32(i) rounds), at
    # random places and at the end of the grid
    total = N ** 3
    starts = {0, max(0, total - 40000)}
    for e in range(24, 30):
        if 2 ** e < total:
            starts.add(max(0, 2 ** e - 20000))
    rng = np.random.default_rng(N)
    starts.update(int(v) for v in rng.integers(0, max(1, total - 40000), 4))
    for first in sorted(starts):
        n = min(40000, total - first)
        xyz = grid_coords(N, mode, first=first, n=n).cpu().numpy()
        np.testing.assert_array_equal(xyz, ogrid.coords(N, np.arange(first, first + n), mode), err_msg=str(first))


GRAD_TOL = {"fp32": 2e-5, "fp16": 3e-3, "bf16": 2e-2, "bf16x3": 4e-5, "bf16x3f": 6e-4}


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "bf16x3", "bf16x3f"])
@pytest.mark.parametrize("case", [([64, 64], 30.0, 1.0), ([256, 256, 256], 30.0, 1.0),
                                  ([128, 100, 128, 128], 30.0, 1.0), ([512, 512, 512], 90.0, 0.1),
                                  ([256], 30.0, 1.0)],
                         ids=lambda c: "x".join(map(str, c[0])))
def test_gradient_matches_fp64_chain_rule(case, precision):
    hidden, omega, last = case
    m = seeded_siren(hidden, omega, seed=77 + hidden[0], last_scale=last).cuda()
    x64 = _points(4099, 5)
    layers = osiren.layers_from_model(m)
    f_ref, g_ref = osiren.forward_grad(layers, x64.numpy())
    f, g = LevelStack([m], precision).eval(x64.float().cuda(), grad=True)
    g = g.double().cpu().numpy()
    gs = np.abs(g_ref).max()
    err = np.abs(g - g_ref).max()
    print(f"{precision} {hidden}: grad max err {err:.3e} (max |grad| {gs:.3e})")
    # three-term BF16 runs levels wider than 256 in FP16, the mixed path
    # levels of another depth in three-term BF16
    tol = (GRAD_TOL["fp16"] if precision in ("bf16x3", "bf16x3f") and max(hidden) > 256
           else GRAD_TOL["bf16x3"] if precision == "bf16x3f" and len(hidden) != 3 else GRAD_TOL[precision])
    assert err <= tol * gs
    assert np.abs(f.double().cpu().numpy() - f_ref).max() <= tol * np.abs(f_ref).max()


def test_gradient_of_level_sum():
    levels = [seeded_siren([128, 128, 128], 30.0, seed=s).cuda() for s in (4, 5)]
    x64 = _points(1000, 9)
    ref = sum(osiren.forward_grad(osiren.layers_from_model(m), x64.numpy())[1] for m in levels)
    _, g = LevelStack(levels, "fp32").eval(x64.float().cuda(), grad=True)
    assert np.abs(g.double().cpu().numpy() - ref).max() <= 2e-5 * np.abs(ref).max()


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_single_column_tiles_equal_default_tiles(case, precision):
    # single-column tiles run two waves per SIMD (eight-wave workgroups) up to
    # width 256; each point's arithmetic is the same, so the values are too
    hidden, omega, last = case
    m = seeded_siren(hidden, omega, seed=4321 + hidden[0], last_scale=last).cuda()
    x = _points(10007, 9).float().cuda()
    st = LevelStack([m], precision)
    ref = st.eval(x).clone()
    st.set_tiles(1)
    assert torch.equal(st.eval(x), ref)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_single_column_tiles_in_the_culled_extraction(precision):
    levels = [seeded_siren([256] * 3, 30.0, seed=5, last_scale=20.0).cuda(),
              seeded_siren([128] * 3, 60.0, seed=6, last_scale=0.1).cuda()]
    st = LevelStack(levels, precision)
    vol_a = torch.empty(96 ** 3, device="cuda")
    vol_b = torch.empty(96 ** 3, device="cuda")
    _, counts_a = st.extract_volume(96, [0.05], chunk=4096, out=vol_a)
    st.set_tiles(1)
    _, counts_b = st.extract_volume(96, [0.05], chunk=4096, out=vol_b)
    assert counts_a == counts_b and counts_a[1] > 0
    assert torch.equal(vol_a, vol_b)


def test_forward_after_update_omegas_is_unchanged():
    # the module's cached stack re-packs the baked weights with the new
    # frequency (update_omegas rewrites the parameters and resets w0 in place)
    from mplicits.model import SIREN
    torch.manual_seed(8)
    m = SIREN(3, 1, [64, 64], w0=30, ww=30).cuda()
    x = _points(4096, 9).float().cuda()
    with torch.no_grad():
        before = m(x)["model_out"].clone()
        m.update_omegas(1)
        after = m(x)["model_out"]
    assert (after - before).abs().max().item() <= 1e-5
