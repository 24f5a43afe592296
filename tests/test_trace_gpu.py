"""Persistent multiscale sphere tracer (mp_trace) against a closed form and the
FP64 oracle's restatement of the renderer's step rule."""
import numpy as np
import pytest
import torch

import os.path as osp

from conftest import ROOT
from mplicits.model import SIREN, from_pth
from mplicits.stack import LevelStack
from mplicits.synthetic import orbit_cameras, seeded_siren
from mplicits.tracer import trace
from oracle import siren as osiren
from oracle import trace as otrace
from checks import check_trace_against_oracle as _check_against_oracle

pytestmark = pytest.mark.gpu


def _plane_model(c=0.2, eps=0.1):
    # f(x) = sin(eps * z) / eps - c: zero set is the plane z = asin(eps c) / eps,
    # gradient (0, 0, cos(eps z)); baked weights, 3 -> 64 -> 1.
    m = SIREN(3, 1, [64], w0=1, ww=1, delay_init=True)
    with torch.no_grad():
        for p in m.parameters():
            p.zero_()
        m.net[0][0].weight[0, 2] = eps
        m.net[1][0].weight[0, 0] = 1.0 / eps
        m.net[1][0].bias[0] = -c
    return m.cuda(), np.arcsin(eps * c) / eps


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_plane_depth_and_normals_closed_form(precision):
    m, zp = _plane_model()
    cams = orbit_cameras(3, seed=5, radius=2.5, elev_deg=60)
    W, H = 96, 64
    out = trace(LevelStack([m], precision), cams, W, H, deltas=[], max_steps=256, hit_eps=1e-5)
    depth = out["depth"].cpu().numpy().reshape(-1)
    P, D, t0, inside = otrace.rays(cams, W, H)
    # eye positions and the analytic intersection with z = zp inside the sphere
    eye = np.concatenate([np.repeat(np.asarray(iv)[:3, 3][None], W * H, 0) for iv, _ in cams])
    t_plane = (zp - eye[:, 2]) / D[:, 2]
    xp = eye + t_plane[:, None] * D
    # a ray that enters the sphere on the positive side (f > 0) walks forward
    # to the plane; one that enters on the negative side steps backward out of
    # the sphere (signed steps, as in the reference tracers) and misses
    expect_hit = (P[:, 2] > zp) & (t_plane > t0) & ((xp ** 2).sum(1) < 1.2 ** 2 - 1e-3) & inside
    got_hit = depth >= 0
    # away from the sphere rim every ray must agree
    rim = np.abs(np.sqrt((xp ** 2).sum(1)) - 1.2) < 0.02
    assert np.array_equal(got_hit[~rim], expect_hit[~rim])
    both = got_hit & expect_hit
    assert both.sum() > 1000
    assert np.abs(depth[both] - t_plane[both]).max() < 5e-4
    nrm = out["normals"].cpu().numpy().reshape(-1, 3)[both]
    assert np.abs(np.abs(nrm[:, 2]) - 1).max() < 1e-3
    assert out["hits"] == int(got_hit.sum())


@pytest.mark.parametrize("precision,prec_err", [("fp32", 0.0), ("bf16", 1e-3), ("fp16", 3e-4), ("bf16x3", 1e-5),
                                                ("bf16x3f", 2e-4)])
def test_stack_matches_oracle_step_rule(precision, prec_err):
    coarse = from_pth(osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")).cuda()
    levels = [coarse, seeded_siren([128] * 3, 60.0, seed=42, last_scale=0.1).cuda()]
    deltas = [0.05]
    cams = orbit_cameras(2, seed=9)
    W, H = 64, 48
    out = trace(LevelStack(levels, precision), cams, W, H, deltas=deltas)
    _check_against_oracle(out, levels, deltas, cams, W, H, 600, 0, prec_err)
    assert out["evals"][0] > 0 and out["evals"][1] > 0


def test_single_level_matches_oracle_step_rule():
    coarse = from_pth(osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")).cuda()
    cams = orbit_cameras(2, seed=10)
    W, H = 64, 48
    out = trace(LevelStack([coarse], "fp32"), cams, W, H, deltas=[])
    _check_against_oracle(out, [coarse], [], cams, W, H, 600, 4)


@pytest.mark.parametrize("flags,iters", [(["-iters=20,5,5", "-delta=0.05", "-normal_lod=2"], "20/5/5"),
                                         (["-adaptive", "-delta=0.05"], "adaptive:128")])
def test_render_cli_benchmark_line(tmp_path, capsys, flags, iters):
    import importlib.util
    import os.path as osp
    from mplicits import renderer_io
    from mplicits.synthetic import seeded_siren
    levels = [seeded_siren([64] * 3, 30.0, 1, last_scale=20.0), seeded_siren([64] * 2, 60.0, 2, last_scale=0.1),
              seeded_siren([64] * 2, 90.0, 3, last_scale=0.1)]
    desc = renderer_io.export_experiment(levels, str(tmp_path / "data" / "m"), "m", data_root="data")
    spec = importlib.util.spec_from_file_location(
        "render_cli", osp.join(osp.dirname(__file__), "..", "m-plicits_amd", "render.py"))
    render = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(render)
    render.main([f"-experiment_file={desc}", f"-data_root={tmp_path}", "-resolution=64x48", "-benchmark=2",
                 f"-out={tmp_path / 'img.ppm'}", f"-csv={tmp_path / 'benchmark.csv'}"] + flags)
    out = capsys.readouterr().out
    assert f"BENCHMARK resolution=64x48 iters={iters} delta=0.0500" in out and "frames=2 avg_fps=" in out
    assert open(tmp_path / "img.ppm", "rb").read(11) == b"P6 64 48 25"
    assert open(tmp_path / "benchmark.csv").read().startswith(f"64x48,{iters},0.0500,")


@pytest.mark.parametrize("n_levels", [1, 3])
def test_rays_that_miss_the_bounding_sphere(n_levels):
    """A camera looking away from the scene: every ray misses the bounding
    sphere (depth -1, no steps, no evaluations; with residual levels the
    level-by-level launches after the first find nothing parked).  Beside a
    camera that sees the surface, that camera's rays equal its own frame."""
    from mplicits.synthetic import look_at, perspective
    coarse = from_pth(osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")).cuda()
    levels = [coarse]
    if n_levels == 3:
        levels += [seeded_siren([128] * 3, 60.0, seed=42, last_scale=0.1).cuda(),
                   seeded_siren([64] * 2, 90.0, seed=43, last_scale=0.1).cuda()]
    deltas = [0.05, 0.02][:n_levels - 1]
    st = LevelStack(levels, "bf16")
    away = (np.linalg.inv(look_at((0.0, 0.0, 3.0), target=(0.0, 0.0, 10.0))),
            np.linalg.inv(perspective(45.0, 1.5)))
    W, H = 48, 32
    out = trace(st, [away], W, H, deltas=deltas)
    assert bool((out["depth"] == -1).all()) and bool((out["steps"] == 0).all())
    assert out["hits"] == 0 and out["evals"] == [0] * n_levels
    see = orbit_cameras(1, seed=3)[0]
    alone = trace(st, [see], W, H, deltas=deltas)
    mixed = trace(st, [away, see], W, H, deltas=deltas)
    assert alone["hits"] > 100
    for k in ("depth", "steps", "rgba", "normals"):
        assert torch.equal(mixed[k][1], alone[k][0]), k
        assert torch.equal(mixed[k][0], out[k][0]), k
    assert mixed["hits"] == alone["hits"] and mixed["evals"] == alone["evals"]


def test_three_level_stack_matches_oracle_step_rule():
    """Three levels (coarse launch, then the parked rays on every level; a
    512-wide fine level) against the FP64 restatement of the step rule."""
    coarse = from_pth(osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")).cuda()
    levels = [coarse, seeded_siren([128] * 3, 60.0, seed=42, last_scale=0.1).cuda(),
              seeded_siren([512] * 3, 90.0, seed=44, last_scale=0.1).cuda()]
    deltas = [0.05, 0.02]
    cams = orbit_cameras(2, seed=11)
    W, H = 64, 48
    # three-term BF16 runs the 512-wide fine level in FP16 (output weights x 0.1)
    for precision, prec_err in (("fp32", 0.0), ("bf16", 1e-3), ("bf16x3", 1e-4), ("bf16x3f", 2e-4)):
        out = trace(LevelStack(levels, precision), cams, W, H, deltas=deltas)
        # the surface covers ~8% of the frame
        _check_against_oracle(out, levels, deltas, cams, W, H, 1000, 1, prec_err)
        assert all(e > 0 for e in out["evals"]) and out["evals"][0] > out["evals"][1] == out["evals"][2]


def test_four_level_stack_matches_oracle_step_rule():
    """Four levels (the most a stack holds)."""
    coarse = from_pth(osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")).cuda()
    levels = [coarse, seeded_siren([128] * 3, 60.0, seed=45, last_scale=0.1),
              seeded_siren([64] * 2, 90.0, seed=46, last_scale=0.1),
              seeded_siren([128] * 2, 120.0, seed=47, last_scale=0.05)]
    # residuals without output bias: stacked constant offsets would move the
    # zero of the sum out of the coarse band, and no ray would hit
    with torch.no_grad():
        for m in levels[1:]:
            m.net[-1][0].bias.zero_()
    levels = [levels[0]] + [m.cuda() for m in levels[1:]]
    deltas = [0.05, 0.03, 0.02]
    cams = orbit_cameras(2, seed=13)
    W, H = 64, 48
    out = trace(LevelStack(levels, "fp32"), cams, W, H, deltas=deltas)
    _check_against_oracle(out, levels, deltas, cams, W, H, 1000, 2)
    assert len(out["evals"]) == 4 and all(e > 0 for e in out["evals"])
