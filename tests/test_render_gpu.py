"""The renderer's fixed-iteration frame (mp_render) against the FP64 oracle of
MIPplicit::SdfInference (oracle/render.py), and its graph replays."""
import os.path as osp

import numpy as np
import pytest
import torch

from conftest import ROOT
from mplicits.model import from_pth
from mplicits.stack import LevelStack
from mplicits.synthetic import orbit_cameras, seeded_siren
from mplicits.tracer import FrameRenderer, render
from oracle import render as orender
from oracle import siren as osiren

pytestmark = pytest.mark.gpu

W, H = 64, 48


def _levels():
    coarse = from_pth(osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")).cuda()
    return [coarse, seeded_siren([128] * 3, 60.0, seed=42, last_scale=0.1).cuda(),
            seeded_siren([64] * 3, 90.0, seed=43, last_scale=0.05).cuda()]


def _channels(rgba):
    v = np.asarray(rgba).astype(np.int64) & 0xFFFFFFFF
    return np.stack([(v >> s) & 255 for s in (0, 8, 16, 24)], -1)


def _compare(out, ref, threshold=0.05, mask_agree=0.99, pos_tol=1e-4, color_agree=0.97):
    pos = out["pos"].cpu().numpy().reshape(-1, 3)
    dist = out["dist"].cpu().numpy().reshape(-1)
    rgba = out["rgba"].cpu().numpy().reshape(-1)
    dp = np.abs(pos - ref["pos"]).max(1)
    assert np.median(dp) < pos_tol, np.median(dp)
    assert np.median(np.abs(dist - ref["dist"])) < pos_tol
    # shaded pixels carry a unit normal (a saturated Phong pixel is white too)
    shaded = np.abs(out["normals"].cpu().numpy().reshape(-1, 3)).sum(1) > 0
    assert np.all(rgba[~shaded].view(np.uint32) == orender.WHITE)
    # away from the shading test's boundary (|dist| = threshold, |x_k| = 1)
    # the decisions must agree
    border = (np.abs(np.abs(ref["dist"]) - threshold) < 2e-3) | (np.abs(np.abs(ref["pos"]) - 1.0) < 2e-3).any(1)
    agree = (shaded == ref["shaded"])[~border]
    assert agree.mean() >= mask_agree, (agree.mean(), border.mean())
    both = shaded & ref["shaded"]
    assert both.sum() > 100
    dc = np.abs(_channels(rgba[both]) - _channels(ref["rgba"][both])).max(1)
    assert (dc <= 1).mean() >= color_agree, (dc <= 1).mean()
    return both


@pytest.mark.parametrize("opts", [
    dict(iters=(20, 5, 5), delta=0.05, normal_lod=2),
    dict(iters=(20, 5, 0), delta=0.05, normal_lod=2),
    dict(iters=(20, 0, 0), delta=0.02, normal_lod=1, shading="normals"),
    dict(iters=(20, 5, 5), delta=0.05, normal_lod=2, lod2_reads_lod0=True),
    dict(iters=(10, 5, 3), delta=0.05, normal_lod=1, residual=False),
    dict(iters=(20, 6, 0), delta=0.05, normal_lod=0, skip_lod0=True, threshold=0.02),
])
def test_frame_matches_oracle_fp32(opts):
    levels = _levels()
    if not opts.get("residual", True):
        levels = [levels[0]] * 3  # standalone levels must be SDFs themselves
    cams = orbit_cameras(1, seed=3)
    out = render(LevelStack(levels, "fp32"), cams, W, H, **opts)
    ref = orender.render_fixed([orender.siren_level(osiren.layers_from_model(m)) for m in levels], cams, W, H,
                               **opts)
    _compare(out, ref, threshold=opts.get("threshold", 0.05))
    assert out["shaded"] == int((out["normals"].abs().sum(-1) > 0).sum())


def test_frame_matches_oracle_bf16():
    levels = _levels()
    cams = orbit_cameras(1, seed=4)
    opts = dict(iters=(20, 5, 5), delta=0.05, normal_lod=2)
    out = render(LevelStack(levels, "bf16"), cams, W, H, **opts)
    ref = orender.render_fixed([orender.siren_level(osiren.layers_from_model(m)) for m in levels], cams, W, H,
                               **opts)
    _compare(out, ref, mask_agree=0.9, pos_tol=5e-3, color_agree=0.6)


def test_frame_matches_oracle_bf16x3():
    # three-term BF16 evaluates within a few 1e-6 of FP64: the FP32 bounds hold
    levels = _levels()
    cams = orbit_cameras(1, seed=4)
    opts = dict(iters=(20, 5, 5), delta=0.05, normal_lod=2)
    out = render(LevelStack(levels, "bf16x3"), cams, W, H, **opts)
    ref = orender.render_fixed([orender.siren_level(osiren.layers_from_model(m)) for m in levels], cams, W, H,
                               **opts)
    _compare(out, ref)


def test_finer_levels_run_only_where_a_step_reads_them():
    levels = _levels()
    cams = orbit_cameras(1, seed=3)
    out = render(LevelStack(levels, "fp32"), cams, W, H, iters=(20, 5, 5), delta=0.05)
    n = W * H
    assert out["evals"][0] == 30 * n
    # levels 1 and 2 run at least on the step that feeds the shading distance
    assert n <= out["evals"][1] <= 10 * n and n <= out["evals"][2] <= 5 * n


def test_graph_replays_equal_the_eager_frame():
    levels = _levels()
    cams = orbit_cameras(2, seed=6)
    r = FrameRenderer(LevelStack(levels, "bf16"), cams, W, H, iters=(20, 5, 5), delta=0.05, normal_lod=2)
    r.frame()
    torch.cuda.synchronize()
    eager = r.rgba.clone(), r.pos.clone(), r.normals.clone()
    for k in range(3):
        r.rgba.fill_(0)
        r.pos.fill_(0)
        r.frame(graph=True)
        torch.cuda.synchronize()
        assert torch.equal(r.rgba, eager[0]) and torch.equal(r.pos, eager[1]) and torch.equal(r.normals, eager[2]), k
        _ = torch.randn(1 << 16, device="cuda").sum().item()  # eager work between replays


def test_rejects_missing_levels():
    levels = _levels()[:2]
    with pytest.raises(ValueError, match="level-2 iterations"):
        render(LevelStack(levels, "fp32"), orbit_cameras(1, seed=3), 8, 8, iters=(1, 1, 1))
