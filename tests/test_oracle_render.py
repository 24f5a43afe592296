"""Closed-form checks of the fixed-iteration frame oracle (oracle/render.py):
analytic levels (planes and constants) whose sphere-tracing iterates, shading
distances and colours are known in closed form."""
import math

import numpy as np
import pytest

from mplicits.synthetic import look_at, perspective
from oracle import render as orender

W = H = 8
CENTER = (H // 2) * W + W // 2  # ndc (0, 0): the ray straight down -z
CORNER = 0                      # ndc (-1, -1): misses the bounding sphere


def _cams():
    return [(np.linalg.inv(look_at((0.0, 0.0, 3.0))), np.linalg.inv(perspective(45.0, 1.0)))]


def plane(normal, offset=0.0):
    n = np.asarray(normal, np.float64)
    return (lambda P: P @ n + offset, lambda P: np.broadcast_to(n, P.shape).copy())


def const(r):
    return (lambda P: np.full(len(P), r), lambda P: np.zeros_like(P))


def _frame(levels, **kw):
    return orender.render_fixed(levels, _cams(), W, H, **kw)


def test_tilted_plane_converges_geometrically_over_exactly_the_given_steps():
    th = math.radians(60.0)
    lv = [plane((math.sin(th), 0.0, math.cos(th)))]
    fr = _frame(lv, iters=(20, 0, 0), delta=0.0)
    # entry at z = 1.2; along -z the plane distance halves every step
    assert fr["pos"][CENTER] == pytest.approx([0.0, 0.0, 1.2 * 2.0 ** -20], abs=1e-15)
    # the shading distance is the last evaluation, taken before the last step
    assert fr["dist"][CENTER] == pytest.approx(0.6 * 2.0 ** -19, rel=1e-9)
    # Phong at 60 degrees: 0.2 + 0.54 * 0.5 = 0.47, specular term 0 -> floor(119.85)
    assert fr["rgba"][CENTER] == 119 | 119 << 8 | 119 << 16 | 255 << 24


@pytest.mark.parametrize("deg,grey", [(0.0, 255), (10.0, 201), (20.0, 180), (60.0, 119)])
def test_phong_known_answers(deg, grey):
    # 10 degrees: 0.2 + 0.54 cos + 0.5 (2 cos^2 - 1)^35 = 0.78846 -> 201 (186 without the specular term)
    th = math.radians(deg)
    fr = _frame([plane((math.sin(th), 0.0, math.cos(th)))], iters=(40, 0, 0))
    assert fr["shaded"][CENTER]
    assert fr["rgba"][CENTER] == grey | grey << 8 | grey << 16 | 255 << 24


def test_residual_phase_steps_by_the_sum_inside_the_band():
    d, r = 0.05, 0.01
    fr = _frame([plane((0, 0, 1)), const(r)], iters=(20, 3, 0), delta=d, shading="normals")
    # phase 0 stops at z = delta, phase 1 moves to the zero of z + r and stays
    assert fr["pos"][CENTER] == pytest.approx([0, 0, -r], abs=1e-15)
    assert fr["dist"][CENTER] == pytest.approx(0.0, abs=1e-15)
    # normal (0, 0, 1) under the normals colouring (|x|, |z|, |y|)
    assert fr["rgba"][CENTER] == 0 | 255 << 8 | 0 << 16 | 255 << 24


def test_standalone_levels_step_by_the_finer_level_inside_the_band():
    d, r = 0.05, 0.01
    fr = _frame([plane((0, 0, 1)), plane((0, 0, 1), r)], iters=(20, 2, 0), delta=d, residual=False)
    assert fr["pos"][CENTER] == pytest.approx([0, 0, -r], abs=1e-15)
    assert fr["dist"][CENTER] == pytest.approx(0.0, abs=1e-15)  # level 1's own distance


def test_outside_the_band_phase_one_steps_by_level_zero():
    fr = _frame([plane((0, 0, 1)), const(0.01)], iters=(20, 1, 0), delta=0.05, skip_lod0=True)
    assert fr["pos"][CENTER] == pytest.approx([0, 0, 0.0], abs=1e-15)  # one full f0 step from z = 1.2
    assert fr["dist"][CENTER] == pytest.approx(1.21)                    # f0 + f1 before that step
    assert not fr["shaded"][CENTER] and fr["rgba"][CENTER] == orender.WHITE


@pytest.mark.parametrize("alias", [False, True])
def test_level_two_binding(alias):
    d, r, s = 0.05, 0.01, 0.003
    lv = [plane((0, 0, 1)), const(r), const(s)]
    fr = _frame(lv, iters=(1, 0, 1), delta=d, lod2_reads_lod0=alias)
    # phase 0 lands on z = delta; the phase-2 step is f0 + f1 + (f0 if aliased else f2)
    z = d - (d + r + (d if alias else s))
    assert fr["pos"][CENTER] == pytest.approx([0, 0, z], abs=1e-15)
    assert fr["dist"][CENTER] == pytest.approx(d + r + (d if alias else s), abs=1e-15)


@pytest.mark.parametrize("residual,lod,expect", [(True, 0, (0, 255, 0)), (True, 1, (180, 180, 0)),
                                                  (False, 1, (255, 0, 0))])
def test_normal_lod(residual, lod, expect):
    # level 1 = x: its gradient (1, 0, 0) is added to (0, 0, 1) in residual models
    lv = [plane((0, 0, 1)), plane((1, 0, 0))]
    fr = _frame(lv, iters=(20, 0, 0), normal_lod=lod, residual=residual, shading="normals")
    r, g, b = expect
    assert fr["rgba"][CENTER] == r | g << 8 | b << 16 | 255 << 24


def test_rays_that_miss_the_bounding_sphere_start_at_the_eye():
    fr = _frame([plane((0, 0, 1))], iters=(0, 0, 0))
    assert fr["pos"][CORNER] == pytest.approx([0, 0, 3.0])
    assert fr["pos"][CENTER] == pytest.approx([0, 0, 1.2])


def test_truncating_rgba8():
    assert int(orender.rgba8(0.47, 1.5, -1.0)) == 119 | 255 << 8 | 0 << 16 | 255 << 24
