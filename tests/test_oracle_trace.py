"""The sphere-tracing oracle (oracle.trace) against closed forms (no GPU).

Ray generation: rays through pixel centres of hand-built camera matrices,
entry points on the bounding sphere.  Step rule: analytic level functions
whose traces are known step by step, chosen so that each rule of
IterateSphereTracing_kernel (renderer/cuda/kernels.cu:315-398) is visible in
the result: the lod-0 step F_0 - delta_0, the switch to the residual phase on
first entering the coarse band, the full F_0 step outside the band after that,
and the sum of every residual level inside the band (no test on delta_1)."""
import math

import numpy as np
import pytest

from oracle import trace as otrace


def _camera(eye_z, fov_deg, aspect=1.0, near=0.1, far=10.0):
    """Inverse view (camera at (0, 0, eye_z) looking down -z) and inverse
    OpenGL perspective projection, as numpy 4x4 matrices."""
    iv = np.eye(4)
    iv[2, 3] = eye_z
    f = 1.0 / math.tan(math.radians(fov_deg) / 2)
    proj = np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, (far + near) / (near - far), 2 * far * near / (near - far)], [0, 0, -1, 0]])
    return iv, np.linalg.inv(proj)


def test_ray_generation_centre_and_corner_pixels():
    cams = [_camera(2.5, 30.0)]
    # 3 x 3 image: the centre pixel's ray is the optical axis
    P, D, t0, inside = otrace.rays(cams, 3, 3)
    assert inside.all()
    np.testing.assert_allclose(D[4], [0, 0, -1], atol=1e-15)
    np.testing.assert_allclose(P[4], [0, 0, 1.2], atol=1e-12)
    assert abs(t0[4] - 1.3) < 1e-12
    # corner pixel (i, j) = (0, 0): NDC (-2/3, -2/3), direction angle
    # atan(tan(15 deg) * |ndc|) per axis
    tn = math.tan(math.radians(15)) * (2 / 3)
    d = np.array([-tn, -tn, -1.0]) / math.sqrt(1 + 2 * tn * tn)
    np.testing.assert_allclose(D[0], d, atol=1e-12)
    # every entry point is on the sphere, ahead of the eye along the ray
    np.testing.assert_allclose(np.linalg.norm(P, axis=1), 1.2, atol=1e-12)
    eye = np.array([0, 0, 2.5])
    np.testing.assert_allclose(eye + t0[:, None] * D, P, atol=1e-12)
    assert (t0 > 0).all()
    # pixel (2, 0) mirrors pixel (0, 0) in x; flip_y mirrors rows
    np.testing.assert_allclose(D[2], d * [-1, 1, 1], atol=1e-12)
    _, Df, _, _ = otrace.rays(cams, 3, 3, flip_y=True)
    np.testing.assert_allclose(Df[0], d * [1, -1, 1], atol=1e-12)
    # without pixel centres pixel 0 maps to NDC -1
    _, D0, _, _ = otrace.rays(cams, 2, 2, pixel_center=False)
    tn1 = math.tan(math.radians(15))
    np.testing.assert_allclose(D0[0], np.array([-tn1, -tn1, -1.0]) / math.sqrt(1 + 2 * tn1 * tn1), atol=1e-12)


def test_rays_of_several_cameras_in_any_id_order():
    # ids of three cameras, shuffled: each ray equals the same pixel of its
    # camera generated alone
    cams = [_camera(2.5, 30.0), _camera(3.0, 45.0), _camera(2.2, 60.0)]
    W, H = 7, 5
    ids = np.random.default_rng(3).permutation(3 * W * H)
    P, D, t0, inside = otrace.rays(cams, W, H, ids=ids)
    for c, cam in enumerate(cams):
        Pc, Dc, tc, ic = otrace.rays([cam], W, H)
        k = np.nonzero(ids // (W * H) == c)[0]
        pix = ids[k] % (W * H)
        np.testing.assert_allclose(D[k], Dc[pix], rtol=0, atol=1e-14)
        np.testing.assert_allclose(P[k], Pc[pix], rtol=0, atol=1e-13)
        np.testing.assert_allclose(t0[k], tc[pix], rtol=0, atol=1e-13)
        assert np.array_equal(inside[k], ic[pix])


def test_ray_that_misses_the_sphere_and_eye_inside():
    # a wide field of view: the corner rays pass outside the sphere
    P, D, t0, inside = otrace.rays([_camera(2.5, 120.0)], 3, 3)
    assert inside[4] and not inside[0]
    # an eye inside the sphere starts at the eye
    P, D, t0, inside = otrace.rays([_camera(0.5, 30.0)], 1, 1)
    assert inside[0] and t0[0] == 0.0
    np.testing.assert_allclose(P[0], [0, 0, 0.5])


def _plane(n, c):
    n = np.asarray(n, np.float64) / np.linalg.norm(n)
    return lambda x: x @ n - c


def _const(v):
    return lambda x: np.full(len(x), v, np.float64)


def test_single_level_perpendicular_plane_hits_in_one_step():
    P = np.array([[0.0, 0.0, 1.0]])
    D = np.array([[0.0, 0.0, -1.0]])
    r = otrace.trace([_plane([0, 0, 1], 0.25)], [], P, D)
    assert r["hit"][0] and r["steps"][0] == 1 and r["t"][0] == 0.75


@pytest.mark.parametrize("theta_deg", [30.0, 60.0, 75.0])
def test_single_level_oblique_plane_converges_geometrically(theta_deg):
    # along a ray at angle theta to the plane normal, F = r cos(theta) for the
    # remaining distance r; each step leaves r (1 - cos theta)
    th = math.radians(theta_deg)
    D = np.array([[math.sin(th), 0.0, -math.cos(th)]])
    P = np.array([[-0.5, 0.0, 0.6]])
    c = math.cos(th)
    r0 = (0.6 - 0.1) / c  # plane z = 0.1
    eps = 1e-4
    k = 0
    while r0 * c * (1 - c) ** k >= eps:  # closed form of the hit step: first k with F_k < eps
        k += 1
    res = otrace.trace([_plane([0, 0, 1], 0.1)], [], P, D, eps=eps, bound_radius=10.0)
    assert res["hit"][0] and res["steps"][0] == k
    assert abs(res["t"][0] - r0 * (1 - (1 - c) ** k)) < 1e-12


def test_single_level_sphere_closed_form_depth():
    # f = |x| - 0.5; rays with impact parameter b hit at t = s - sqrt(0.25 - b^2)
    b = np.array([0.0, 0.1, 0.3, 0.45])
    P = np.stack([b, np.zeros_like(b), np.full_like(b, 1.1)], 1)
    D = np.tile([0.0, 0.0, -1.0], (len(b), 1))
    sph = lambda x: np.linalg.norm(x, axis=1) - 0.5
    r = otrace.trace([sph], [], P, D, eps=1e-7, max_steps=500)
    assert r["hit"].all()
    t_hit = 1.1 - np.sqrt(0.25 - b * b)
    assert np.abs(r["t"] - t_hit).max() < 1e-5
    assert r["steps"][0] == 1


def test_single_level_exit_and_step_limit():
    # moving away from the surface: the ray leaves the bounding sphere
    P = np.array([[0.0, 0.0, 0.3]])
    r = otrace.trace([_plane([0, 0, 1], 0.0)], [], P, np.array([[0.0, 0.0, 1.0]]))
    # 0.3 -> 0.6 -> 1.2 (on the sphere: still inside the 1e-4 relative margin
    # of the exit test) -> 2.4 (outside, moving away)
    assert not r["hit"][0] and abs(r["p"][0, 2] - 2.4) < 1e-12 and r["steps"][0] == 3
    # grazing ray (89.9 deg): the geometric convergence needs more than 16 steps
    th = math.radians(89.9)
    r = otrace.trace([_plane([0, 0, 1], 0.0)], [], np.array([[0.0, 0.0, 0.05]]),
                     np.array([[math.sin(th), 0.0, -math.cos(th)]]), max_steps=16)
    assert not r["hit"][0] and r["steps"][0] == 16


def test_two_levels_lod0_step_then_banded_sum():
    # perpendicular plane F_0 = z - 0.2 from z = 1, delta_0 = 0.05, residual -0.01:
    # lod-0 step F_0 - delta_0 = 0.75 lands at F_0 = 0.05 < 1.4 delta_0 (banded);
    # the banded step F_0 + f_1 = 0.04 lands where the sum is 0
    P = np.array([[0.0, 0.0, 1.0]])
    D = np.array([[0.0, 0.0, -1.0]])
    r = otrace.trace([_plane([0, 0, 1], 0.2), _const(-0.01)], [0.05], P, D)
    assert r["hit"][0] and r["steps"][0] == 2
    assert abs(r["t"][0] - (0.75 + 0.04)) < 1e-12
    assert r["fmin"][0] < 1e-12


def test_residual_phase_steps_by_f0_outside_the_band():
    # F_0 = 0.06 + 2 max(0, x + 1) along +x from x = -1.2, delta_0 = 0.05
    # (band 0.07), f_1 = 0.2: banded at once (step 0.26 to x = -0.94), then
    # outside the band in the residual phase: steps by F_0 = 0.18, 0.54, 1.62
    # and leaves the sphere at x = 1.40.  (A lod-0 step F_0 - delta_0 there
    # would take five steps.)
    f0 = lambda x: 0.06 + 2 * np.maximum(0.0, x[:, 0] + 1.0)
    P = np.array([[-1.2, 0.0, 0.0]])
    D = np.array([[1.0, 0.0, 0.0]])
    r = otrace.trace([f0, _const(0.2)], [0.05], P, D)
    assert not r["hit"][0]
    assert r["steps"][0] == 4
    assert abs(r["t"][0] - (0.26 + 0.18 + 0.54 + 1.62)) < 1e-12
    assert abs(r["p"][0, 0] - 1.40) < 1e-12


def test_every_residual_level_is_added_inside_the_coarse_band():
    # three levels, delta = (0.05, 0.02): plane F_0 = z - 0.2 from z = 1,
    # f_1 = 0.03, f_2 = -0.06.  Inside the coarse band the step is
    # F_0 + f_1 + f_2 even where |F_0 + f_1| >= 1.4 delta_1 (the renderer never
    # reads delta_lod_1): lod-0 step 0.75 (F_0 = 0.05), banded step 0.02
    # (F_0 = 0.03), then the sum is 0: a hit after two steps.
    P = np.array([[0.0, 0.0, 1.0]])
    D = np.array([[0.0, 0.0, -1.0]])
    r = otrace.trace([_plane([0, 0, 1], 0.2), _const(0.03), _const(-0.06)], [0.05, 0.02], P, D)
    assert r["hit"][0] and r["steps"][0] == 2
    assert abs(r["t"][0] - 0.77) < 1e-12


def test_network_levels_are_evaluated_like_callables():
    from oracle import siren as osiren
    rng = np.random.default_rng(3)
    layers = [(rng.uniform(-1, 1, (16, 3)), rng.uniform(-1, 1, 16)), (rng.uniform(-0.1, 0.1, (1, 16)),
                                                                      np.array([0.3]))]
    P = rng.uniform(-0.5, 0.5, (20, 3))
    D = rng.normal(size=(20, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    a = otrace.trace([layers], [], P, D)
    b = otrace.trace([lambda x: osiren.forward(layers, x)], [], P, D)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def test_crossing_records_a_step_over_the_level_set():
    # f = 2 (z - 0.2) is no distance (|grad f| = 2): from z = 1 the step 1.6
    # carries the ray to z = -0.6, across the plane
    P = np.array([[0.0, 0.0, 1.0]])
    D = np.array([[0.0, 0.0, -1.0]])
    r = otrace.trace([lambda x: 2 * (x[:, 2] - 0.2)], [], P, D, max_steps=3)
    assert r["crossed"][0] and not r["hit"][0]
    r = otrace.trace([_plane([0, 0, 1], 0.2)], [], P, D)
    assert not r["crossed"][0] and r["hit"][0]
