"""Closed-form checks of the training-data oracle (oracle/meshdist.py):
distances to a box mesh and to one triangle in each Voronoi region, the
tubular-neighbourhood search on a slab (converges to its half thickness) and
the invariants of the band sampler."""
import numpy as np
import pytest

from oracle import meshdist as om


def box_mesh(lo, hi):
    """Closed triangle mesh of the axis-aligned box [lo, hi] (12 triangles)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    V = np.array([[hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]]
                  for i in range(8)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    F = np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))])
    return V, F


def box_distance(P, lo, hi):
    """Closed form: unsigned distance to the surface of the box."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    q = np.maximum(np.maximum(lo - P, P - hi), 0.0)
    outside = np.sqrt((q * q).sum(1))
    inside = np.minimum(P - lo, hi - P).min(1)
    return np.where((q > 0).any(1), outside, inside)


def test_box_distances():
    lo, hi = (-0.5, -0.3, -0.1), (0.4, 0.6, 0.2)
    V, F = box_mesh(lo, hi)
    rng = np.random.default_rng(1)
    P = rng.uniform(-1.2, 1.2, (500, 3))
    P[:100] = rng.uniform(lo, hi, (100, 3))  # inside points: distance to the nearest face
    assert np.allclose(om.distance(V, F, P), box_distance(P, lo, hi), atol=1e-12)


@pytest.mark.parametrize("p,d", [
    ((0.2, 0.2, 0.5), 0.5),                 # above the face
    ((-1.0, -1.0, 0.0), np.sqrt(2.0)),      # beyond vertex A
    ((0.5, -0.3, 0.4), 0.5),                # beyond edge AB: distance to (0.5, 0, 0)
    ((1.0, 1.0, 0.0), np.sqrt(0.5)),        # beyond the hypotenuse BC
    ((3.0, -1.0, 2.0), np.sqrt(4 + 1 + 4)),  # beyond vertex B = (1, 0, 0)
])
def test_single_triangle_regions(p, d):
    V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert om.distance(V, np.array([[0, 1, 2]]), np.array([p]))[0] == pytest.approx(d, abs=1e-12)


def test_calc_deltas_on_a_slab_converge_to_half_the_thickness():
    # a thin box, normals at the top and bottom vertices pointing into it: the
    # offset whose distance to the mesh equals itself is the half thickness
    V, F = box_mesh((-1, -1, -0.1), (1, 1, 0.1))
    verts = np.array([[x, y, z, 0.0, 0.0, -3.0 * np.sign(z)] for x in (-0.2, 0.0, 0.3) for y in (-0.1, 0.25)
                      for z in (-0.1, 0.1)])
    out, iters = om.calc_deltas(V, F, verts, max_delta=0.5, n_iters=100)
    assert iters == 100
    assert np.allclose(out[:, 3:6], verts[:, 3:6] / 3.0)
    assert np.all(np.abs(out[:, 6] - 0.1) < 2e-4)


def test_calc_deltas_stops_when_a_delta_reaches_the_maximum():
    V, F = box_mesh((-1, -1, -1), (1, 1, 1))
    # outward normal at a face centre: the distance always equals the offset
    verts = np.array([[0.0, 0.0, 1.0, 0.0, 0.0, 1.0]])
    out, iters = om.calc_deltas(V, F, verts, max_delta=0.05, n_iters=100)
    # delta grows by one step (5e-4) per iteration from 5e-4 and reaches 0.05
    # after 99 of them
    assert iters == 99 and out[0, 6] == pytest.approx(0.05)


def test_neighborhood_sample_invariants():
    rng = np.random.default_rng(2)
    n = rng.normal(size=(40, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    verts = np.hstack([rng.uniform(-1, 1, (40, 3)), n, rng.uniform(0.001, 0.01, (40, 1))]).astype(np.float32)
    idx = rng.permutation(40)[:25]
    rnd = rng.uniform(0, 1, (25, 3)).astype(np.float32)
    for use in (False, True):
        pts, nrm, sdf = om.neighborhood_samples(verts, idx, rnd, 0.03, use_neigh_delta=use)
        assert pts.shape == (75, 3) and sdf.shape == (75,)
        assert np.array_equal(pts[:25], verts[idx, :3]) and np.all(sdf[:25] == 0)
        # offset points lie at distance sdf along the unit normal
        assert np.allclose(np.linalg.norm(pts[25:50] - pts[:25], axis=1), sdf[25:50], atol=1e-6)
        bound = verts[idx, 6] if use else np.full(25, 0.0005)
        assert np.all((sdf[25:50] >= 0) & (sdf[25:50] <= bound))
        # dithered points: inside the delta-cube at p + delta / 2 n, labelled -1
        c = pts[:25] + 0.015 * nrm[:25]
        assert np.all((pts[50:] - c >= -1e-6) & (pts[50:] - c <= 0.03 + 1e-6)) and np.all(sdf[50:] == -1)
    pts, _, sdf = om.neighborhood_samples(verts, None, rng.uniform(0, 1, (40, 3)), 0.03, dithering=False)
    assert pts.shape == (80, 3) and np.array_equal(pts[:40], verts[:, :3])
