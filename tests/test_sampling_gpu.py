"""Training-data generation on the GPU (mplicits.sampling: mp_mesh_*,
mp_calc_deltas, mp_neighborhood_samples) against closed forms and the FP64 /
float32 restatements in oracle/meshdist.py."""
import numpy as np
import pytest
import torch

from mplicits import sampling
from oracle import meshdist as om
from test_oracle_meshdist import box_distance, box_mesh

pytestmark = pytest.mark.gpu


def wavy_sphere(n_lat, n_lon, inward=False):
    """Closed UV sphere r = 0.5 + 0.05 sin(5 theta) cos(3 phi): vertices
    [V, 6] (position, area-weighted vertex normal) and faces [F, 3]."""
    th = np.linspace(0, np.pi, n_lat + 1)[1:-1]
    ph = np.linspace(0, 2 * np.pi, n_lon, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    r = 0.5 + 0.05 * np.sin(5 * T) * np.cos(3 * P)
    ring = np.stack([r * np.sin(T) * np.cos(P), r * np.sin(T) * np.sin(P), r * np.cos(T)], -1).reshape(-1, 3)
    V = np.vstack([ring, [[0, 0, 0.5]], [[0, 0, -0.5]]])
    top, bot = len(ring), len(ring) + 1
    F = []
    for i in range(n_lat - 1):
        for j in range(n_lon):
            a, b = i * n_lon + j, i * n_lon + (j + 1) % n_lon
            if i + 1 < n_lat - 1:
                c, d = a + n_lon, b + n_lon
                F += [(a, c, b), (b, c, d)]
    for j in range(n_lon):
        F.append((top, j, (j + 1) % n_lon))
        k = (n_lat - 2) * n_lon
        F.append((bot, k + (j + 1) % n_lon, k + j))
    F = np.array(F)
    fn = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    N = np.zeros_like(V)
    for k in range(3):
        np.add.at(N, F[:, k], fn)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    if inward:
        N = -N
    return np.hstack([V, N]), F


def _scene(V, F):
    s = sampling.RaycastingScene()
    s.add_triangles(V[:, :3], F)
    return s


def test_box_distance_closed_form():
    lo, hi = (-0.5, -0.3, -0.1), (0.4, 0.6, 0.2)
    V, F = box_mesh(lo, hi)
    s = _scene(V, F)
    rng = np.random.default_rng(3)
    P = rng.uniform(-1.2, 1.2, (20000, 3))
    P[:5000] = rng.uniform(lo, hi, (5000, 3))
    got = s.compute_distance(torch.tensor(P, dtype=torch.float32, device="cuda")).cpu().numpy()
    assert np.abs(got - box_distance(P.astype(np.float32).astype(np.float64), lo, hi)).max() < 2e-6


@pytest.mark.parametrize("n_lat,n_lon", [(12, 20), (90, 160)])
def test_mesh_distance_matches_brute_force(n_lat, n_lon):
    V, F = wavy_sphere(n_lat, n_lon)
    s = _scene(V, F)
    tris, nodes, depth = s.info()
    assert tris == len(F) and depth <= 48
    rng = np.random.default_rng(n_lat)
    P = np.vstack([rng.uniform(-0.8, 0.8, (300, 3)), V[rng.choice(len(V), 100), :3] * rng.uniform(0.9, 1.1, (100, 1))])
    P32 = P.astype(np.float32)
    got = s.compute_distance(torch.tensor(P32, device="cuda")).cpu().numpy()
    ref = om.distance(V[:, :3].astype(np.float32), F, P32.astype(np.float64), chunk=16)
    assert np.abs(got - ref).max() < 2e-6, np.abs(got - ref).max()


def test_calc_deltas_slab_closed_form():
    V, F = box_mesh((-1, -1, -0.1), (1, 1, 0.1))
    verts = np.array([[x, y, z, 0.0, 0.0, -3.0 * np.sign(z)] for x in (-0.2, 0.0, 0.3) for y in (-0.1, 0.25)
                      for z in (-0.1, 0.1)], np.float32)
    out = sampling.calc_deltas(torch.tensor(verts, device="cuda"), _scene(V, F), max_delta=0.5, n_iters=100)
    assert sampling.calc_deltas.iterations == 100
    out = out.cpu().numpy()
    assert np.array_equal(out[:, :3], verts[:, :3])
    assert np.allclose(out[:, 3:6], verts[:, 3:6] / 3.0)
    assert np.all(np.abs(out[:, 6] - 0.1) < 2e-4)


def test_calc_deltas_matches_oracle():
    V, F = wavy_sphere(24, 40, inward=True)
    verts = V.astype(np.float32)
    out = sampling.calc_deltas(torch.tensor(verts, device="cuda"), _scene(V, F), max_delta=0.8, n_iters=60)
    ref, iters = om.calc_deltas(verts[:, :3].astype(np.float64), F, verts.astype(np.float64), max_delta=0.8,
                                n_iters=60)
    assert sampling.calc_deltas.iterations == iters == 60
    got = out.cpu().numpy()
    # the search takes the same branches except where |delta - distance| is
    # within rounding of eps
    close = np.abs(got[:, 6] - ref[:, 6]) < 1e-5
    assert close.mean() > 0.98, close.mean()
    assert np.all((got[:, 6] > 0) & (got[:, 6] < 0.8))


@pytest.mark.parametrize("n_on,use,dither", [(300, False, True), (300, True, True), (500, True, False),
                                             (800, False, True)])
def test_neighborhood_samples_bit_exact(n_on, use, dither):
    V, F = wavy_sphere(20, 25)
    rng = np.random.default_rng(5)
    verts = np.hstack([V, rng.uniform(0.001, 0.02, (len(V), 1))]).astype(np.float32)  # 7 columns
    vt = torch.tensor(verts, device="cuda")
    torch.manual_seed(11)
    got = sampling.create_training_data_on_neighborhood(vt, n_on, device="cuda", delta=0.03, use_neigh_delta=use,
                                                        dithering_sampling=dither)["on_surf"]
    # the same draws, as the reference makes them
    torch.manual_seed(11)
    nv = len(verts)
    idx = None if n_on == nv else torch.randperm(nv, device="cuda")[:n_on].cpu().numpy()
    rnd = torch.rand((min(n_on, nv), 3)).numpy()
    ref = om.neighborhood_samples(verts, idx, rnd, 0.03, use_neigh_delta=use, dithering=dither)
    for g, r in zip(got, ref):
        assert np.array_equal(g.cpu().numpy(), r)
