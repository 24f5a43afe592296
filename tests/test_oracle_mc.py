"""Pins the C Lewiner restatement (oracle/mc_lewiner.c) against scikit-image
0.18.3 outputs (tests/golden/mc_lewiner_fixtures.npz, made by
tests/golden/make_mc_fixtures.py) and against mesh invariants."""
import os.path as osp

import numpy as np
import pytest

from conftest import ROOT
from oracle import mc

FIX = osp.join(ROOT, "tests", "golden", "mc_lewiner_fixtures.npz")


@pytest.fixture(scope="module")
def fixtures():
    d = np.load(FIX)
    names = sorted({k.split("/")[0] for k in d.files if "/" in k})
    return d, names


def test_fixture_provenance(fixtures):
    d, names = fixtures
    assert str(d["skimage_version"]) == "0.18.3"
    assert len([n for n in names if n.startswith("cube")]) == 256 * 8


def test_triangle_sets_match_skimage(fixtures):
    # every fixture, including volumes with corners exactly at the level
    d, names = fixtures
    for nm in names:
        ref = mc.canonical(mc.mesh_triangles(d[nm + "/verts"], d[nm + "/faces"]))
        got = mc.canonical(mc.triangles(d[nm + "/vol"])[0])
        assert ref.shape == got.shape, nm
        np.testing.assert_array_equal(got, ref, err_msg=nm)


def _euler(keys):
    tri = keys.reshape(-1, 3)
    verts = np.unique(tri)
    edges = np.unique(np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1), axis=0)
    return len(verts) - len(edges) + len(tri), tri, edges


def test_sphere_is_a_closed_oriented_surface():
    n = 40
    g = np.linspace(-1, 1, n)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    vol = (np.sqrt(x * x + y * y + z * z) - 0.55).astype(np.float32)
    pos, key = mc.triangles(vol)
    chi, tri, edges = _euler(mc.global_vertex_ids(key, vol.shape))
    assert chi == 2
    # every edge is shared by exactly two triangles with opposite directions
    directed = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    assert len(np.unique(directed, axis=0)) == len(directed)
    assert len(directed) == 2 * len(edges)
    # consistent winding: in (axis0, axis1, axis2) coordinates the right-hand
    # normal of every face points toward increasing values (outward for an SDF)
    c = pos.mean(1) - (n - 1) / 2
    nrm = np.cross(pos[:, 1] - pos[:, 0], pos[:, 2] - pos[:, 0])
    assert np.all(np.sum(nrm * c, 1) > 0)
    # every vertex lies on a sign-changing grid edge at the interpolated point
    assert np.all(key % 13 < 12)


def test_vertex_count_equals_sign_changes():
    rng = np.random.default_rng(3)
    vol = rng.normal(size=(9, 10, 11)).astype(np.float32)
    _, key = mc.triangles(vol)
    gid = np.unique(mc.global_vertex_ids(key[key % 13 < 12], vol.shape))
    pos = vol > 0
    n_changes = (pos[:, :, 1:] != pos[:, :, :-1]).sum() + (pos[:, 1:] != pos[:, :-1]).sum() + (pos[1:] != pos[:-1]).sum()
    assert len(gid) == n_changes
