"""GPU marching cubes (mp_mc_count / mp_mc_emit) against scikit-image 0.18.3
fixtures and the C oracle: identical triangle sets for identical volumes."""
import os.path as osp

import numpy as np
import pytest
import torch

from conftest import ROOT
from mplicits.meshing import convert_sdf_samples_to_ply, marching_cubes_gpu
from oracle import mc

pytestmark = pytest.mark.gpu
FIX = osp.join(ROOT, "tests", "golden", "mc_lewiner_fixtures.npz")


def _gpu_triangles(vol):
    v, f, n, val = marching_cubes_gpu(torch.from_numpy(np.ascontiguousarray(vol)).cuda())
    v, f = v.cpu().numpy(), f.cpu().numpy()
    assert f.dtype == np.int32
    if len(f):
        assert f.min() >= 0 and f.max() < len(v)
        # every vertex is referenced by some face
        assert len(np.unique(f)) == len(v)
    return v, f, n, val


def test_matches_skimage_fixtures():
    d = np.load(FIX)
    names = sorted({k.split("/")[0] for k in d.files if "/" in k})
    for nm in names:
        ref = mc.canonical(mc.mesh_triangles(d[nm + "/verts"], d[nm + "/faces"]))
        v, f, _, _ = _gpu_triangles(d[nm + "/vol"])
        got = mc.canonical(mc.mesh_triangles(v, f))
        assert got.shape == ref.shape, nm
        np.testing.assert_array_equal(got, ref, err_msg=nm)


def test_deterministic_and_closed():
    n = 96
    g = torch.linspace(-1, 1, n, dtype=torch.float64)
    z, y, x = torch.meshgrid(g, g, g, indexing="ij")
    vol = (torch.sqrt(x * x + y * y + z * z) - 0.6 + 0.05 * torch.sin(9 * x) * torch.sin(7 * y)).float()
    a = marching_cubes_gpu(vol.cuda())
    b = marching_cubes_gpu(vol.cuda())
    for t, u in zip(a, b):
        assert torch.equal(t, u)
    v, f = a[0].cpu().numpy(), a[1].cpu().numpy()
    edges = np.unique(np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1), axis=0)
    assert len(v) - len(edges) + len(f) == 2  # sphere topology
    ref = mc.canonical(mc.triangles(vol.numpy())[0])
    np.testing.assert_array_equal(mc.canonical(mc.mesh_triangles(v, f)), ref)
    nrm = a[2].cpu().numpy()
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-5)


def test_no_surface(capsys):
    vol = torch.ones(8, 8, 8, device="cuda")
    v, f, n, val = marching_cubes_gpu(vol)
    assert v.shape == (0, 3) and f.shape == (0, 3)
    verts, faces, normals, values = convert_sdf_samples_to_ply(vol, [-1, -1, -1], 2 / 7)
    assert "Surface level must be within volume data range." in capsys.readouterr().out
    assert verts.shape == (0, 3) and faces.shape == (0, 3) and values.shape == (0,)


def test_reference_transform():
    # mesh_points = origin + verts * spacing in float64 (i3d/meshing.py:296-305)
    rng = np.random.default_rng(1)
    vol = rng.normal(size=(6, 7, 8)).astype(np.float32)
    h = 2.0 / 7
    verts, faces, _, _ = convert_sdf_samples_to_ply(torch.from_numpy(vol), [-1, -1, -1], h,
                                                    offset=0.25, scale=2.0)
    v, f, _, _ = marching_cubes_gpu(torch.from_numpy(vol).cuda())
    expect = (v.cpu().numpy() * np.r_[[h] * 3] + -1) / 2.0 - 0.25
    assert verts.dtype == np.float64
    np.testing.assert_array_equal(verts, expect)
