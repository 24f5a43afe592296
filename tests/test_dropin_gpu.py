"""The Python drop-ins called by their reference names: create_mesh_multistage
(i3d/meshing.py:135-264), the reconstruct.py CLI with adaptive_deltas
(reconstruct.py:18-133), the Eq. 5 table of appendix_experiments/
extract_deltas.py and true_sdf (i3d/loss_functions.py:123-163), against
the oracle."""
import os.path as osp
import importlib.util

import numpy as np
import pytest
import torch

from conftest import ROOT
from mplicits import loss_functions, meshing
from mplicits.model import SIREN, from_pth
from mplicits.ply import read_ply
from mplicits.synthetic import DELTAS, residual_levels, seeded_siren
from oracle import losses as olosses
from oracle import mc as omc
from oracle import siren as osiren

pytestmark = pytest.mark.gpu

COARSE = osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")
CLOUD = osp.join(ROOT, "tests", "golden", "bumpy_surface_2k.ply")


def _levels():
    coarse = from_pth(COARSE).cuda().eval()
    med, fine = residual_levels("cuda")
    return [coarse, med, fine]


def _reconstruct_module():
    spec = importlib.util.spec_from_file_location("reconstruct_cli", osp.join(ROOT, "m-plicits_amd", "reconstruct.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_create_mesh_multistage_prints_returns_and_writes_the_ply(tmp_path, capsys):
    levels = _levels()
    N = 48
    out = str(tmp_path / "mesh.ply")
    res = meshing.create_mesh_multistage(levels, deltas=list(DELTAS), filename=out, N=N, max_batch=32 ** 3,
                                         device="cuda", return_volume=True)
    verts, faces, normals, values, t, vol, counts = res
    printed = capsys.readouterr().out.splitlines()
    assert printed[0].startswith("Sampling took: ") and printed[0].endswith(" s")
    assert printed[1] == f"Saving mesh to {out}" and printed[2] == "Done"
    assert isinstance(t, float) and t > 0
    assert counts[0] == N ** 3 and 0 < counts[2] < counts[1]
    # the reference's 5-tuple (without return_volume)
    five = meshing.create_mesh_multistage(levels, deltas=list(DELTAS), N=N, max_batch=32 ** 3, silent=True)
    assert len(five) == 5 and np.array_equal(five[1], faces) and np.array_equal(five[0], verts)
    assert capsys.readouterr().out == ""  # silent
    # marching cubes of the returned volume: the oracle's triangle set, in
    # world units (spacing 2 / (N - 1), origin -1)
    h = 2.0 / (N - 1)
    tri = omc.canonical(omc.mesh_triangles((verts + 1.0) / h, faces))
    want = omc.canonical(omc.triangles(vol.cpu().numpy())[0])
    assert tri.shape == want.shape and np.abs(tri - want).max() < 1e-4
    assert verts.dtype == np.float64 and faces.shape[1] == 3 and normals.shape == verts.shape
    assert values.shape == (len(verts),)
    # the PLY holds the returned mesh (float32 vertices)
    ply = read_ply(out)
    v = ply["vertex"]
    np.testing.assert_array_equal(np.stack([v["x"], v["y"], v["z"]], 1), verts.astype(np.float32))
    np.testing.assert_array_equal(ply["face"], faces)
    assert open(out, "rb").read(36) == b"ply\nformat binary_little_endian 1.0\n"


def test_create_mesh_multistage_empty_level_set_and_rejections(capsys):
    m = SIREN(3, 1, [64, 64], w0=1, ww=1, delay_init=True)
    with torch.no_grad():
        for p in m.parameters():
            p.zero_()
        m.net[-1][0].bias.fill_(1.0)  # f = 1 everywhere: 0 is outside the value range
    m = m.cuda()
    verts, faces, normals, values, _ = meshing.create_mesh_multistage([m], N=16)
    assert "Surface level must be within volume data range." in capsys.readouterr().out
    assert verts.shape == (0, 3) and faces.shape == (0, 3) and values.shape == (0,)
    with pytest.raises(NotImplementedError):
        meshing.create_mesh_multistage([m], N=16, t=0.5)
    with pytest.raises(ValueError):
        meshing.create_mesh_multistage([m, m], deltas=[0.1, 0.2], N=16)


def test_reconstruct_cli(tmp_path, capsys):
    levels = _levels()
    paths = {}
    for name, m in zip(("coarse", "medium", "fine"), levels):
        paths[name] = str(tmp_path / f"{name}.pth")
        torch.save(m.state_dict(), paths[name])
    cli = _reconstruct_module()
    out = str(tmp_path / "out" / "mesh.ply")
    cli.main([paths["fine"], out, "--coarse_path", paths["coarse"], "--medium_path", paths["medium"], "-r", "40",
              "--device", "cuda", "--multistage", "--deltas", "0.05", "0.02"])
    text = capsys.readouterr().out
    for line in ("Running marching cubes running with resolution 40", "Using residual approach for fine level",
                 "========== Medium model ==========", "Using the residual approach",
                 "========== Coarse model ==========", f"Saving mesh to {out}"):
        assert line in text, line
    lines = text.splitlines()
    assert lines[-1] == "Done" and lines[-2].startswith("SAMPLING_TIME_S:")
    assert float(lines[-2].split(":")[1]) > 0
    # the CLI's mesh is create_mesh_multistage's (coarse first, 32^3 chunks)
    want = meshing.create_mesh_multistage(levels, deltas=[0.05, 0.02], N=40, max_batch=32 ** 3, silent=True)
    ply = read_ply(out)
    np.testing.assert_array_equal(ply["face"], want[1])
    # --multistage without --input / --deltas: the fixed widths and a warning
    cli.main([paths["fine"], str(tmp_path / "m2.ply"), "--coarse_path", paths["coarse"], "--medium_path",
              paths["medium"], "-r", "24", "--device", "cuda", "--multistage"])
    assert "[WARNING] --multistage without --input: using fixed band widths [0.1, 0.06]" in capsys.readouterr().out


def test_adaptive_deltas_against_the_oracle(tmp_path, capsys):
    """Eq. 5: delta_i = 1.3 max_j |F_i(x_j)| over the training point cloud,
    F_i the partial sum of the levels before i; through --input as well."""
    levels = _levels()
    cli = _reconstruct_module()
    got = cli.adaptive_deltas(levels, CLOUD, "cuda")
    v = read_ply(CLOUD)["vertex"]
    pts = np.stack([v["x"], v["y"], v["z"]], 1).astype(np.float64)
    layers = [osiren.layers_from_model(m) for m in levels]
    partial = np.zeros(len(pts))
    want = []
    for lay in layers[:-1]:
        partial = partial + osiren.forward(lay, pts)
        want.append(1.3 * np.abs(partial).max())
    assert len(got) == 2
    np.testing.assert_allclose(got, want, rtol=1e-5)
    paths = {}
    for name, m in zip(("coarse", "medium", "fine"), levels):
        paths[name] = str(tmp_path / f"{name}.pth")
        torch.save(m.state_dict(), paths[name])
    cli.main([paths["fine"], str(tmp_path / "m.ply"), "--coarse_path", paths["coarse"], "--medium_path",
              paths["medium"], "-r", "24", "--device", "cuda", "--multistage", "--input", CLOUD])
    text = capsys.readouterr().out
    assert "adaptive band widths (Eq. 5): " + ", ".join(f"{d:.3e}" for d in got) in text


def test_extract_deltas_cli_against_the_oracle(tmp_path, capsys):
    """The appendix table of Eq. 5 band widths (appendix_experiments/
    extract_deltas.py): checkpoints found under the i3d layout, a shape
    without them skipped, max |f_1| and max |f_1 + r_1| over the input points
    against the oracle's FP64 partial sums, the printed line and the table."""
    import shutil
    spec = importlib.util.spec_from_file_location("extract_deltas_cli",
                                                  osp.join(ROOT, "m-plicits_amd", "extract_deltas.py"))
    cli = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cli)
    coarse, med, _ = _levels()
    for stage, m in (("coarse", coarse), ("medium", med)):
        d = tmp_path / "results" / "bumpy" / stage
        d.mkdir(parents=True)
        torch.save(m.state_dict(), str(d / "best.pth"))
    (tmp_path / "results" / "half" / "coarse").mkdir(parents=True)  # no medium checkpoint
    torch.save(coarse.state_dict(), str(tmp_path / "results" / "half" / "coarse" / "best.pth"))
    inp = tmp_path / "data" / "normalized_sphere" / "input"
    inp.mkdir(parents=True)
    shutil.copy(CLOUD, str(inp / "bumpy.ply"))
    table = tmp_path / "deltas_table.md"
    rows = cli.main(["--root", str(tmp_path), "--shapes", "bumpy", "half", "--out", str(table)])
    out = capsys.readouterr().out
    v = read_ply(CLOUD)["vertex"]
    pts = np.stack([v["x"], v["y"], v["z"]], 1).astype(np.float64)
    f1 = osiren.forward(osiren.layers_from_model(coarse), pts)
    r1 = osiren.forward(osiren.layers_from_model(med), pts)
    m1, m2 = np.abs(f1).max(), np.abs(f1 + r1).max()
    assert len(rows) == 1 and rows[0][0] == "bumpy" and rows[0][1] == len(pts)
    np.testing.assert_allclose(rows[0][2:], [m1, 1.3 * m1, m2, 1.3 * m2], rtol=1e-5)
    assert "[skip] half: missing checkpoints" in out
    r = rows[0]
    assert f"bumpy: n={len(pts)}  max|f1|={r[2]:.3e} d1={r[3]:.3e}  max|f1+r1|={r[4]:.3e} d2={r[5]:.3e}" in out
    text = table.read_text()
    assert text.startswith("# Adaptive band widths (Eq. 5, eps=0.3) from paper checkpoints")
    assert f"| bumpy | {len(pts)} | {r[2]:.3e} | {r[3]:.3e} | {r[4]:.3e} | {r[5]:.3e} |" in text


def test_true_sdf_against_the_oracle():
    """The coarse trainer's loss (terms and parameter gradients through the
    create_graph input gradient) against FP64 autograd."""
    torch.manual_seed(3)
    model = SIREN(3, 1, [64, 64, 64], w0=30, ww=30).cuda()
    g = torch.Generator().manual_seed(4)
    n = 3000
    x = torch.rand(n, 3, generator=g) * 2 - 1
    gt = torch.rand(n, 1, generator=g) * 0.2 - 0.1
    gt[::2] = 0
    nrm = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=-1)
    out = model(x.cuda())
    d = loss_functions.true_sdf(out, {"sdf": gt.cuda(), "normals": nrm.cuda()})
    model.zero_grad()
    sum(d.values()).backward()
    lay = [(torch.from_numpy(w).requires_grad_(True), torch.from_numpy(b).requires_grad_(True))
           for w, b in osiren.layers_from_model(model)]
    ref = olosses.true_sdf(lay, x.double(), gt.double(), nrm.double())
    for k in ref:
        assert abs(d[k].item() - ref[k].item()) <= 1e-5 * abs(ref[k].item()) + 1e-7, k
    grads = torch.autograd.grad(sum(ref.values()), [p for wb in lay for p in wb])
    oms = [30.0] * 3 + [1.0]
    for k, p in enumerate(model.parameters()):
        r = grads[k] * oms[k // 2]
        err = ((p.grad.double().cpu() - r).abs().max() / r.abs().max()).item()
        assert err <= 1e-3, (tuple(p.shape), err)
