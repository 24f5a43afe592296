"""The workload configurations end to end on the GPU, against the oracle.

1. The fitted coarse level (256x3, w0 30; tests/golden/bumpy_coarse_256x3.pth):
   dense 128^3 grid without culling, marching cubes.
2. The three-level stack (fitted coarse > 256x3 w60, delta 0.05 > 512x3 w90,
   delta 0.02; residual last layers x 0.1): band-culled extraction in 32^3
   chunks, at 64^3 in full and at 512^3 on sampled chunks, per precision; the
   mixed-precision meshes against the FP32 mesh.
4. The training step of the fine 512x3 residual over the frozen levels: FP32
   parameter gradients of the benchmark loss against FP64 autograd.
5. Twelve models of the uneven batch (widths 64-512, 1-3 levels, own targets)
   at reduced grids and images, every one against the oracle.
(3, tracing, is covered in test_trace_gpu.py.)
"""
import os.path as osp

import numpy as np
import pytest
import torch

from checks import check_trace_against_oracle, check_volume_against_oracle
from conftest import ROOT
from mplicits import training
from mplicits.meshing import marching_cubes_gpu
from mplicits.model import from_pth
from mplicits.stack import LevelStack
from mplicits.synthetic import DELTAS, MASTER_SEED, BumpySphere, residual_levels, uneven_workload
from oracle import extract as oextract
from oracle import losses as olosses
from oracle import mc as omc
from oracle import siren as osiren

pytestmark = pytest.mark.gpu

COARSE = osp.join(ROOT, "tests", "golden", "bumpy_coarse_256x3.pth")
CHUNK = 32 ** 3
# max |SDF - FP64| stated per precision (samples whose band decision flips are
# counted apart); the north star's bound for the mixed-precision path is 1e-4
SDF_TOL = {"fp32": 1e-5, "bf16x3": 1e-4, "bf16x3f": 1e-4, "fp16": 5e-4, "bf16": 3e-3}


def _three_levels():
    coarse = from_pth(COARSE).cuda().eval()
    med, fine = residual_levels("cuda")
    return [coarse, med, fine]


def _surface_chunks(vol, N, k):
    flat = vol.reshape(-1)
    near = torch.nonzero(flat.abs() < 2.0 / N).flatten()
    pick = near[torch.linspace(0, len(near) - 1, k).long()].cpu().numpy()
    return sorted(set(int(i) // CHUNK for i in pick))


def test_config1_fitted_coarse_dense_grid_and_mesh():
    coarse = from_pth(COARSE).cuda().eval()
    N = 128
    vol, counts = LevelStack([coarse], "fp32").extract_volume(N, [], chunk=CHUNK)
    assert counts == [N ** 3]
    chunks = sorted(set(_surface_chunks(vol, N, 5)) | {0, 31, 63})
    emax, emean, _, _ = check_volume_against_oracle(vol, [coarse], N, [], CHUNK, chunks, "fp32", SDF_TOL["fp32"])
    print(f"config 1 fp32: max {emax:.2e} mean {emean:.2e} over {len(chunks)} chunks")
    # marching cubes: the triangle set of the C Lewiner restatement on the same grid
    v, f, _, _ = marching_cubes_gpu(vol)
    got = omc.canonical(omc.mesh_triangles(v.cpu().numpy(), f.cpu().numpy()))
    want = omc.canonical(omc.triangles(vol.cpu().numpy())[0])
    assert got.shape == want.shape and np.array_equal(got, want)
    # a closed surface of sphere topology: V - E + F = 2
    fa = f.cpu().numpy().astype(np.int64)
    edges = np.unique(np.sort(np.concatenate([fa[:, [0, 1]], fa[:, [1, 2]], fa[:, [2, 0]]]), axis=1), axis=0)
    assert len(np.unique(fa)) - len(edges) + len(fa) == 2


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x3f", "fp16", "bf16"])
def test_config2_three_level_band_culled_64(precision):
    levels = _three_levels()
    N = 64
    vol, counts = LevelStack(levels, precision).extract_volume(N, list(DELTAS), chunk=CHUNK)
    assert counts[0] == N ** 3 and 0 < counts[2] < counts[1] < N ** 3
    emax, emean, flips, total = check_volume_against_oracle(vol, levels, N, DELTAS, CHUNK, range(N ** 3 // CHUNK),
                                                            precision, SDF_TOL[precision])
    print(f"config 2 at 64^3 {precision}: max {emax:.2e} mean {emean:.2e}, {flips} flips of {total}, "
          f"counts {counts}")


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x3f"])
def test_config2_three_level_band_culled_512_sampled(precision):
    levels = _three_levels()
    N = 512
    vol, counts = LevelStack(levels, precision).extract_volume(N, list(DELTAS), chunk=CHUNK)
    chunks = _surface_chunks(vol, N, 3) + [0]
    layers = [osiren.layers_from_model(m) for m in levels]
    ref = oextract.sample_chunks(layers, N, list(DELTAS), CHUNK, chunks, return_masks=True)
    flat = vol.reshape(-1).double().cpu().numpy()
    for c, (idx, r, masks) in ref.items():
        e = np.abs(flat[idx] - r)
        # at 512^3 the decisions of the GPU are not recomputed: a flipped sample
        # differs by a whole residual, and the reference value is then within
        # tol of a band edge
        ok = e <= SDF_TOL[precision]
        assert ok.mean() >= 0.999, (c, ok.mean())
    # band counts of the full grid (n1 ~ 3.9%, n2 ~ 1.6% for the bumpy sphere)
    assert counts[0] == N ** 3
    assert 0.02 < counts[1] / N ** 3 < 0.06 and 0.008 < counts[2] / N ** 3 < 0.03, counts


def _mesh(vol, N):
    v, _, nrm, _ = marching_cubes_gpu(vol, 0.0)
    return v.double().cpu().numpy() * (2.0 / (N - 1)) - 1.0, nrm.double().cpu().numpy()


def test_config2_mixed_precision_mesh_tolerances():
    """Meshes of the mixed-precision volumes against the FP32 mesh (256^3):
    Chamfer distance (mean of the two directed nearest-vertex means, world
    units; one voxel is 7.8e-3) and the angle between matched MC normals."""
    from scipy.spatial import cKDTree
    levels = _three_levels()
    N = 256
    ref = _mesh(LevelStack(levels, "fp32").extract_volume(N, list(DELTAS), chunk=CHUNK)[0], N)
    # stated tolerances: (chamfer, mean normal angle in degrees)
    tol = {"bf16x3": (5e-6, 0.02), "bf16x3f": (5e-5, 0.15), "fp16": (1e-4, 0.3), "bf16": (5e-4, 1.5)}
    for prec, (tc, ta) in tol.items():
        m = _mesh(LevelStack(levels, prec).extract_volume(N, list(DELTAS), chunk=CHUNK)[0], N)
        da, ia = cKDTree(ref[0]).query(m[0])
        db, _ = cKDTree(m[0]).query(ref[0])
        chamfer = (da.mean() + db.mean()) / 2
        na, nb = m[1], ref[1][ia]
        c = (na * nb).sum(1) / np.maximum(np.linalg.norm(na, axis=1) * np.linalg.norm(nb, axis=1), 1e-30)
        ang = np.degrees(np.arccos(np.clip(c, -1, 1))).mean()
        print(f"{prec}: chamfer {chamfer:.2e}, normal angle mean {ang:.3f} deg")
        assert chamfer <= tc and ang <= ta, (prec, chamfer, ang)


def _config4_batch(levels, n, seed):
    """Half uniform in [-1, 1]^3, half inside the fine band |F_c + F_m| < 0.02
    (rejection sampling on the frozen levels), with the analytic targets."""
    coarse, med = levels[:2]
    frozen = LevelStack([coarse, med], "fp32")
    g = torch.Generator(device="cuda").manual_seed(seed)
    uni = torch.rand(n // 2, 3, device="cuda", generator=g) * 2 - 1
    band = []
    got = 0
    while got < n - n // 2:
        cand = torch.rand(1 << 16, 3, device="cuda", generator=g) * 2 - 1
        keep = cand[frozen.eval(cand).abs() < DELTAS[1]]
        band.append(keep)
        got += len(keep)
    x = torch.cat([uni, torch.cat(band)[:n - n // 2]])
    shape = BumpySphere()
    return x, shape.sdf(x).float(), shape.normals(x).float()


def test_config4_fine_level_gradients_fp32():
    """One training step of the fine 512x3 residual (w 90) over the frozen
    fitted coarse and medium levels, the benchmark loss (L1 SDF + eikonal +
    normal cosine) on 4096 points: FP32 parameter gradients within 1e-3
    (relative, per parameter tensor) of FP64 autograd."""
    from mplicits.model import SIREN
    coarse, med, fine_baked = _three_levels()
    n = 4096
    x, f_t, n_t = _config4_batch([coarse, med], n, MASTER_SEED + 40)
    frozen = LevelStack([coarse, med], "fp32")
    # the trainer's module: the fine level before baking (w0 = ww = 90)
    fine = SIREN(3, 1, [512] * 3, w0=90, ww=90, delay_init=True).cuda()
    with torch.no_grad():
        for (a, b) in zip(fine.parameters(), fine_baked.parameters()):
            a.copy_(b)
        for seq in fine.net[:-1]:  # the baked tensors / 90: the same function
            seq[0].weight.div_(90.0)
            seq[0].bias.div_(90.0)
    loss, _, _ = training.train_step(fine, x, lambda F, G: training.config_loss(F, G, f_t, n_t), frozen=frozen,
                                     precision="fp32")
    fl = [(torch.from_numpy(w).requires_grad_(True), torch.from_numpy(b).requires_grad_(True))
          for w, b in osiren.layers_from_model(fine)]
    tl = lambda m: [(torch.from_numpy(w), torch.from_numpy(b)) for w, b in osiren.layers_from_model(m)]  # noqa: E731
    xd = x.double().cpu()
    ref_loss = olosses.config_loss(fl, tl(med), tl(coarse), xd, f_t.double().cpu()[:, None], n_t.double().cpu())
    ref = torch.autograd.grad(ref_loss, [p for wb in fl for p in wb])
    assert abs(loss.item() - ref_loss.item()) <= 1e-5 * abs(ref_loss.item()) + 1e-5
    oms = [90.0] * 3 + [1.0]
    for k, p in enumerate(fine.parameters()):
        r = ref[k] * oms[k // 2]  # d/dW = omega d/dW_eff
        g = p.grad.double().cpu()
        err = ((g - r).abs().max() / r.abs().max()).item()
        print(f"{tuple(p.shape)}: max rel err {err:.2e}")
        assert err <= 1e-3, (tuple(p.shape), err)


def test_config4_fine_level_gradients_fp16_over_bf16x3_levels():
    """The same fine-level step with FP16 training GEMMs over frozen levels
    evaluated in three-term BF16, under the reference's smooth residual loss
    (loss_add_detail_fine) on 16,384 points: every parameter tensor's gradient
    within 1e-3 (relative, ||g - g_64|| / ||g_64||) of FP64 autograd, and all
    parameters within 3e-4.  With the frozen levels in FP16 their evaluation
    error (~2e-4 in F) reaches the seeds and the first layers' gradients
    (~2.5e-3); that figure is printed, not asserted."""
    from mplicits.model import SIREN
    coarse, med, fine_baked = _three_levels()
    n = 1 << 14
    x, f_t, n_t = _config4_batch([coarse, med], n, MASTER_SEED + 41)
    x = x.contiguous()
    tl = lambda m: [(torch.from_numpy(w).cuda(), torch.from_numpy(b).cuda())  # noqa: E731
                    for w, b in osiren.layers_from_model(m)]
    fl = [(w.requires_grad_(True), b.requires_grad_(True)) for w, b in tl(fine_baked)]
    xd, fd, nd = x.double(), f_t.double()[:, None], n_t.double()
    ref_loss = sum(olosses.loss_add_detail_fine(fl, tl(med), tl(coarse), xd, fd, nd).values())
    ref = torch.autograd.grad(ref_loss, [p for wb in fl for p in wb])

    def loss_fn(F, G):
        cos = torch.nn.functional.cosine_similarity(n_t, G, dim=-1)
        return 3e3 * ((f_t - F) ** 2).mean() + 1e2 * (1 - cos).mean() + 5e1 * ((G.norm(dim=-1) - 1) ** 2).mean()

    oms = [90.0] * 3 + [1.0]
    for frozen_prec in ("bf16x3", "fp16"):
        fine = SIREN(3, 1, [512] * 3, w0=90, ww=90, delay_init=True).cuda()
        with torch.no_grad():
            for (a, b) in zip(fine.parameters(), fine_baked.parameters()):
                a.copy_(b)
            for seq in fine.net[:-1]:
                seq[0].weight.div_(90.0)
                seq[0].bias.div_(90.0)
        training.train_step(fine, x, loss_fn, frozen=LevelStack([coarse, med], frozen_prec), precision="fp16")
        errs, num, den = [], 0.0, 0.0
        for k, p in enumerate(fine.parameters()):
            r = ref[k] * oms[k // 2]
            e = (p.grad.double() - r).norm()
            errs.append(float(e / r.norm()))
            num += float(e) ** 2
            den += float(r.norm()) ** 2
        tot = (num / den) ** 0.5
        print(f"fp16 over {frozen_prec} levels: all {tot:.2e}, per tensor " + " ".join(f"{e:.1e}" for e in errs))
        if frozen_prec == "bf16x3":
            assert max(errs) <= 1e-3 and tot <= 3e-4, (errs, tot)


def test_config5_uneven_batch_against_the_oracle():
    """Twelve models of the uneven workload (every width, 1-3 levels, their own
    targets) extracted in one batch at ragged reduced grids (FP32) and traced
    in one batch at 96x64 (FP32 and BF16): each model against the oracle."""
    from mplicits.batch import BatchedExtraction
    from mplicits.tracer import BatchedTrace
    work = uneven_workload(12, device="cuda")
    assert {m["width"] for m in work} >= {64, 128, 256, 512} and {len(m["levels"]) for m in work} == {1, 2, 3}
    grids = [24 + 4 * (j % 5) for j in range(len(work))]  # 24..40
    chunk = 4096
    stacks = [LevelStack(m["levels"], "fp32") for m in work]
    ext = BatchedExtraction(stacks, grids, [m["deltas"] for m in work], chunk=chunk)
    ext.run(graph=True)
    torch.cuda.synchronize()
    for j, m in enumerate(work):
        N = grids[j]
        check_volume_against_oracle(ext.volume(j), m["levels"], N, m["deltas"], chunk,
                                    range((N ** 3 + chunk - 1) // chunk), "fp32", SDF_TOL["fp32"],
                                    max_flip_frac=5e-3)
    ext.close()
    W, H = 96, 64
    for prec in ("fp32", "bf16"):
        err = 0.0
        pst = [LevelStack(m["levels"], prec) for m in work]
        tr = BatchedTrace([(st, m["camera"], W, H, m["deltas"]) for st, m in zip(pst, work)])
        tr.run(graph=True)
        torch.cuda.synchronize()
        shared = 0
        g = torch.Generator().manual_seed(77)
        xs = torch.rand(8192, 3, generator=g) * 2.4 - 1.2
        for j, m in enumerate(work):
            out = {"depth": tr.depth[j], "steps": tr.steps[j]}
            if prec != "fp32":
                # this model's evaluation error in this precision (its fitted
                # output layer sets the scale), measured, with a margin of 2
                ref = sum(osiren.forward(osiren.layers_from_model(lv), xs.double().numpy()) for lv in m["levels"])
                got = pst[j].eval(xs.cuda()).double().cpu().numpy()
                err = 2.0 * float(np.abs(got - ref).max())
            shared += check_trace_against_oracle(out, m["levels"], m["deltas"], [m["camera"]], W, H, 1200, j, err,
                                                 min_hits=0)
        assert shared > 500, shared
        tr.close()
