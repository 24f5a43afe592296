"""Band-culled volume evaluation (mp_extract_volume) against the restated
sampling loop of create_mesh_multistage (oracle.extract)."""
import numpy as np
import pytest
import torch

from mplicits.stack import LevelStack
from mplicits.synthetic import seeded_siren
from oracle import extract as oextract
from oracle import siren as osiren

pytestmark = pytest.mark.gpu


def _levels(widths=(128, 128, 256), seeds=(21, 22, 23)):
    coarse = seeded_siren([widths[0]] * 3, 30.0, seed=seeds[0], last_scale=20.0).cuda()
    med = seeded_siren([widths[1]] * 3, 60.0, seed=seeds[1], last_scale=0.1).cuda()
    fine = seeded_siren([widths[2]] * 3, 90.0, seed=seeds[2], last_scale=0.1).cuda()
    return [coarse, med, fine]


def _oracle_volume(levels, N, deltas, chunk):
    layers = [osiren.layers_from_model(m) for m in levels]
    n_chunks = (N ** 3 + chunk - 1) // chunk
    out = oextract.sample_chunks(layers, N, deltas, chunk, range(n_chunks))
    vol = np.empty(N ** 3)
    for idx, v in out.values():
        vol[idx] = v
    return vol


@pytest.mark.parametrize("deltas", [[], [0.05, 0.02], [0.02, 0.08]], ids=["unculled", "nested", "wider-fine"])
def test_volume_matches_oracle(deltas):
    levels = _levels()
    N, chunk = 40, 4096
    ref = _oracle_volume(levels, N, deltas, chunk)
    st = LevelStack(levels, "fp32")
    vol, counts = st.extract_volume(N, deltas, chunk=chunk)
    v = vol.double().cpu().numpy().reshape(-1)
    assert np.abs(v - ref).max() <= 1e-5
    assert counts[0] == N ** 3
    if deltas:
        assert 0 < counts[1] < N ** 3


def test_chunk_break_semantics():
    # With delta_1 > delta_0 some chunks have level-2 band samples but no level-1
    # sample: the reference breaks out for those chunks and never adds level 2.
    levels = _levels()
    N, chunk = 40, 512
    deltas = [0.004, 0.2]
    layers = [osiren.layers_from_model(m) for m in levels]
    ref = _oracle_volume(levels, N, deltas, chunk)
    st = LevelStack(levels, "fp32")
    vol, counts = st.extract_volume(N, deltas, chunk=chunk)
    v = vol.double().cpu().numpy().reshape(-1)
    assert np.abs(v - ref).max() <= 1e-5
    # the case is exercised: a chunk-blind evaluation would differ
    from oracle import grid as ogrid
    xyz = ogrid.coords(N, np.arange(N ** 3)).astype(np.float64)
    s = osiren.forward(layers[0], xyz)
    m1 = np.abs(s) < np.float32(deltas[0])
    s1 = s.copy()
    s1[m1] += osiren.forward(layers[1], xyz[m1])
    m2 = np.abs(s1) < np.float32(deltas[1])
    blind = s1.copy()
    blind[m2] += osiren.forward(layers[2], xyz[m2])
    assert np.abs(blind - ref).max() > 1e-6


def test_reduced_precision_volume_close():
    # Reduced precision moves values by ~1e-4, so a sample sitting on a band
    # edge may be culled differently; compare the unculled volumes and the band
    # counts.
    levels = _levels()
    N, chunk = 48, 32 ** 3
    ref = _oracle_volume(levels, N, [], chunk)
    _, counts32 = LevelStack(levels, "fp32").extract_volume(N, [0.05, 0.02], chunk=chunk)
    for prec, tol in (("fp16", 2e-3), ("bf16", 1e-2)):
        st = LevelStack(levels, prec)
        vol, _ = st.extract_volume(N, [], chunk=chunk)
        err = np.abs(vol.double().cpu().numpy().reshape(-1) - ref).max()
        assert err <= tol * np.abs(ref).max(), (prec, err)
        _, counts = st.extract_volume(N, [0.05, 0.02], chunk=chunk)
        for a, b in zip(counts, counts32):
            assert abs(a - b) <= 0.01 * b + 8, (prec, counts, counts32)


def test_bf16x3_culled_volume_matches_oracle():
    # Three-term BF16 keeps the culled volume within 1e-4 of the FP64 sampling
    # loop (all three levels 128/256 wide: every level runs the split path)
    levels = _levels(widths=(256, 128, 256))
    N, chunk = 40, 4096
    ref = _oracle_volume(levels, N, [0.05, 0.02], chunk)
    vol, counts = LevelStack(levels, "bf16x3").extract_volume(N, [0.05, 0.02], chunk=chunk)
    _, counts32 = LevelStack(levels, "fp32").extract_volume(N, [0.05, 0.02], chunk=chunk)
    v = vol.double().cpu().numpy().reshape(-1)
    # a sample within rounding of a band edge may be culled differently
    same = np.abs(v - ref) <= 1e-4
    assert same.mean() >= 0.9995, same.mean()
    assert [abs(a - b) <= 2 for a, b in zip(counts, counts32)] == [True] * 3, (counts, counts32)
    assert 0 < counts[2] < counts[1] < N ** 3
