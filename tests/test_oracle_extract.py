"""The band-culled sampling oracle (oracle.extract) against hand-computed cases
and invariants of the reference loop (i3d/meshing.py:212-236); no GPU."""
import numpy as np
import torch

from oracle import extract as oextract
from oracle import siren as osiren


def _table(values):
    """A level whose value at sample k is values[k] (x[:, 0] carries k)."""
    v = np.asarray(values, np.float64)
    return lambda x: v[x[:, 0].astype(np.int64)]


def _pts(n):
    x = np.zeros((n, 3))
    x[:, 0] = np.arange(n)
    return x


def test_band_of_a_level_is_taken_on_the_updated_sum():
    # deltas (0.05, 0.02).  Level 1 is added where |f_0| < 0.05 (samples 0, 1, 3);
    # level 2 where |f_0 + f_1| < 0.02 (samples 0 and 3).  Masking level 2 on
    # |f_0| instead would add it at sample 0 only.
    f0 = _table([0.01, 0.03, 0.2, 0.04])
    f1 = _table([0.005, 0.0, 0.5, -0.03])
    f2 = _table([1.0, 1.0, 1.0, 1.0])
    s = oextract.sample_chunk([f0, f1, f2], _pts(4), [0.05, 0.02])
    np.testing.assert_array_equal(s, [1.015, 0.03, 0.2, 1.01])


def test_empty_band_breaks_out_of_the_finer_levels():
    # no sample of the chunk has |f_0| < delta_0 = 0.05: the loop breaks, and
    # level 2 is not added although every |f_0| < delta_1 = 0.1
    f0 = _table([0.08, 0.09, -0.07])
    f1 = _table([5.0, 5.0, 5.0])
    f2 = _table([7.0, 7.0, 7.0])
    s = oextract.sample_chunk([f0, f1, f2], _pts(3), [0.05, 0.1])
    np.testing.assert_array_equal(s, [0.08, 0.09, -0.07])
    # one sample inside the coarse band re-enables both levels, each on its own
    # band: level 1 at sample 1 only, level 2 where |f_0 + f_1| < 0.1 (samples 0, 2)
    f0 = _table([0.08, 0.04, -0.07])
    s = oextract.sample_chunk([f0, f1, f2], _pts(3), [0.05, 0.1])
    np.testing.assert_array_equal(s, [7.08, 5.04, 6.93])


def test_no_deltas_sums_every_level_everywhere():
    rng = np.random.default_rng(1)
    layers = [[(rng.normal(size=(8, 3)), rng.normal(size=8)), (rng.normal(size=(1, 8)) * 0.1, rng.normal(size=1))]
              for _ in range(3)]
    x = rng.uniform(-1, 1, (50, 3))
    s = oextract.sample_chunk(layers, x, [])
    want = sum(osiren.forward(lay, x) for lay in layers)
    np.testing.assert_allclose(s, want, rtol=0, atol=1e-14)


def test_band_edges_compare_in_float32():
    # torch compares float32 samples with the Python float delta in float32:
    # float32(0.02) = 0.01999999955... < 0.02, and a sample equal to it is
    # outside the band (a float64 comparison would put it inside)
    d32 = float(np.float32(0.02))
    assert not bool(torch.tensor([d32], dtype=torch.float32) < 0.02)  # the torch semantics
    assert d32 < 0.02
    f0 = _table([d32, np.nextafter(d32, 0.0)])
    f1 = _table([1.0, 1.0])
    s = oextract.sample_chunk([f0, f1], _pts(2), [0.02])
    np.testing.assert_array_equal(s, [d32, np.nextafter(d32, 0.0) + 1.0])


def test_chunks_are_consecutive_linear_indices():
    # chunks of max_batch = 4 consecutive linear indices: 0..3 and 4..7; each
    # chunk's band samples receive level 1, and a chunk without any receives
    # nothing
    N = 2  # 8 grid points
    coarse = lambda x: np.where(x[:, 2] > 0.5, 0.01, 0.3)  # z = 1 only at odd indices (reference grid)
    rest = lambda x: np.full(len(x), 2.0)
    out = oextract.sample_chunks([coarse, rest], N, [0.05], 4, range(2))
    (i0, v0), (i1, v1) = out[0], out[1]
    np.testing.assert_array_equal(i0, [0, 1, 2, 3])
    np.testing.assert_array_equal(i1, [4, 5, 6, 7])
    # grid z at index k of the reference grid (N = 2): k % 2
    np.testing.assert_array_equal(v0, [0.3, 2.01, 0.3, 2.01])
    np.testing.assert_array_equal(v1, [0.3, 2.01, 0.3, 2.01])
    # a coarse band only in chunk 1 leaves chunk 0 untouched
    coarse1 = lambda x: np.where((x[:, 2] > 0.5) & (x[:, 0] > 0.5), 0.01, 0.3)
    out = oextract.sample_chunks([coarse1, rest], N, [0.05], 4, range(2))
    np.testing.assert_array_equal(out[0][1], [0.3, 0.3, 0.3, 0.3])
    assert out[1][1].max() > 2.0


def test_masks_name_the_samples_each_level_received():
    f0 = _table([0.01, 0.03, 0.2, 0.04])
    f1 = _table([0.005, 0.0, 0.5, -0.03])
    f2 = _table([1.0, 1.0, 1.0, 1.0])
    s, (m1, m2) = oextract.sample_chunk([f0, f1, f2], _pts(4), [0.05, 0.02], return_masks=True)
    np.testing.assert_array_equal(m1, [True, True, False, True])
    np.testing.assert_array_equal(m2, [True, False, False, True])
    _, (m1, m2) = oextract.sample_chunk([f0, f1, f2], _pts(4), [0.001, 0.5], return_masks=True)
    assert not m1.any() and not m2.any()  # the loop broke at level 1


def test_whole_grid_loop_equals_the_chunk_loop():
    # the torch whole-grid restatement (the benchmark's FP64 checker) against
    # the per-chunk restatement, including the break and the updated-sum masks
    rng = np.random.default_rng(4)
    mk = lambda w, s: [(rng.normal(size=(w, 3)) * s, rng.normal(size=w)), (rng.normal(size=(1, w)) * 0.05,  # noqa
                                                                             rng.normal(size=1) * 0.02)]
    layers = [mk(16, 3.0), mk(8, 5.0), mk(8, 7.0)]
    layers[0][1] = (layers[0][1][0] * 4, np.array([0.05]))
    N, chunk = 12, 100
    for deltas in ([], [0.05, 0.02], [0.004, 0.2]):
        ref = oextract.sample_chunks(layers, N, deltas, chunk, range((N ** 3 + chunk - 1) // chunk),
                                     return_masks=True)
        tl = [[(torch.from_numpy(w), torch.from_numpy(b)) for w, b in lay] for lay in layers]
        s, masks = oextract.sampling_loop_full(tl, N, deltas, chunk, "cpu", batch=333)
        for c, (idx, v, m) in ref.items():
            np.testing.assert_allclose(s.numpy()[idx], v, rtol=0, atol=1e-13)
            for a, b in zip(masks, m):
                np.testing.assert_array_equal(a.numpy()[idx], b)
        if deltas:
            assert 0 < masks[0].sum() < N ** 3
