"""Pins the oracle's restatement of gen_mc_coordinate_grid (i3d/meshing.py:42-63)."""
import numpy as np
import torch

from oracle import grid as ogrid


def _torch_reference_expression(N, idx):
    # The reference's own expression (i3d/meshing.py:55-63), evaluated by torch
    # on the CPU: the semantics the restatement must reproduce.
    overall_index = torch.as_tensor(idx, dtype=torch.long)
    samples = torch.zeros(len(idx), 4)
    samples[:, 2] = overall_index % N
    samples[:, 1] = (overall_index.long() / N) % N
    samples[:, 0] = ((overall_index.long() / N) / N) % N
    voxel_size = 2.0 / (N - 1)
    vox = samples[:, :3].clone()
    samples[:, 0] = (samples[:, 0] * voxel_size) + -1
    samples[:, 1] = (samples[:, 1] * voxel_size) + -1
    samples[:, 2] = (samples[:, 2] * voxel_size) + -1
    return vox.numpy(), samples[:, :3].numpy()


def test_known_answers():
    # N=4, index 1 sits at voxel (1/16, 1/4, 1), not (0, 0, 1)
    vx, vy, vz = ogrid.voxel(4, [1])
    assert (vx[0], vy[0], vz[0]) == (np.float32(0.0625), np.float32(0.25), np.float32(1.0))
    # N=512: the last index rounds to 2^27 in float32 and wraps to voxel (0, 0, 511)
    vx, vy, vz = ogrid.voxel(512, [512 ** 3 - 1])
    assert (vx[0], vy[0], vz[0]) == (0.0, 0.0, 511.0)
    # the lattice mode is the regular grid
    vx, vy, vz = ogrid.voxel(4, [1, 4, 16, 63], "lattice")
    assert vx.tolist() == [0, 0, 1, 3] and vy.tolist() == [0, 1, 0, 3] and vz.tolist() == [1, 0, 0, 3]


def test_bit_exact_against_torch_semantics_small():
    for N in (2, 3, 4, 7, 8, 16, 33, 64, 128):
        idx = np.arange(N ** 3)
        vox, xyz = _torch_reference_expression(N, idx)
        vx, vy, vz = ogrid.voxel(N, idx)
        np.testing.assert_array_equal(np.stack([vx, vy, vz], 1), vox)
        np.testing.assert_array_equal(ogrid.coords(N, idx), xyz)


def test_bit_exact_against_torch_semantics_512_regions():
    N = 512
    rng = np.random.default_rng(0)
    idx = np.concatenate([
        np.arange(0, 4096), np.arange(2 ** 24 - 2048, 2 ** 24 + 2048),
        np.arange(N ** 3 - 4096, N ** 3), rng.integers(0, N ** 3, 20000)])
    vox, xyz = _torch_reference_expression(N, idx)
    np.testing.assert_array_equal(np.stack(ogrid.voxel(N, idx), 1), vox)
    np.testing.assert_array_equal(ogrid.coords(N, idx), xyz)


def test_reference_grid_is_not_the_lattice():
    idx = np.arange(1, 64)
    assert not np.array_equal(ogrid.coords(4, idx), ogrid.coords(4, idx, "lattice"))


def test_torch_restatement_is_bit_identical():
    import torch
    for N in (7, 64, 512):
        idx = np.unique(np.concatenate([np.arange(min(N ** 3, 5000)), np.arange(N ** 3 - 100, N ** 3),
                                        np.random.default_rng(N).integers(0, N ** 3, 5000)]))
        for mode in ("reference", "lattice"):
            a = ogrid.coords(N, idx, mode)
            b = ogrid.coords_torch(N, torch.from_numpy(idx), mode).numpy()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (N, mode)
