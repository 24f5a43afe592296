"""Extraction grid coordinates (oracle; test infrastructure only).

restates gen_mc_coordinate_grid (i3d/meshing.py:42-63) as torch evaluates it:
`overall_index.long() / N` is float32 true division of the int64 index (cast to
float32 first), `% N` on a float tensor is the fmod-based remainder, and the
scale and shift are two separately rounded float32 operations.  Column 0 takes
the slowest-varying index component.
"""
import numpy as np


def voxel_size32(N):
    return np.float32(2.0 / (N - 1))


def voxel(N, idx, mode="reference"):
    """float32 voxel-unit components (vx, vy, vz) of linear grid indices
    (i3d/meshing.py:55-57)."""
    idx = np.asarray(idx, dtype=np.int64)
    fN = np.float32(N)
    if mode == "reference":
        vz = (idx % N).astype(np.float32)
        t = idx.astype(np.float32) / fN
        vy = np.fmod(t, fN)
        vx = np.fmod(t / fN, fN)
    elif mode == "lattice":
        vx = (idx // (N * N)).astype(np.float32)
        vy = ((idx // N) % N).astype(np.float32)
        vz = (idx % N).astype(np.float32)
    else:
        raise ValueError(mode)
    return vx, vy, vz


def coords(N, idx, mode="reference"):
    """float32 [len(idx), 3] sample coordinates (i3d/meshing.py:61-63)."""
    idx = np.asarray(idx, dtype=np.int64)
    vx, vy, vz = voxel(N, idx, mode)
    h = voxel_size32(N)
    one = np.float32(-1.0)
    out = np.empty((idx.size, 3), dtype=np.float32)
    out[:, 0] = (vx * h) + one
    out[:, 1] = (vy * h) + one
    out[:, 2] = (vz * h) + one
    return out


def coords_torch(N, idx, mode="reference"):
    """coords() in torch on idx's device (int64 tensor of linear indices): the
    same float32 operations in the same order."""
    import torch
    idx = idx.to(torch.int64)
    fN = torch.tensor(float(N), dtype=torch.float32, device=idx.device)
    if mode == "reference":
        vz = (idx % N).to(torch.float32)
        t = idx.to(torch.float32) / fN
        vy = torch.fmod(t, fN)
        vx = torch.fmod(t / fN, fN)
    elif mode == "lattice":
        vx = (idx // (N * N)).to(torch.float32)
        vy = ((idx // N) % N).to(torch.float32)
        vz = (idx % N).to(torch.float32)
    else:
        raise ValueError(mode)
    h = torch.tensor(float(voxel_size32(N)), dtype=torch.float32, device=idx.device)
    one = torch.tensor(-1.0, dtype=torch.float32, device=idx.device)
    return torch.stack([(vx * h) + one, (vy * h) + one, (vz * h) + one], 1)
