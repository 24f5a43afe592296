"""Band-culled mesh extraction: drop-in for the reference's ``i3d.meshing``.

Same functions, arguments, prints and return values as i3d/meshing.py; the
sampling loop, the coordinate grid and marching cubes run on the HIP library
(mp_extract_volume, mp_grid_coords, mp_mc_count / mp_mc_emit).  Differences
from the reference, by design:
  * the N^3 x 4 coordinate tensor is never built (coordinates are generated in
    the kernels from the linear index, reproducing the reference's float32
    arithmetic; ``grid="lattice"`` selects the regular lattice instead);
  * the culling runs as whole-grid passes with the per-chunk semantics of the
    reference loop (``max_batch`` is the chunk size), with two host
    synchronisations in total instead of two per chunk;
  * marching cubes runs on the GPU; a vertex inside a cube (tiling index 12)
    is created once per cube and shared by that cube's triangles;
  * ``precision`` ("fp32", "bf16", "fp16", "bf16x3", "bf16x3f") selects the hidden-GEMM
    arithmetic.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from .ply import save_ply  # noqa: F401  (re-exported, i3d/meshing.py:316-384)
from .stack import LevelStack, grid_coords

__all__ = ["gen_mc_coordinate_grid", "create_mesh", "create_mesh_multistage",
           "convert_sdf_samples_to_ply", "save_ply", "marching_cubes_gpu"]


def gen_mc_coordinate_grid(N, voxel_size, t=None, device="cpu", voxel_origin=[-1, -1, -1]):
    """[N^3, 4] float32 samples, columns 0-2 the coordinates, column 3 zero
    (i3d/meshing.py:11-69).  Computed on the GPU with the reference's float32
    arithmetic, then moved to ``device``."""
    if t is not None:
        raise NotImplementedError("space-time (4D) grids are outside this package's scope")
    if float(voxel_size) != 2.0 / (N - 1) or list(voxel_origin) != [-1, -1, -1]:
        raise NotImplementedError("the GPU grid uses voxel_size = 2/(N-1) and origin (-1,-1,-1), "
                                  "as every reference caller does")
    xyz = grid_coords(N, "reference", device="cuda")
    samples = torch.zeros(N ** 3, 4, device=xyz.device)
    samples[:, :3] = xyz
    return samples.to(device)


def marching_cubes_gpu(volume, level=0.0, with_normals=True):
    """Lewiner marching cubes of a 3D float32 GPU tensor.

    Returns (verts [V, 3] float32 in voxel units, axis0..2 order, faces [F, 3]
    int32, normals [V, 3] float32 or None, values [V] float32) as GPU tensors."""
    lib = _lib.load()
    vol = volume.detach()
    if not vol.is_cuda:
        raise RuntimeError("marching_cubes_gpu needs a GPU tensor (no CPU path)")
    if vol.dim() != 3:
        raise ValueError(f"volume must be 3D, got shape {tuple(vol.shape)}")
    vol = vol.float().contiguous()
    nz, ny, nx = vol.shape
    ws = torch.empty(int(lib.mp_mc_workspace_size(nz, ny, nx)), dtype=torch.uint8, device=vol.device)
    nv = ctypes.c_int64()
    nf = ctypes.c_int64()
    with torch.cuda.device(vol.device):
        stream = _lib.stream_handle()
        _lib.check(lib.mp_mc_count(_lib.ptr(vol), nz, ny, nx, float(level), ctypes.byref(nv),
                                   ctypes.byref(nf), _lib.ptr(ws), stream))
        verts = torch.empty(nv.value, 3, dtype=torch.float32, device=vol.device)
        faces = torch.empty(nf.value, 3, dtype=torch.int32, device=vol.device)
        normals = torch.empty(nv.value, 3, dtype=torch.float32, device=vol.device) if with_normals else None
        values = torch.empty(nv.value, dtype=torch.float32, device=vol.device)
        if nv.value == 0 and nf.value == 0:
            return verts, faces, normals, values
        _lib.check(lib.mp_mc_emit(_lib.ptr(vol), nz, ny, nx, float(level), _lib.ptr(verts),
                                  _lib.ptr(faces), _lib.ptr(normals), _lib.ptr(values), _lib.ptr(ws),
                                  stream))
    return verts, faces, normals, values


def convert_sdf_samples_to_ply(pytorch_3d_sdf_tensor, voxel_grid_origin, voxel_size, offset=None,
                               scale=None):
    """Marching cubes at level 0 plus the reference's coordinate transform
    (i3d/meshing.py:267-313).  Returns numpy (mesh_points float64, faces,
    normals, values); prints and returns empty arrays when 0 is outside the
    value range."""
    vol = pytorch_3d_sdf_tensor
    if not isinstance(vol, torch.Tensor):
        vol = torch.from_numpy(np.asarray(vol, dtype=np.float32))
    if not vol.is_cuda:
        vol = vol.cuda()
    verts, faces, normals, values = (np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    level = 0.0
    vmin, vmax = (float(x) for x in torch.aminmax(vol.float()))
    if level < vmin or level > vmax:
        print("Surface level must be within volume data range.")
    else:
        v, f, n, val = marching_cubes_gpu(vol, level)
        verts = v.cpu().numpy()
        faces = f.cpu().numpy()
        normals = n.cpu().numpy()
        values = val.cpu().numpy()
        # skimage: vertices * spacing (float64) when spacing != (1, 1, 1)
        spacing = [voxel_size] * 3
        if not np.array_equal(spacing, (1, 1, 1)):
            verts = verts * np.r_[spacing]
    mesh_points = np.zeros_like(verts)
    mesh_points[:, 0] = voxel_grid_origin[0] + verts[:, 0]
    mesh_points[:, 1] = voxel_grid_origin[1] + verts[:, 1]
    mesh_points[:, 2] = voxel_grid_origin[2] + verts[:, 2]
    if scale is not None:
        mesh_points = mesh_points / scale
    if offset is not None:
        mesh_points = mesh_points - offset
    return mesh_points, faces, normals, values


def create_mesh_multistage(decoders, deltas=[], filename="", t=None, N=256, max_batch=64 ** 3,
                           offset=None, scale=None, device="cpu", silent=False, precision=None,
                           grid="reference", return_volume=False):
    """Band-culled multi-level extraction (i3d/meshing.py:135-264).

    ``decoders`` (coarse first) must live on the GPU.  Returns (verts, faces,
    normals, values, sampling_time); with ``return_volume`` also the GPU
    volume and the per-level sample counts.
    """
    if t is not None and t != -1:
        raise NotImplementedError("space-time (4D) models are outside this package's scope")
    decoders = [d.eval() for d in decoders]
    if deltas and len(deltas) != len(decoders) - 1:
        raise ValueError(f"deltas must be empty or have {len(decoders) - 1} entries")
    precision = precision or getattr(decoders[0], "precision", "fp32")
    voxel_origin = [-1, -1, -1]
    voxel_size = 2.0 / (N - 1)
    stack = LevelStack(decoders, precision=precision)
    torch.cuda.synchronize()
    start = time.perf_counter()
    vol, counts = stack.extract_volume(N, deltas, chunk=max_batch, grid=grid)
    torch.cuda.synchronize()
    sampling_time = time.perf_counter() - start
    if not silent:
        print(f"Sampling took: {sampling_time} s")
    verts, faces, normals, values = convert_sdf_samples_to_ply(vol, voxel_origin, voxel_size, offset,
                                                              scale)
    if filename:
        if not silent:
            print(f"Saving mesh to {filename}")
        save_ply(verts, faces, filename)
        if not silent:
            print("Done")
    if return_volume:
        return verts, faces, normals, values, sampling_time, vol, counts
    return verts, faces, normals, values, sampling_time


def create_mesh(decoder, filename="", t=None, N=256, max_batch=64 ** 3, offset=None, scale=None,
                device="cpu", silent=False):
    """Single-level extraction (i3d/meshing.py:72-132)."""
    if t is not None and t != -1:
        raise NotImplementedError("space-time (4D) models are outside this package's scope")
    decoder.eval()
    voxel_origin = [-1, -1, -1]
    voxel_size = 2.0 / (N - 1)
    stack = LevelStack([decoder], precision=getattr(decoder, "precision", "fp32"))
    torch.cuda.synchronize()
    start = time.time()
    vol, _ = stack.extract_volume(N, [], chunk=max_batch)
    torch.cuda.synchronize()
    end = time.time()
    if not silent:
        print(f"Sampling took: {end - start} s")
    verts, faces, normals, values = convert_sdf_samples_to_ply(vol, voxel_origin, voxel_size, offset,
                                                              scale)
    if filename:
        if not silent:
            print(f"Saving mesh to {filename}")
        save_ply(verts, faces, filename)
        if not silent:
            print("Done")
    return verts, faces, normals, values
