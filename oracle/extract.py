"""Band-culled sampling loop of create_mesh_multistage (oracle; test infrastructure only).

restates i3d/meshing.py:203-238 per chunk of `max_batch` consecutive linear
indices (reconstruct.py:128 passes 32**3):
    s = f_0(x)
    for i in 1..L-1:
        if deltas:
            m = |s| < delta_{i-1};  if not m.any(): break;  s[m] += f_i(x[m])
        else:
            s += f_i(x)
The comparisons use float32 deltas, as torch compares a float32 tensor with a
Python float.
"""
import numpy as np
import torch

from . import grid as ogrid
from . import siren as osiren


def _level_fn(level):
    if callable(level):
        return level
    return lambda x: osiren.forward(level, x)


def sample_chunk(level_layers, xyz, deltas, dtype=np.float64, return_masks=False):
    """Restated loop body for one chunk (i3d/meshing.py:212-236).

    level_layers: per level an oracle layer list or a callable f(x [n, 3]) -> [n].
    The band of level i is taken on the sum so far (meshing.py:225 masks the
    updated samples tensor).  With return_masks, also the samples each residual
    level was added to ([n] bool per level 1..L-1)."""
    fns = [_level_fn(lv) for lv in level_layers]
    s = np.asarray(fns[0](xyz)).astype(dtype)
    masks = [np.zeros(len(s), bool) for _ in fns[1:]]
    for i in range(1, len(fns)):
        if deltas:
            m = np.abs(s) < np.float32(deltas[i - 1])
            if not m.any():
                break
            s[m] = s[m] + np.asarray(fns[i](xyz[m])).astype(dtype)
        else:
            m = np.ones(len(s), bool)
            s = s + np.asarray(fns[i](xyz)).astype(dtype)
        masks[i - 1] = m
    return (s, masks) if return_masks else s


def sample_chunks(level_layers, N, deltas, max_batch, chunk_ids, mode="reference", return_masks=False):
    """{chunk id: (indices, values[, masks])} for the requested chunks (FP64 evaluation)."""
    out = {}
    total = N ** 3
    for c in chunk_ids:
        idx = np.arange(c * max_batch, min((c + 1) * max_batch, total), dtype=np.int64)
        xyz = ogrid.coords(N, idx, mode).astype(np.float64)
        r = sample_chunk(level_layers, xyz, deltas, return_masks=return_masks)
        out[c] = (idx,) + (r if return_masks else (r,))
    return out


def torch_sampling_loop(level_layers, N, deltas, max_batch, chunk_ids=None, mode="reference",
                        threads=None):
    """The reference loop in torch float32 on the CPU (the CPU baseline).

    level_layers: per level [(W, b)] float32 torch tensors (baked).  Evaluates
    the listed chunks (all when None) and returns (values by chunk, per-level
    evaluation counts)."""
    if threads:
        torch.set_num_threads(threads)
    total = N ** 3
    n_chunks = (total + max_batch - 1) // max_batch
    chunk_ids = range(n_chunks) if chunk_ids is None else chunk_ids
    counts = [0] * len(level_layers)
    out = {}
    with torch.no_grad():
        for c in chunk_ids:
            idx = np.arange(c * max_batch, min((c + 1) * max_batch, total), dtype=np.int64)
            x = torch.from_numpy(ogrid.coords(N, idx, mode))
            s = osiren.torch_forward(level_layers[0], x).squeeze(-1)
            counts[0] += x.shape[0]
            for i in range(1, len(level_layers)):
                if deltas:
                    m = torch.abs(s) < deltas[i - 1]
                    if not m.any():
                        break
                    xm = x[m]
                    counts[i] += xm.shape[0]
                    s[m] += osiren.torch_forward(level_layers[i], xm).squeeze(-1)
                else:
                    counts[i] += x.shape[0]
                    s += osiren.torch_forward(level_layers[i], x).squeeze(-1)
            out[c] = (idx, s.numpy())
    return out, counts


def sampling_loop_full(level_layers, N, deltas, max_batch, device, dtype=torch.float64, mode="reference",
                       batch=1 << 20):
    """The whole-grid result of the chunked loop (i3d/meshing.py:212-236) in
    torch (FP64 by default) on `device`: the same per-chunk semantics as
    sample_chunk, evaluated level by level over the grid.

    level_layers: per level [(W, b)] torch tensors of `dtype` on `device`
    (baked).  Returns (values [N^3], masks: per residual level the [N^3] bool
    samples it was added to)."""
    total = N ** 3
    n_chunks = (total + max_batch - 1) // max_batch

    def level_at(layers, idx):
        # coordinates in float32 on the CPU (bit-identical to ogrid.coords),
        # the level on the device
        out = torch.empty(len(idx), dtype=dtype, device=device)
        for i in range(0, len(idx), batch):
            x = ogrid.coords_torch(N, idx[i:i + batch].cpu(), mode).to(device=device, dtype=dtype)
            out[i:i + batch] = osiren.torch_forward(layers, x)[:, 0]
        return out

    with torch.no_grad():
        s = level_at(level_layers[0], torch.arange(total, device=device))
        masks = []
        alive = torch.ones(n_chunks, dtype=torch.bool, device=device)
        chunk_of = lambda i: torch.div(i, max_batch, rounding_mode="floor")  # noqa: E731
        for li in range(1, len(level_layers)):
            if deltas:
                raw = s.abs() < float(np.float32(deltas[li - 1]))
                hit = torch.zeros(n_chunks, dtype=torch.bool, device=device)
                hit[chunk_of(torch.nonzero(raw).flatten())] = True
                alive = alive & hit
                m = raw & alive.repeat_interleave(max_batch)[:total]
            else:
                m = torch.ones(total, dtype=torch.bool, device=device)
            idx = torch.nonzero(m).flatten()
            if len(idx):
                s[idx] += level_at(level_layers[li], idx)
            masks.append(m)
    return s, masks
