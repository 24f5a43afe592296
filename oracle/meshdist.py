"""Training-data generation restated in numpy FP64 (oracle; test infrastructure
only).

* ``distance``: unsigned point-to-mesh distance (what open3d's
  RaycastingScene.compute_distance returns to experiment_scripts/utils.py:77-78)
  by brute force over all triangles.  The point-triangle distance is written
  independently of the GPU's Voronoi-region walk: the projection onto the
  triangle's plane when its barycentric coordinates are all non-negative,
  else the nearest of the three edge segments (clamped projection).
* ``calc_deltas``: experiment_scripts/utils.py:46-89 (step max_delta /
  n_iters, eps 2e-4, loop while i < n_iters and max(delta) < max_delta).
* ``neighborhood_samples``: create_training_data_on_neighborhood,
  experiment_scripts/utils.py:92-169, given its random draws (vertex ids and
  uniform numbers), in float32 like the reference.
"""
import numpy as np


def _seg_d2(P, A, B):
    """Squared distance of points P [n, 3] to the segments A-B [m, 3] -> [n, m]."""
    AB = B - A
    L2 = np.maximum((AB * AB).sum(1), 1e-300)
    AP = P[:, None, :] - A[None]
    t = np.clip((AP * AB[None]).sum(2) / L2[None], 0.0, 1.0)
    D = AP - t[..., None] * AB[None]
    return (D * D).sum(2)


def point_triangle_d2(P, A, B, C):
    """Squared distances [n, m] of points P [n, 3] to triangles (A, B, C) [m, 3]."""
    P, A, B, C = (np.asarray(x, np.float64) for x in (P, A, B, C))
    e0, e1 = B - A, C - A
    nrm = np.cross(e0, e1)
    n2 = (nrm * nrm).sum(1)
    AP = P[:, None, :] - A[None]
    # barycentric coordinates of the projection (degenerate triangles: edges only)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = (np.cross(AP, e1[None]) * nrm[None]).sum(2) / n2[None]
        v = (np.cross(e0[None], AP) * nrm[None]).sum(2) / n2[None]
        h = (AP * nrm[None]).sum(2) ** 2 / n2[None]
    inside = (n2[None] > 0) & (u >= 0) & (v >= 0) & (u + v <= 1)
    edges = np.minimum(np.minimum(_seg_d2(P, A, B), _seg_d2(P, B, C)), _seg_d2(P, C, A))
    return np.where(inside, h, edges)


def distance(verts, faces, pts, chunk=256):
    """Unsigned distance [n] of pts [n, 3] to the mesh (verts [V, 3], faces [F, 3])."""
    V = np.asarray(verts, np.float64)
    F = np.asarray(faces, np.int64)
    A, B, C = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    pts = np.asarray(pts, np.float64)
    out = np.empty(len(pts))
    for s in range(0, len(pts), chunk):
        out[s:s + chunk] = np.sqrt(point_triangle_d2(pts[s:s + chunk], A, B, C).min(1))
    return out


def calc_deltas(verts, faces, vertices, max_delta=0.05, n_iters=100, eps=2e-4):
    """[V, 7] = (x, y, z, unit normal, delta) and the iterations run."""
    vertices = np.asarray(vertices, np.float64)
    out = np.empty((len(vertices), 7))
    out[:, :6] = vertices[:, :6]
    out[:, 3:6] /= np.sqrt((out[:, 3:6] ** 2).sum(1))[:, None]
    step = np.full(len(vertices), max_delta / n_iters)
    out[:, 6] = step
    i = 0
    while i < n_iters and out[:, 6].max() < max_delta:
        d = out[:, 6]
        sdf = distance(verts, faces, vertices[:, :3] + d[:, None] * out[:, 3:6])
        far = np.abs(d - sdf) > eps
        step = np.where(far, step * 0.5, step)
        out[:, 6] = np.where(far, d - step, d + step)
        i += 1
    return out, i


def neighborhood_samples(vertices, idx, rnd, delta, delta_eps=0.0005, use_neigh_delta=False, dithering=True):
    """(pts, normals, sdfs) in float32 from the sampled rows idx (None: all
    rows) and the uniform numbers rnd [n, 3]."""
    f = np.float32
    vertices = np.asarray(vertices, f)
    surf = vertices if idx is None else vertices[np.asarray(idx)]
    rnd = np.asarray(rnd, f)
    p, q = surf[:, :3], surf[:, 3:6]
    dn = surf[:, -1:] if use_neigh_delta else np.full((len(surf), 1), f(delta_eps), f)
    r = rnd[:, -1:]
    off = p + (dn * r) * q
    pts = [p, off]
    nrm = [q, q]
    sdfs = [np.zeros(len(surf), f), (dn * r)[:, 0]]
    if dithering:
        pts.append(rnd * f(delta) + (p + (f(0.5) * f(delta)) * q))
        nrm.append(q)
        sdfs.append(np.full(len(surf), -1.0, f))
    return np.concatenate(pts), np.concatenate(nrm), np.concatenate(sdfs)
