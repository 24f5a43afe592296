"""Multiscale sphere tracing restated in numpy FP64 (oracle; test infrastructure only).

rays() restates the ray generation of InitializeRays_kernel
(renderer/cuda/kernels.cu:400-487: NDC -> inverse projection -> inverse view,
unit direction) and starts each ray where it enters the sphere of radius R.

trace() restates the step rule of IterateSphereTracing_kernel
(kernels.cu:315-398) with its fixed-length phases (MIPplicit::SdfInference,
renderer/cuda/state.cu:848-881) switched per ray:
  F_0 = f_0(p), banded = F_0 < band_scale * delta_0.
  lod-0 phase (kernels.cu:331-336, 369-373) until the ray is first banded:
      p += (F_0 - delta_0) d.
  residual phase (kernels.cu:337-358, 374-394) from then on:
      banded:     p += (F_0 + f_1(p) + ... + f_{L-1}(p)) d   (delta_lod_1 unused)
      not banded: p += F_0 d.
  One level: p += F_0 d.
plus the persistent tracer's hit test (banded and |F| < eps) and exit test
(outside the bounding sphere and moving away, or out of steps).

A level is an oracle layer list (oracle.siren) or a callable f(x [n, 3]) -> [n].
"""
import numpy as np

from . import siren as osiren


def rays(cameras, width, height, bound_radius=1.2, pixel_center=True, flip_y=False, ids=None):
    """(origins at the sphere entry, directions, t_enter, inside) for ray ids."""
    per = width * height
    ids = np.arange(len(cameras) * per) if ids is None else np.asarray(ids)
    cam, pix = ids // per, ids % per
    i, j = pix % width, pix // width
    off = 0.5 if pixel_center else 0.0
    nx = 2.0 * (i + off) / width - 1.0
    ny = 2.0 * (j + off) / height - 1.0
    if flip_y:
        ny = -ny
    O = np.empty((len(ids), 3))
    D = np.empty((len(ids), 3))
    for c in np.unique(cam):
        k = np.nonzero(cam == c)[0]
        iv, ip = (np.asarray(m, np.float64) for m in cameras[c])
        clip = np.stack([nx[k], ny[k], -np.ones(len(k)), np.ones(len(k))], 1)
        eye = clip @ ip.T
        world = eye[:, :3] @ iv[:3, :3].T + iv[:3, 3]
        o = iv[:3, 3]
        d = world - o
        O[k], D[k] = o, d / np.linalg.norm(d, axis=1, keepdims=True)
    b = (O * D).sum(1)
    c = (O * O).sum(1) - bound_radius ** 2
    disc = b * b - c
    inside = (c <= 0) | ((b < 0) & (disc > 0))
    t0 = np.where((c > 0) & inside, -b - np.sqrt(np.maximum(disc, 0)), 0.0)
    return O + t0[:, None] * D, D, t0, inside


def _level_fn(level):
    if callable(level):
        return level
    return lambda x: osiren.forward(level, x)


def trace(levels, deltas, P, D, max_steps=128, eps=1e-4, bound_radius=1.2, band_scale=1.4):
    """Trace rays (P, D) through the level stack.

    Returns a dict: hit [n] bool, t [n] distance travelled from P, p [n, 3] final
    positions, steps [n] steps taken, fmin [n] the smallest |F| the ray's hit
    test saw (inf if it was never banded), edge [n] the smallest
    |F_0 - band_scale delta_0| along the ray (inf for one level: how close a
    band decision came to flipping), crossed [n] whether a step carried the
    ray across the level set and beyond the hit band (the step distance changed
    sign with |F| >= eps: the field grew
    faster than a distance along the ray, and where the ray lands is then
    sensitive to rounding)."""
    fns = [_level_fn(lv) for lv in levels]
    n = len(P)
    P = np.array(P, np.float64, copy=True)
    t = np.zeros(n)
    steps = np.zeros(n, np.int64)
    hit = np.zeros(n, bool)
    live = np.ones(n, bool)
    entered = np.zeros(n, bool)
    fmin = np.full(n, np.inf)
    edge = np.full(n, np.inf)
    crossed = np.zeros(n, bool)
    last = np.zeros(n)
    L = len(fns)
    while True:
        idx = np.nonzero(live)[0]
        if len(idx) == 0:
            break
        x = P[idx]
        F0 = np.asarray(fns[0](x), np.float64)
        if L == 1:
            s = F0
            h = np.abs(F0) < eps
            fmin[idx] = np.minimum(fmin[idx], np.abs(F0))
        else:
            band = F0 < band_scale * deltas[0]
            edge[idx] = np.minimum(edge[idx], np.abs(F0 - band_scale * deltas[0]))
            F = F0.copy()
            if band.any():
                for f in fns[1:]:
                    F[band] += np.asarray(f(x[band]), np.float64)
            entered[idx[band]] = True
            s = np.where(band, F, np.where(entered[idx], F0, F0 - deltas[0]))
            h = band & (np.abs(F) < eps)
            fmin[idx[band]] = np.minimum(fmin[idx[band]], np.abs(F[band]))
        crossed[idx] |= (last[idx] * s < 0) & (np.abs(s) >= eps)
        last[idx] = s
        hit[idx[h]] = True
        live[idx[h]] = False
        mv = idx[~h]
        sm = s[~h]
        P[mv] += sm[:, None] * D[mv]
        t[mv] += sm
        steps[mv] += 1
        r2 = (P[mv] ** 2).sum(1)
        outward = (P[mv] * D[mv]).sum(1) * sm  # signed steps: moving away from the centre
        out = ((r2 > bound_radius ** 2 * 1.0001) & (outward > 0)) | (steps[mv] >= max_steps)
        live[mv[out]] = False
    return {"hit": hit, "t": t, "p": P, "steps": steps, "fmin": fmin, "edge": edge, "crossed": crossed}
