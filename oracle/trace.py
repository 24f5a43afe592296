"""Multiscale sphere tracing restated in numpy FP64 (oracle; test infrastructure only).

restates the ray generation of InitializeRays_kernel
(renderer/cuda/kernels.cu:400-487: NDC -> inverse projection -> inverse view,
unit direction, start where the ray enters the sphere of radius R) and the
step rule of IterateSphereTracing_kernel (kernels.cu:315-398): outside the
coarse band (F_0 >= 1.4 delta_0) step by F_0 - delta_0, inside step by the
partial sum of the levels whose bands the ray is in; plus the hit test
|F| < eps and the exit test of the persistent tracer.
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
    for k, c in enumerate(cam):
        iv, ip = (np.asarray(m, np.float64) for m in cameras[c])
        eye = ip @ np.array([nx[k], ny[k], -1.0, 1.0])
        world = iv[:3, :3] @ eye[:3] + iv[:3, 3]
        o = iv[:3, 3]
        d = world - o
        O[k], D[k] = o, d / np.linalg.norm(d)
    b = (O * D).sum(1)
    c = (O * O).sum(1) - bound_radius ** 2
    disc = b * b - c
    inside = (c <= 0) | ((b < 0) & (disc > 0))
    t0 = np.where((c > 0) & inside, -b - np.sqrt(np.maximum(disc, 0)), 0.0)
    return O + t0[:, None] * D, D, t0, inside


def trace(level_layers, deltas, P, D, max_steps=128, eps=1e-4, bound_radius=1.2, band_scale=1.4):
    """Trace rays (P, D) through the FP64 level stack.  Returns (hit, t, position)."""
    n = len(P)
    P = P.copy()
    t = np.zeros(n)
    hit = np.zeros(n, bool)
    live = np.ones(n, bool)
    L = len(level_layers)
    for _ in range(max_steps + 1):
        idx = np.nonzero(live)[0]
        if len(idx) == 0:
            break
        x = P[idx]
        F0 = osiren.forward(level_layers[0], x)
        if L == 1:
            s = F0
            h = np.abs(F0) < eps
        else:
            band = F0 < band_scale * deltas[0]
            F = F0.copy()
            need = band.copy()
            for l in range(1, L):
                if need.any():
                    F[need] += osiren.forward(level_layers[l], x[need])
                need = need & (l + 1 < L) & (np.abs(F) < band_scale * deltas[min(l, L - 2)])
            s = np.where(band, F, F0 - deltas[0])
            h = band & (np.abs(F) < eps)
        hit[idx[h]] = True
        live[idx[h]] = False
        mv = idx[~h]
        P[mv] += s[~h, None] * D[mv]
        t[mv] += s[~h]
        r2 = (P[mv] ** 2).sum(1)
        outward = (P[mv] * D[mv]).sum(1) * s[~h]  # signed steps: moving away from the centre
        out = (r2 > bound_radius ** 2 * 1.0001) & (outward > 0)
        live[mv[out]] = False
    return hit, t, P
