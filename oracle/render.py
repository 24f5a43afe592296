"""The renderer's fixed-iteration frame restated in numpy FP64 (oracle; test
infrastructure only).

Restates MIPplicit::SdfInference (renderer/cuda/state.cu:816-1055):
rays of InitializeRays_kernel (renderer/cuda/kernels.cu:400-487; NDC 2i/W - 1
without a pixel centre, start where the ray enters the bounding sphere, or at
the eye when it never does); a fixed number of IterateSphereTracing_kernel
steps per phase (kernels.cu:315-398, launched by state.cu:848-881): phase 0
evaluates level 0 and steps by f0 - delta, phase 1 levels 0 and 1, phase 2
levels 0, 1 (residual models only) and 2, phases 1 and 2 stepping by the
residual sum where f0 < 1.4 delta and by f0 elsewhere; the shading distance,
test and normals of state.cu:936-1054 (sum_residual_normals,
kernels.cu:65-120); and Shade_kernel's Phong / normal colours with
RgbaFloatToInt's truncation to 8 bits (kernels.cu:28-39,122-275).

state.cu:834 binds the distances of level 2 to level 0's output buffer;
``lod2_reads_lod0=True`` reproduces that binding (level 2 then never enters a
step or the shading distance), ``False`` reads level 2's own output.

Levels are (f, grad) callables over positions [n, 3] so that the same code
runs on analytic SDFs (closed-form checks) and on FP64 SIRENs.
"""
import numpy as np

from . import siren as osiren
from .trace import rays as _rays

WHITE = 0xFFFFFFFF


def siren_level(layers):
    """(f, grad) callables of one FP64 SIREN level (frequency baked)."""
    return (lambda P: osiren.forward(layers, P), lambda P: osiren.forward_grad(layers, P)[1])


def rgba8(r, g, b, a=1.0):
    """RgbaFloatToInt (kernels.cu:28-39): clamp to [0, 1], scale by 255,
    truncate; bytes R, G, B, A from the low end."""
    def q(v):
        return np.floor(np.clip(np.asarray(v, np.float64), 0.0, 1.0) * 255.0).astype(np.uint32)
    return q(r) | (q(g) << 8) | (q(b) << 16) | (q(a) << 24)


def phong(n, d):
    """Shade_kernel's PHONG colour for unit normals n and ray directions d
    (light and viewer at the eye: L = V = -d)."""
    L = -d
    dot_d = (L * n).sum(1)
    R = 2.0 * dot_d[:, None] * n - L
    ls = (L * R).sum(1)
    dot_s = np.where(ls > 0, np.abs(ls) ** 35.0, 0.0)
    c = 0.2 + np.maximum(dot_d, 0.0) * 0.54 + 0.5 * dot_s
    return rgba8(c, c, c)


def render_fixed(levels, cameras, width, height, iters=(20, 0, 0), delta=0.0, threshold=0.05, normal_lod=0,
                 shading="phong", residual=True, skip_lod0=False, flip_y=False, bound_radius=1.2,
                 band_scale=1.4, lod2_reads_lod0=False):
    """One frame.  Returns a dict: rgba [n] uint32, pos [n, 3] final
    positions, dist [n] shading distance, shaded [n] bool, normals [n, 3]
    (unit where shaded, else 0)."""
    P, D, _, _ = _rays(cameras, width, height, bound_radius, pixel_center=False, flip_y=flip_y)
    n = len(P)
    out = [np.zeros(n), np.zeros(n), np.zeros(n)]
    lim = delta * band_scale
    if not skip_lod0:
        for _ in range(iters[0]):
            f0 = levels[0][0](P)
            out[0] = f0
            P = P + (f0 - delta)[:, None] * D
    for _ in range(iters[1]):
        f0, f1 = levels[0][0](P), levels[1][0](P)
        out[0], out[1] = f0, f1
        s = np.where(f0 < lim, f0 + f1 if residual else f1, f0)
        P = P + s[:, None] * D
    for _ in range(iters[2]):
        f0 = levels[0][0](P)
        out[0] = f0
        if residual:
            out[1] = levels[1][0](P)
        out[2] = levels[2][0](P)
        r2 = out[0] if lod2_reads_lod0 else out[2]
        s = np.where(f0 < lim, f0 + out[1] + r2 if residual else r2, f0)
        P = P + s[:, None] * D
    o2 = out[0] if lod2_reads_lod0 else out[2]
    if residual:
        dist = out[0].copy()
        if iters[1] > 0 or iters[2] > 0:
            dist = dist + out[1]
        if iters[2] > 0:
            dist = dist + o2
    else:
        dist = o2 if iters[2] > 0 else (out[1] if iters[1] > 0 else out[0])
    shaded = (np.abs(dist) < threshold) & (np.abs(P) < 1.0).all(1)
    rgba = np.full(n, WHITE, np.uint32)
    nrm = np.zeros((n, 3))
    idx = np.nonzero(shaded)[0]
    if len(idx):
        X = P[idx]
        lods = range(normal_lod + 1) if residual else [normal_lod]
        g = sum(levels[l][1](X) for l in lods)
        u = g / np.linalg.norm(g, axis=1, keepdims=True)
        nrm[idx] = u
        if shading == "normals":
            rgba[idx] = rgba8(np.abs(u[:, 0]), np.abs(u[:, 2]), np.abs(u[:, 1]))
        else:
            rgba[idx] = phong(u, D[idx])
    return {"rgba": rgba, "pos": P, "dist": dist, "shaded": shaded, "normals": nrm}
