"""Seeded synthetic workload: the bumpy sphere, seeded SIREN levels, sample
points and orbit cameras.

No data ships with the reference (its checkpoints are downloaded), so every
benchmark and test runs on this analytic shape:

    f(x) = |x| - 0.5 - sum_k a_k exp(-|x - c_k|^2 / (2 s_k^2)),   k = 1..8

with c_k = 0.5 u_k (u_k uniform on the unit sphere), a_k ~ U(0.03, 0.08),
s_k ~ U(0.05, 0.15), drawn from numpy's PCG64 seeded with the master seed
668123 (the reference trainers' default seed, train_sdf.py:131) plus a
per-use offset.  "N(0, 0.01)" surface jitter is read as standard deviation
0.01.
"""
import math

import numpy as np
import torch

MASTER_SEED = 668123


class BumpySphere:
    """Analytic SDF-like target with closed-form gradient (float64 on any device)."""

    def __init__(self, seed=MASTER_SEED, n_bumps=8, radius=0.5):
        rng = np.random.Generator(np.random.PCG64(seed))
        u = rng.normal(size=(n_bumps, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        self.radius = radius
        self.centers = radius * u
        self.amps = rng.uniform(0.03, 0.08, size=n_bumps)
        self.sigmas = rng.uniform(0.05, 0.15, size=n_bumps)

    def _params(self, x):
        dt = x.dtype if x.dtype == torch.float64 else torch.float64
        c = torch.as_tensor(self.centers, dtype=dt, device=x.device)
        a = torch.as_tensor(self.amps, dtype=dt, device=x.device)
        s = torch.as_tensor(self.sigmas, dtype=dt, device=x.device)
        return c, a, s

    def sdf(self, x):
        """f(x) for x [..., 3] (computed in float64, returned in x's dtype)."""
        x64 = x.to(torch.float64)
        c, a, s = self._params(x64)
        d2 = ((x64[..., None, :] - c) ** 2).sum(-1)
        f = x64.norm(dim=-1) - self.radius - (a * torch.exp(-d2 / (2 * s * s))).sum(-1)
        return f.to(x.dtype)

    def grad(self, x):
        """grad f(x) [..., 3]."""
        x64 = x.to(torch.float64)
        c, a, s = self._params(x64)
        diff = x64[..., None, :] - c
        d2 = (diff ** 2).sum(-1)
        w = a * torch.exp(-d2 / (2 * s * s)) / (s * s)
        g = x64 / x64.norm(dim=-1, keepdim=True).clamp_min(1e-12) + (w[..., None] * diff).sum(-2)
        return g.to(x.dtype)

    def normals(self, x):
        g = self.grad(x)
        return g / g.norm(dim=-1, keepdim=True).clamp_min(1e-12)

    def surface_points(self, n, seed, device="cpu"):
        """n points on f = 0 along uniformly random directions (float64 bisection)."""
        rng = np.random.Generator(np.random.PCG64(seed))
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        u = torch.as_tensor(u, dtype=torch.float64, device=device)
        lo = torch.full((n,), 0.2, dtype=torch.float64, device=device)
        hi = torch.full((n,), 1.0, dtype=torch.float64, device=device)
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            inside = self.sdf(mid[:, None] * u) < 0
            lo = torch.where(inside, mid, lo)
            hi = torch.where(inside, hi, mid)
        return 0.5 * (lo + hi)[:, None] * u


def training_batch(shape, n, seed, device="cpu", jitter=0.01):
    """Config-1 batch: n/2 uniform in [-1,1]^3 + n/2 jittered surface points,
    with target SDF and unit normals (float32)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_u = n // 2
    n_s = n - n_u
    uni = torch.as_tensor(rng.uniform(-1.0, 1.0, size=(n_u, 3)), dtype=torch.float64)
    surf = shape.surface_points(n_s, seed + 1).cpu()
    surf = surf + torch.as_tensor(rng.normal(0.0, jitter, size=(n_s, 3)), dtype=torch.float64)
    x = torch.cat([uni, surf]).to(device)
    return x.float(), shape.sdf(x).float(), shape.normals(x).float()


def seeded_siren(hidden, omega, seed, last_scale=1.0, device="cpu", SIREN=None):
    """A SIREN with the reference initialisation under torch.manual_seed(seed),
    frequency baked (update_omegas(1), as checkpoints are saved) and the last
    layer's weights scaled by ``last_scale``."""
    if SIREN is None:
        from .model import SIREN
    gen_state = torch.random.get_rng_state()
    torch.manual_seed(seed)
    try:
        m = SIREN(3, 1, list(hidden), w0=omega, ww=omega)
    finally:
        torch.random.set_rng_state(gen_state)
    m.update_omegas(1)
    if last_scale != 1.0:
        with torch.no_grad():
            m.net[-1][0].weight.mul_(last_scale)
    return m.to(device)


# Config-2 levels: (hidden, omega, seed offset, last-layer scale, band delta)
MEDIUM = ([256, 256, 256], 60.0, 2, 0.1)
FINE = ([512, 512, 512], 90.0, 3, 0.1)
DELTAS = (0.05, 0.02)


def residual_levels(device="cpu"):
    """Seeded medium and fine residual levels of the three-level workload."""
    med = seeded_siren(MEDIUM[0], MEDIUM[1], MASTER_SEED + MEDIUM[2], MEDIUM[3], device)
    fine = seeded_siren(FINE[0], FINE[1], MASTER_SEED + FINE[2], FINE[3], device)
    return med, fine


# ---------------------------------------------------------------------------
# Cameras
# ---------------------------------------------------------------------------
def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """View matrix (glm::lookAt convention), float64 numpy, row-major math layout."""
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def perspective(fovy_deg, aspect, near=0.1, far=10.0):
    """Projection matrix (glm::perspective, OpenGL clip space)."""
    t = math.tan(math.radians(fovy_deg) / 2)
    m = np.zeros((4, 4))
    m[0, 0] = 1 / (aspect * t)
    m[1, 1] = 1 / t
    m[2, 2] = -(far + near) / (far - near)
    m[2, 3] = -2 * far * near / (far - near)
    m[3, 2] = -1
    return m


def orbit_cameras(n, seed, radius=2.5, fovy=45.0, aspect=16 / 9, elev_deg=30.0):
    """n cameras on a sphere of `radius` looking at the origin: azimuth
    U[0, 2pi), elevation U[-elev, elev].  Returns a list of (inv_view, inv_proj)
    float64 4x4 matrices."""
    rng = np.random.Generator(np.random.PCG64(seed))
    az = rng.uniform(0, 2 * math.pi, size=n)
    el = np.radians(rng.uniform(-elev_deg, elev_deg, size=n))
    cams = []
    for a, e in zip(az, el):
        eye = radius * np.array([math.cos(e) * math.sin(a), math.sin(e), math.cos(e) * math.cos(a)])
        cams.append((np.linalg.inv(look_at(eye)), np.linalg.inv(perspective(fovy, aspect))))
    return cams


def _hidden_features(model, x):
    """Last sine layer's activations of a baked SIREN (float64 torch ops; data
    construction only)."""
    h = x.to(torch.float64)
    for seq in model.net[:-1]:
        lin = seq[0]
        h = torch.sin(float(seq[1].w0) * (h @ lin.weight.detach().to(torch.float64).T + lin.bias.detach().to(torch.float64)))
    return h


def fit_output_layer(model, shape, n=20000, seed=0, ridge=1e-6):
    """Solve the output layer of ``model`` by ridge least squares so that the
    network approximates ``shape.sdf`` (random-feature regression on the
    config-1 point mix).  Cheap stand-in for a full fit when many models need
    a surface."""
    dev = model.net[0][0].weight.device
    x, y, _ = training_batch(shape, n, seed, device=dev)
    h = _hidden_features(model, x)
    a = torch.cat([h, torch.ones(h.shape[0], 1, dtype=h.dtype, device=h.device)], 1)
    lhs = a.T @ a + ridge * a.shape[0] * torch.eye(a.shape[1], dtype=a.dtype, device=a.device)
    sol = torch.linalg.solve(lhs, a.T @ y.to(torch.float64))
    last = model.net[-1][0]
    with torch.no_grad():
        last.weight.copy_(sol[:-1].to(torch.float32)[None, :])
        last.bias.copy_(sol[-1:].to(torch.float32))
    return model


def tab4_levels(device="cpu", seed=MASTER_SEED + 600):
    """Levels shaped like the rows of the renderer benchmark (width, hidden
    layers): coarse (128, 2) and the single-SIREN baseline (400, 4), omega 30,
    with output layers fitted to the bumpy sphere; residual medium (256, 2),
    omega 60, and fine (400, 2), omega 90, seeded with last layers x 0.1."""
    shape = BumpySphere()
    return {"coarse": fit_output_layer(seeded_siren([128] * 2, 30.0, seed, device=device), shape, seed=seed),
            "medium": seeded_siren([256] * 2, 60.0, seed + 1, last_scale=0.1, device=device),
            "fine": seeded_siren([400] * 2, 90.0, seed + 2, last_scale=0.1, device=device),
            "siren": fit_output_layer(seeded_siren([400] * 4, 30.0, seed + 3, device=device), shape, seed=seed + 3)}


def uneven_workload(n_models=24, seed=MASTER_SEED + 500, device="cpu"):
    """Batched uneven workload: per model a width in {64, 128, 256, 512}, 1-3
    levels, its own bumpy-sphere target (sub-seed), a grid resolution in
    {128, 192, 256, 384, 512} and one orbit camera.  The coarse level (width x 3,
    omega 30) gets its output layer fitted to the target; residual levels are
    seeded (omega 60, 90; last layer x 0.1) with bands (0.05, 0.02)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(n_models):
        width = int(rng.choice([64, 128, 256, 512]))
        n_levels = int(rng.integers(1, 4))
        N = int(rng.choice([128, 192, 256, 384, 512]))
        sub = seed + 1000 * (k + 1)
        shape = BumpySphere(seed=sub)
        coarse = fit_output_layer(seeded_siren([width] * 3, 30.0, sub, device=device), shape, seed=sub + 1)
        levels = [coarse]
        for i in range(1, n_levels):
            levels.append(seeded_siren([width] * 3, 30.0 * (i + 1), sub + 10 * i, last_scale=0.1, device=device))
        out.append({"levels": levels, "width": width, "N": N, "deltas": list(DELTAS[:n_levels - 1]),
                    "camera": orbit_cameras(1, seed=sub + 2)[0], "shape": shape})
    return out
