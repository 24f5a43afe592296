"""Comparisons of GPU results with the oracle shared by the GPU tests.

Extraction: the culled volume against the restated sampling loop, with every
sample whose band decisions differ between the two (a value within rounding
of a band edge) explained and counted apart.  Tracing: traced rays against the
restated step rule, with every hit-set difference explained."""
import numpy as np

from oracle import extract as oextract
from oracle import siren as osiren
from oracle import trace as otrace


def gpu_band_masks(levels, precision, N, deltas, chunk):
    """The band decisions the GPU pipeline took: (s0, [m1, m2, ...]) from the
    same kernels run on the partial stacks (coarse alone; coarse + level 1 with
    its band; ...), with the reference loop's per-chunk break."""
    from mplicits.stack import LevelStack
    n = N ** 3
    nch = (n + chunk - 1) // chunk
    chunk_of = np.arange(n) // chunk
    alive = np.ones(nch, bool)
    masks = []
    s = None
    for i in range(1, len(levels)):
        vol, _ = LevelStack(levels[:i], precision).extract_volume(N, list(deltas[:i - 1]), chunk=chunk)
        s = vol.reshape(-1).cpu().numpy()
        if i == 1:
            s0 = s
        raw = np.abs(s) < np.float32(deltas[i - 1])
        alive = alive & np.bincount(chunk_of, weights=raw, minlength=nch).astype(bool)
        masks.append(raw & alive[chunk_of])
    return s0, masks


def check_volume_against_oracle(vol, levels, N, deltas, chunk, chunk_ids, precision, tol, edge_tol=None,
                                max_flip_frac=1e-3):
    """|vol - oracle| <= tol on every sample of the listed chunks whose band
    decisions agree; samples whose decisions differ must sit within edge_tol of
    the band edge they flipped at (or in a chunk whose break decision flipped).
    Returns (max error, mean error, flips, samples)."""
    edge_tol = tol if edge_tol is None else edge_tol
    layers = [osiren.layers_from_model(m) for m in levels]
    ref = oextract.sample_chunks(layers, N, list(deltas), chunk, chunk_ids, return_masks=True)
    flat = vol.reshape(-1).double().cpu().numpy()
    if deltas:
        _, gmasks = gpu_band_masks(levels, precision, N, deltas, chunk)
        partial = [oextract.sample_chunks(layers[:i], N, list(deltas[:i - 1]), chunk, chunk_ids)
                   for i in range(1, len(levels))]
    errs, flips, total = [], 0, 0
    for c, (idx, r, rmasks) in ref.items():
        e = np.abs(flat[idx] - r)
        total += len(idx)
        if deltas:
            flip = np.zeros(len(idx), bool)
            near = np.zeros(len(idx), bool)
            whole = np.zeros(len(idx), bool)
            for i, (rm, gm) in enumerate(zip(rmasks, gmasks)):
                d = rm != gm[idx]
                flip |= d
                s_before = partial[i][c][1]
                near |= d & (np.abs(np.abs(s_before) - np.float32(deltas[i])) <= edge_tol)
                # a chunk whose only band sample flipped breaks (or not) as a whole
                whole |= d & (rm.any() != gm[idx].any())
            assert (near | whole)[flip].all(), (c, np.nonzero(flip & ~(near | whole))[0][:8])
            flips += int(flip.sum())
            e = e[~flip]
        errs.append(e)
    e = np.concatenate(errs)
    assert e.max() <= tol, (precision, float(e.max()), tol)
    assert flips <= max_flip_frac * total, (precision, flips, total)
    return float(e.max()), float(e.mean()), flips, total


def check_trace_against_oracle(out, levels, deltas, cams, W, H, n_sample, seed, prec_err=0.0, max_steps=128,
                               eps=1e-4, min_hits=50):
    """The GPU trace of a sample of rays against the FP64 oracle.

    * every GPU hit lies where the FP64 stack is within eps + prec_err of zero;
    * rays that both hit: |depth difference| <= (2 eps + prec_err) / |grad F . d|
      (each stops somewhere in |F| < eps) where the ray meets the surface at more
      than 11 degrees and the oracle's ray never stepped across the level set;
    * a ray that hits in one and not the other must be explained: the oracle's
      ray came within a threshold rounding of a hit (fmin) or of flipping a
      band decision (edge), ran out of steps, stepped across the level set
      (crossed), or meets the surface grazing; in a reduced precision
      (prec_err > 0, a different field) at most 1% of the rays may differ
      without such an explanation.
    prec_err: the stated error bound of the precision's level evaluation."""
    depth = out["depth"].cpu().numpy().reshape(-1)
    gsteps = out["steps"].cpu().numpy().reshape(-1)
    rng = np.random.default_rng(seed)
    ids = rng.choice(len(depth), n_sample, replace=False)
    P, D, t0, inside = otrace.rays(cams, W, H, ids=ids)
    layers = [osiren.layers_from_model(m) for m in levels]
    r = otrace.trace(layers, deltas, P[inside], D[inside], max_steps=max_steps, eps=eps)
    n = len(ids)
    ref_hit = np.zeros(n, bool)
    ref_hit[inside] = r["hit"]
    ref_depth = np.full(n, -1.0)
    ref_depth[inside] = t0[inside] + r["t"]
    fmin = np.full(n, np.inf)
    fmin[inside] = r["fmin"]
    edge = np.full(n, np.inf)
    edge[inside] = r["edge"]
    rsteps = np.zeros(n, np.int64)
    rsteps[inside] = r["steps"]
    crossed = np.zeros(n, bool)
    crossed[inside] = r["crossed"]
    got = depth[ids]
    got_hit = got >= 0
    eye = P - t0[:, None] * D

    def stack_value_grad(x):
        f = np.zeros(len(x))
        g = np.zeros((len(x), 3))
        for lay in layers:
            fl, gl = osiren.forward_grad(lay, x)
            f += fl
            g += gl
        return f, g

    # every GPU hit is a hit of the FP64 field
    if got_hit.any():
        xg = eye[got_hit] + got[got_hit, None] * D[got_hit]
        fg, _ = stack_value_grad(xg)
        assert np.abs(fg).max() <= eps + prec_err + 2e-6, np.abs(fg).max()
    # depths of shared hits
    both = got_hit & ref_hit
    assert both.sum() >= min_hits, both.sum()
    if not both.any():
        assert not (got_hit | ref_hit).any()
        return 0
    _, gref = stack_value_grad(eye[both] + ref_depth[both, None] * D[both])
    slope = np.abs((gref * D[both]).sum(1))
    steep = slope > 0.2 * np.linalg.norm(gref, axis=1)
    bound = (2 * eps + prec_err) / slope + 1e-5
    dd = np.abs(got[both] - ref_depth[both])
    # a ray that stepped across the level set may land on another sheet of it
    bad = steep & (dd > bound) & (edge[both] > 1e-5 + prec_err) & ~crossed[both]
    assert not bad.any(), (dd[bad], bound[bad])
    # hit-set differences
    only_gpu = got_hit & ~ref_hit
    ok = ((fmin[only_gpu] < 2 * eps + prec_err) | (edge[only_gpu] < 1e-5 + prec_err) | (rsteps[only_gpu] >= max_steps)
          | crossed[only_gpu])
    unexplained = int((~ok).sum())
    only_ref = ref_hit & ~got_hit
    if only_ref.any():
        _, g = stack_value_grad(eye[only_ref] + ref_depth[only_ref, None] * D[only_ref])
        graze = np.abs((g * D[only_ref]).sum(1)) < 0.2 * np.linalg.norm(g, axis=1)
        ok = graze | (edge[only_ref] < 1e-5 + prec_err) | (gsteps[ids[only_ref]] >= max_steps) | crossed[only_ref]
        unexplained += int((~ok).sum())
    if prec_err == 0.0:
        # FP32: every difference is explained
        assert unexplained == 0, (unexplained, ids[only_gpu | only_ref])
    else:
        # a reduced-precision field is another field (within prec_err): its
        # rays may take other paths; the differences stay few
        assert unexplained <= 0.01 * n, (unexplained, n)
    assert (only_gpu.sum() + only_ref.sum()) <= 0.02 * n, (only_gpu.sum(), only_ref.sum())
    return both.sum()


