"""Batched extraction (mp_extract_batch_*): every model of an uneven batch
(widths 64-512, 1-3 levels, mixed resolutions, culled and unculled) must give
exactly the volume and counts of its own mp_extract_volume call, eagerly and
on every replay of the captured graph."""
import pytest
import torch

from mplicits.batch import BatchedExtraction
from mplicits.stack import LevelStack
from mplicits.synthetic import seeded_siren

pytestmark = pytest.mark.gpu


def _model(width, levels, seed):
    out = [seeded_siren([width] * 3, 30.0, seed=seed, last_scale=20.0).cuda()]
    for i in range(1, levels):
        out.append(seeded_siren([width] * 2, 30.0 * (i + 1), seed=seed + 10 * i, last_scale=0.1).cuda())
    return out


# (width, levels, N, deltas)
JOBS = [
    (64, 1, 24, []),
    (128, 2, 33, [0.05]),
    (256, 3, 40, [0.05, 0.02]),
    (512, 2, 29, [0.04]),
    (256, 3, 31, [0.02, 0.08]),  # wider fine band: the per-chunk break matters
    (128, 3, 20, []),            # unculled multi-level
]


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_batch_matches_single_model_extraction(precision):
    chunk = 4096
    stacks = [LevelStack(_model(w, l, 100 + k), precision) for k, (w, l, _, _) in enumerate(JOBS)]
    refs, ref_counts = [], []
    for st, (_, _, N, d) in zip(stacks, JOBS):
        vol, counts = st.extract_volume(N, d, chunk=chunk)
        refs.append(vol.clone())
        ref_counts.append(list(counts))
    b = BatchedExtraction(stacks, [j[2] for j in JOBS], [j[3] for j in JOBS], chunk=chunk)
    b.run(graph=False)
    assert b.counts() == ref_counts
    for j, ref in enumerate(refs):
        assert torch.equal(b.volume(j), ref), f"model {j}"

    # captured graph: replays with eager work in between, each checked
    probe = torch.rand(1000, 3, device="cuda") * 2 - 1
    for rep in range(3):
        for v in b.volumes:
            v.fill_(float("nan"))
        stacks[2].eval(probe)  # unrelated eager work on the same stacks
        b.run(graph=True)
        torch.cuda.synchronize()
        assert b.counts() == ref_counts, f"replay {rep}"
        for j, ref in enumerate(refs):
            assert torch.equal(b.volume(j), ref), f"replay {rep}, model {j}"
    b.close()


def test_batch_rejects_mixed_precision():
    a = LevelStack(_model(64, 1, 1), "bf16")
    c = LevelStack(_model(64, 1, 2), "fp32")
    with pytest.raises(ValueError):
        BatchedExtraction([a, c], [16, 16], [[], []])


@pytest.mark.parametrize("precision", ["bf16", "bf16x3"])
def test_trace_batch_matches_single_model_trace(precision):
    from mplicits.synthetic import orbit_cameras
    from mplicits.tracer import BatchedTrace, trace
    specs = [(64, 1, (40, 24), []), (256, 3, (33, 20), [0.05, 0.02]), (512, 2, (48, 27), [0.04]),
             (128, 2, (20, 31), [0.05]), (128, 4, (36, 22), [0.05, 0.03, 0.02])]
    stacks = [LevelStack(_model(w, l, 300 + k), precision) for k, (w, l, _, _) in enumerate(specs)]
    cams = orbit_cameras(len(specs), seed=5, aspect=1.5)
    refs = []
    for st, cam, (_, _, (W, H), d) in zip(stacks, cams, specs):
        out = trace(st, [cam], W, H, deltas=d, max_steps=64)
        refs.append({k: (v[0].clone() if isinstance(v, torch.Tensor) else v) for k, v in out.items()})
    b = BatchedTrace([(st, cam, W, H, d) for st, cam, (_, _, (W, H), d) in zip(stacks, cams, specs)],
                     max_steps=64)
    for rep in range(3):
        for t in b.depth + b.rgba:
            t.fill_(7)
        b.run(graph=rep > 0)
        torch.cuda.synchronize()
        st = b.stats()
        for j, ref in enumerate(refs):
            assert torch.equal(b.depth[j], ref["depth"]), f"run {rep}, job {j}"
            assert torch.equal(b.steps[j], ref["steps"]), f"run {rep}, job {j}"
            assert torch.equal(b.rgba[j], ref["rgba"]), f"run {rep}, job {j}"
            assert torch.equal(b.normals[j], ref["normals"]), f"run {rep}, job {j}"
            assert st[j]["hits"] == ref["hits"] and st[j]["evals"] == ref["evals"]
    b.close()


def test_trace_batch_bf16x3f_matches_single_model_trace():
    # the mixed three-term path in the batched tracer: every level up to
    # width 256 with three hidden layers (the 512-wide residual runs FP16)
    from mplicits.synthetic import orbit_cameras
    from mplicits.tracer import BatchedTrace, trace
    models = [_model(64, 1, 400), _model(128, 1, 401), _model(256, 1, 402),
              _model(256, 1, 403) + [seeded_siren([512] * 2, 60.0, seed=404, last_scale=0.1).cuda()]]
    sizes, deltas = [(40, 24), (33, 20), (36, 22), (48, 27)], [[], [], [], [0.05]]
    stacks = [LevelStack(m, "bf16x3f") for m in models]
    cams = orbit_cameras(len(models), seed=6, aspect=1.5)
    refs = [trace(st, [cam], W, H, deltas=d, max_steps=64) for st, cam, (W, H), d in zip(stacks, cams, sizes, deltas)]
    b = BatchedTrace([(st, cam, W, H, d) for st, cam, (W, H), d in zip(stacks, cams, sizes, deltas)], max_steps=64)
    for rep in range(2):
        b.run(graph=rep > 0)
        torch.cuda.synchronize()
        for j, ref in enumerate(refs):
            for k in ("depth", "steps", "rgba", "normals"):
                assert torch.equal(getattr(b, k)[j], ref[k][0]), (rep, j, k)
    b.close()


def test_autotune_caches_and_applies(tmp_path):
    from mplicits import autotune
    cache = str(tmp_path / "tiles.txt")
    cols = autotune.best_tiles(256, 3, "bf16", cache=cache)
    assert cols in (1, 2)
    assert open(cache).read().split() == ["0", "256", "3", "1", str(cols)]
    assert autotune.best_tiles(256, 3, "bf16", cache=cache) == cols  # cached: no second line
    assert len(open(cache).read().splitlines()) == 1
    st = LevelStack(_model(256, 2, 9), "bf16")
    assert autotune.tune_stack(st, cache=cache) == cols
    x = torch.rand(4096, 3, device="cuda") * 2 - 1
    ref = LevelStack(_model(256, 2, 9), "bf16").eval(x)
    assert torch.equal(st.eval(x), ref)  # the tile shape changes the schedule, not the values
    # the other families: one measured line each, applied to the stack
    got = autotune.tune_stack(st, cache=cache, families=("batch", "trace"))
    assert got["batch"] in (1, 2) and got["trace"] in (1, 2)
    lines = [ln.split() for ln in open(cache).read().splitlines()]
    assert [ln[0] for ln in lines] == ["0", "1", "2"]


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_batch_tile_shapes_change_no_value(precision):
    """Every job of the uneven batch under the halved tile shape of the
    batched kernel gives the volumes of the default shape, bit for bit; the
    tracer's one- and two-block tiles give the same frames."""
    from mplicits.synthetic import orbit_cameras
    from mplicits.tracer import BatchedTrace
    stacks = [LevelStack(_model(w, l, 400 + k), precision) for k, (w, l, _, _) in enumerate(JOBS)]
    runs = []
    for tiles in (1, 2):
        for st in stacks:
            st.set_tiles(tiles, family="batch")
        b = BatchedExtraction(stacks, [N for _, _, N, _ in JOBS], [d for _, _, _, d in JOBS], chunk=4096)
        b.run()
        torch.cuda.synchronize()
        runs.append(([b.volume(j).clone() for j in range(len(JOBS))], b.counts()))
        b.close()
    for j in range(len(JOBS)):
        assert torch.equal(runs[0][0][j], runs[1][0][j]), j
    assert runs[0][1] == runs[1][1]
    cams = orbit_cameras(3, seed=8, aspect=1.5)
    frames = []
    for tiles in (1, 2):
        for st in stacks[:3]:
            st.set_tiles(tiles, family="trace")
        bt = BatchedTrace([(st, cam, 36, 24, d) for st, cam, (_, _, _, d) in zip(stacks[:3], cams, JOBS[:3])],
                          max_steps=64)
        bt.run()
        torch.cuda.synchronize()
        frames.append([t.clone() for t in bt.depth])
        bt.close()
    for a, b in zip(*frames):
        assert torch.equal(a, b)
