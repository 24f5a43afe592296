"""mp_extract_volume_noalloc (work lists in the library's fixed device array)
against mp_extract_volume (work lists in the stack's allocated scratch), both
called through the C ABI: bit-identical volumes and level counts on two calls
in a row, the second smaller, plus the host-side size limits."""
import ctypes

import pytest
import torch

from mplicits import _lib
from mplicits.stack import LevelStack
from mplicits.synthetic import seeded_siren

pytestmark = pytest.mark.gpu


def _levels():
    coarse = seeded_siren([128] * 3, 30.0, seed=31, last_scale=20.0).cuda()
    med = seeded_siren([128] * 3, 60.0, seed=32, last_scale=0.1).cuda()
    fine = seeded_siren([256] * 3, 90.0, seed=33, last_scale=0.1).cuda()
    return [coarse, med, fine]


def _call(lib, fn, st, N, deltas, chunk):
    vol = torch.full((N, N, N), float("nan"), device="cuda", dtype=torch.float32)
    d_arr = (ctypes.c_float * max(1, len(deltas)))(*deltas)
    counts = (ctypes.c_int64 * st.n_levels)()
    print(f"{fn}: N={N} deltas={deltas} chunk={chunk}", flush=True)
    rc = getattr(lib, fn)(st._handle, N, _lib.MP_GRID_REFERENCE, chunk, ctypes.cast(d_arr, _lib.c_float_p),
                          len(deltas), _lib.ptr(vol), ctypes.cast(counts, _lib.c_int64_p), _lib.stream_handle())
    _lib.check(rc)
    torch.cuda.synchronize()
    return vol, list(counts)


def test_noalloc_matches_allocating_entry_point():
    lib = _lib.load()
    st = LevelStack(_levels(), "bf16")
    st.refresh()
    # larger call first, then a smaller one: the second must not see the first's lists or counters
    for N, deltas, chunk in ((96, [0.05, 0.02], 32 ** 3), (40, [0.05, 0.02], 4096), (40, [0.02, 0.08], 512),
                             (24, [], 4096)):
        want, wc = _call(lib, "mp_extract_volume", st, N, deltas, chunk)
        got, gc = _call(lib, "mp_extract_volume_noalloc", st, N, deltas, chunk)
        assert gc == wc, (N, gc, wc)
        assert torch.equal(got, want), (N, (got - want).abs().max().item())
        if deltas:
            assert 0 < gc[1] < N ** 3


def test_noalloc_rejects_sizes_beyond_its_scratch():
    lib = _lib.load()
    st = LevelStack(_levels()[:2], "fp32")
    st.refresh()
    vol = torch.empty(8, device="cuda")
    d_arr = (ctypes.c_float * 1)(0.05)
    rc = lib.mp_extract_volume_noalloc(st._handle, 257, 0, 32768, ctypes.cast(d_arr, _lib.c_float_p), 1,
                                       _lib.ptr(vol), None, _lib.stream_handle())
    assert rc == _lib.MP_EINVAL and b"exceeds the fixed scratch" in lib.mp_last_error()
    rc = lib.mp_extract_volume_noalloc(st._handle, 256, 0, 8, ctypes.cast(d_arr, _lib.c_float_p), 1,
                                       _lib.ptr(vol), None, _lib.stream_handle())
    assert rc == _lib.MP_EINVAL and b"chunk flags" in lib.mp_last_error()
