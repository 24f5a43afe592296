"""The C-ABI library loads and exports every symbol include/mplicits.h declares
(no compute: these run on hosts without a GPU)."""
import ctypes
import os.path as osp
import re

from conftest import ROOT


def _declared_symbols():
    text = open(osp.join(ROOT, "include", "mplicits.h")).read()
    return sorted(set(re.findall(r"\b(mp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    syms = _declared_symbols()
    for s in ("mp_abi_version", "mp_last_error", "mp_stack_create", "mp_eval", "mp_extract_volume"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in _declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    from mplicits import _lib
    assert set(_declared_symbols()) == set(_lib.SIGNATURES)


def test_header_constants_match_the_binding():
    # every precision, tile family and error code the header defines has the
    # same value in the Python binding (the precision names map onto them)
    from mplicits import _lib
    text = open(osp.join(ROOT, "include", "mplicits.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define (MP_[A-Z0-9_]+) \(?(-?\d+)\)?", text)}
    for name in ("MP_FP32", "MP_BF16", "MP_FP16", "MP_BF16X3", "MP_BF16X3F", "MP_OK", "MP_EINVAL",
                 "MP_EUNSUPPORTED", "MP_TILES_EXTRACT", "MP_TILES_BATCH", "MP_TILES_TRACE"):
        assert getattr(_lib, name) == defs[name], name
    assert _lib.PRECISIONS == {"fp32": defs["MP_FP32"], "bf16": defs["MP_BF16"], "fp16": defs["MP_FP16"],
                               "bf16x3": defs["MP_BF16X3"], "bf16x3f": defs["MP_BF16X3F"]}


def test_abi_version_and_error_path(lib):
    from mplicits import _lib
    assert lib.mp_abi_version() == 1
    # invalid arguments are rejected on the host, before any device work
    out = ctypes.c_void_p()
    rc = lib.mp_stack_create(None, 0, 0, None, ctypes.byref(out))
    assert rc == _lib.MP_EINVAL
    assert b"n_levels" in lib.mp_last_error()
    rc = lib.mp_grid_coords(1, 0, 0, 0, None, None)
    assert rc == _lib.MP_EINVAL
    try:
        _lib.check(rc)
    except ValueError as e:
        assert "N must be" in str(e)
    else:
        raise AssertionError("check() did not raise")


def test_training_data_and_renderer_arguments_are_checked_on_the_host(lib):
    import numpy as np
    from mplicits import _lib
    _lib.load()  # argument tables
    out = ctypes.c_void_p()
    v = np.zeros((3, 3), np.float32)
    f = np.array([[0, 1, 3]], np.int32)  # vertex 3 of 3
    rc = lib.mp_mesh_create(v.ctypes.data_as(ctypes.c_void_p), 3, f.ctypes.data_as(ctypes.c_void_p), 1, None,
                            ctypes.byref(out))
    assert rc == _lib.MP_EINVAL and b"indexes vertex 3" in lib.mp_last_error() and not out.value
    rc = lib.mp_mesh_create(v.ctypes.data_as(ctypes.c_void_p), 3, f.ctypes.data_as(ctypes.c_void_p), 0, None,
                            ctypes.byref(out))
    assert rc == _lib.MP_EINVAL and b"empty mesh" in lib.mp_last_error()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the checks come first
    rc = lib.mp_neighborhood_samples(dummy, 10, 6, None, dummy, 4, 0.02, 5e-4, 1, 1, dummy, dummy, dummy, None)
    assert rc == _lib.MP_EINVAL and b"seventh column" in lib.mp_last_error()
    rc = lib.mp_neighborhood_samples(dummy, 10, 6, None, dummy, 11, 0.02, 5e-4, 0, 1, dummy, dummy, dummy, None)
    assert rc == _lib.MP_EINVAL and b"without ids" in lib.mp_last_error()
    assert lib.mp_train_param_count(None, 0) == -1
    rc = lib.mp_train_fwd(None, 0, dummy, 16, dummy, dummy, 1, None)
    assert rc == _lib.MP_EINVAL and b"stack is NULL" in lib.mp_last_error()
    rc = lib.mp_train_bwd(None, 0, dummy, 16, dummy, None, dummy, None)
    assert rc == _lib.MP_EINVAL and b"stack is NULL" in lib.mp_last_error()
    rc = lib.mp_render(None, dummy, 1, None, dummy, None, None, None, None, None)
    assert rc == _lib.MP_EINVAL and b"NULL argument" in lib.mp_last_error()
