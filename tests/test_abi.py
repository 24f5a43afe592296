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
