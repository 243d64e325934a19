"""The C-ABI library loads and exports every symbol include/numbacs_amd.h
declares (no compute calls: no GPU needed)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "numbacs_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("nbx_flowmap_grid_2d", "nbx_ftle_grid_2d", "nbx_ridges", "nbx_flow_create_spline3"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from numbacs_amd import _lib

    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from numbacs_amd import _lib

    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_error_path_without_gpu_work():
    """Argument validation happens on the host and reports through
    nbx_last_error()."""
    from numbacs_amd import _lib

    lib = _lib.load()
    out = ctypes.c_void_p()
    assert lib.nbx_flow_create_analytic(7, ctypes.byref(out)) != 0
    assert b"unknown" in lib.nbx_last_error()
    assert lib.nbx_flow_create_analytic(0, ctypes.byref(out)) == 0
    lib.nbx_flow_destroy(out)
