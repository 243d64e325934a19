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


def test_batch_and_launch_config_validation_without_gpu():
    """Argument checks of the batch and launch-config entry points run on
    the host before any device call."""
    from numbacs_amd import _lib

    lib = _lib.load()
    out = ctypes.c_void_p()
    flow = ctypes.c_void_p()
    assert lib.nbx_flow_create_analytic(0, ctypes.byref(flow)) == 0
    p = (ctypes.c_double * 6)(-1.0, 0.1, 0.25, 0.0, 0.6283185307179586, 0.0)
    reqs = (_lib.Request * 1)()
    assert lib.nbx_batch_create(flow, reqs, 0, p, 6, 1e-6, 1e-8, 100, 0, ctypes.byref(out)) != 0
    assert b"at least one request" in lib.nbx_last_error()
    assert lib.nbx_batch_create(flow, reqs, 1, p, 6, 1e-6, 1e-8, 100, 7, ctypes.byref(out)) != 0
    assert b"precision" in lib.nbx_last_error()
    assert lib.nbx_batch_create(flow, reqs, 1, p, 3, 1e-6, 1e-8, 100, 0, ctypes.byref(out)) != 0
    assert b"6 parameters" in lib.nbx_last_error()
    assert lib.nbx_set_launch_config(2, 0, 3, 64, 0) != 0
    assert lib.nbx_set_launch_config(0, 0, 2, 64, 0) != 0
    assert lib.nbx_set_launch_config(2, 0, 2, 96, 0) != 0
    assert lib.nbx_set_launch_config(2, 0, 2, 64, 65) != 0
    g, b, q = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.nbx_get_launch_config(2, 1, ctypes.byref(g), ctypes.byref(b), ctypes.byref(q)) == 0
    assert g.value in (1, 2, 4) and b.value in (64, 128, 256) and q.value >= 0
    assert lib.nbx_set_tile_order(2, 1) != 0 and b"precision" in lib.nbx_last_error()
    assert lib.nbx_set_tile_order(0, 2) != 0 and b"edge_first" in lib.nbx_last_error()
    e = ctypes.c_int(-1)
    assert lib.nbx_get_tile_order(0, ctypes.byref(e)) == 0 and e.value in (0, 1)
    assert lib.nbx_set_tile_order(1, 0) == 0
    assert lib.nbx_get_tile_order(1, ctypes.byref(e)) == 0 and e.value == 0
    assert lib.nbx_set_tile_order(1, 1) == 0
    assert lib.nbx_mlp_gemm(0, None, None, 65, 64, 32, None, None, None, None, None) != 0
    assert b"multiples" in lib.nbx_last_error()
    lib.nbx_flow_destroy(flow)


def test_noalloc_ridges_rejects_grids_beyond_its_scratch():
    """nbx_ridges_noalloc refuses a grid whose scratch exceeds its fixed
    array on the host, before any device call (dummy device addresses are
    never touched)."""
    from numbacs_amd import _lib

    lib = _lib.load()
    nr, npt = ctypes.c_int64(), ctypes.c_int64()
    dummy = ctypes.c_void_p(256)
    rc = lib.nbx_ridges_noalloc(dummy, dummy, dummy, 1000, dummy, 1000, 1.0, 1.0, 0.0, 90.0, 3,
                                dummy, dummy, ctypes.byref(nr), ctypes.byref(npt), None)
    assert rc != 0
    assert b"scratch" in lib.nbx_last_error() and b"nbx_ridges" in lib.nbx_last_error()
    assert nr.value == 0 and npt.value == 0
