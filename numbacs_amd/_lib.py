"""Loader for libnumbacs_amd.so (the HIP kernels behind the C ABI declared in
include/numbacs_amd.h) and the device-memory plumbing shared by the Python
entry points.

There is no CPU fallback: if the library or a GPU is missing every entry
point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnumbacs_amd.so")

NBX_FLOW_DOUBLE_GYRE = 0
NBX_FLOW_BICKLEY_JET = 1
NBX_FLOW_SPLINE3 = 2
NBX_FLOW_LINEAR3 = 3

NBX_FP64 = 0
NBX_FP32_DISP = 1

EXTRAP = {"constant": 0, "linear": 1, "nearest": 2}

_c_d = ctypes.c_double
_c_i = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_p = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/numbacs_amd.h
SIGNATURES = {
    "nbx_last_error": (ctypes.c_char_p, []),
    "nbx_version": (_c_i, []),
    "nbx_flow_create_analytic": (_c_i, [_c_i, _c_p]),
    "nbx_flow_create_spline3": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_i, _c_d, _c_p, _c_p]),
    "nbx_flow_create_linear3": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_i, _c_d, _c_p, _c_p]),
    "nbx_ingest_xy": (_c_i, [_c_p, _c_i, _c_i, _c_i, _c_d, _c_p, _c_p]),
    "nbx_flow_destroy": (None, [_c_p]),
    "nbx_prefilter3": (_c_i, [_c_p, _c_i, _c_i, _c_i, _c_p, _c_p]),
    "nbx_flow_eval": (_c_i, [_c_p, _c_p, _c_i, _c_p, _c_p, _c_i64, _c_p, _c_p]),
    "nbx_flowmap_grid_2d": (_c_i, [_c_p, _c_d, _c_d, _c_p, _c_i, _c_p, _c_i, _c_p, _c_i, _c_d,
                                   _c_d, _c_p, _c_i64, _c_i, _c_p, _c_p, _c_p, _c_p]),
    "nbx_flowmap_pts": (_c_i, [_c_p, _c_d, _c_d, _c_p, _c_i64, _c_p, _c_i, _c_d, _c_d, _c_p,
                               _c_i64, _c_i, _c_p, _c_p, _c_p, _c_p]),
    "nbx_flowmap_aux_grid_2d": (_c_i, [_c_p, _c_d, _c_d, _c_p, _c_i, _c_p, _c_i, _c_d, _c_i,
                                       _c_i, _c_p, _c_i, _c_d, _c_d, _c_p, _c_i64, _c_i, _c_p,
                                       _c_p, _c_p, _c_p]),
    "nbx_flowmap_compose": (_c_i, [_c_p, _c_i, _c_i, _c_i, _c_d, _c_d, _c_i, _c_d, _c_d, _c_i,
                                   _c_p, _c_p]),
    "nbx_set_launch_config": (_c_i, [_c_i, _c_i, _c_i, _c_i, _c_i]),
    "nbx_get_launch_config": (_c_i, [_c_i, _c_i, _c_p, _c_p, _c_p]),
    "nbx_set_tile_order": (_c_i, [_c_i, _c_i]),
    "nbx_get_tile_order": (_c_i, [_c_i, _c_p]),
    "nbx_batch_create": (_c_i, [_c_p, _c_p, _c_i, _c_p, _c_i, _c_d, _c_d, _c_i64, _c_i, _c_p]),
    "nbx_batch_particles": (_c_i64, [_c_p]),
    "nbx_batch_run": (_c_i, [_c_p, _c_p]),
    "nbx_batch_capture": (_c_i, [_c_p, _c_p]),
    "nbx_batch_run_graph": (_c_i, [_c_p, _c_p]),
    "nbx_batch_destroy": (None, [_c_p]),
    "nbx_mlp_gemm": (_c_i, [_c_i, _c_p, _c_p, _c_i, _c_i, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "nbx_mlp_features": (_c_i, [_c_p, _c_i64, _c_p, _c_p, _c_p, _c_i, _c_p, _c_p]),
    "nbx_mlp_out_forward": (_c_i, [_c_p, _c_i64, _c_i, _c_p, _c_p, _c_p, _c_p]),
    "nbx_mlp_loss": (_c_i, [_c_p, _c_p, _c_i64, _c_p, _c_p, _c_p]),
    "nbx_mlp_out_backward": (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_p, _c_p, _c_p, _c_p,
                                    _c_p]),
    "nbx_mlp_colsum": (_c_i, [_c_p, _c_i64, _c_i, _c_p, _c_p]),
    "nbx_mlp_adam": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, ctypes.c_float, ctypes.c_float,
                            ctypes.c_float, ctypes.c_float, _c_i, ctypes.c_float, _c_p]),
    "nbx_mlp_adam_sched": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, _c_p, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float, _c_p]),
    "nbx_mlp_to_bf16": (_c_i, [_c_p, _c_i64, _c_p, _c_p]),
    "nbx_mlp_infer": (_c_i, [_c_p, _c_p, _c_i64, _c_p, _c_p]),
    "nbx_ftle_grid_2d": (_c_i, [_c_p, _c_i, _c_i, _c_d, _c_d, _c_d, _c_p, _c_p, _c_p]),
    "nbx_C_eig_2d": (_c_i, [_c_p, _c_i, _c_i, _c_d, _c_d, _c_p, _c_p, _c_p, _c_p]),
    "nbx_C_tensor_aux_2d": (_c_i, [_c_p, _c_i, _c_i, _c_i, _c_d, _c_p, _c_p, _c_p]),
    "nbx_C_eig_aux_2d": (_c_i, [_c_p, _c_i, _c_i, _c_i, _c_d, _c_d, _c_d, _c_i, _c_p, _c_p,
                                _c_p, _c_p]),
    "nbx_ftle_eig_2d": (_c_i, [_c_p, _c_i, _c_i, _c_d, _c_d, _c_d, _c_p, _c_p, _c_p, _c_p]),
    "nbx_percentile": (_c_i, [_c_p, _c_i64, _c_d, _c_p, _c_p]),
    "nbx_ridge_pts": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_p, _c_i, _c_d, _c_d, _c_d, _c_d, _c_p,
                             _c_p, _c_p]),
    "nbx_ridges": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_p, _c_i, _c_d, _c_d, _c_d, _c_d, _c_i, _c_p,
                          _c_p, _c_p, _c_p, _c_p]),
    "nbx_ridges_async": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_p, _c_i, _c_d, _c_d, _c_d, _c_d, _c_i,
                                _c_p, _c_p, _c_p, _c_p]),
    "nbx_ridges_noalloc": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_p, _c_i, _c_d, _c_d, _c_d, _c_d, _c_i,
                                  _c_p, _c_p, _c_p, _c_p, _c_p]),
    "nbx_ordered_ridges": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_p, _c_i, _c_d, _c_d, _c_i, _c_d,
                                  _c_d, _c_d, _c_p, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p]),
    "nbx_ordered_ridges_link": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_i, _c_d, _c_d, _c_d, _c_i,
                                       _c_d, _c_d, _c_p, _c_i64, _c_p, _c_i64, _c_p, _c_p]),
    "nbx_label8": (_c_i, [_c_p, _c_i, _c_i, _c_p, _c_p, _c_p]),
}

class Request(ctypes.Structure):
    """struct nbx_request (include/numbacs_amd.h)."""
    _fields_ = [("t0", _c_d), ("T", _c_d), ("x", _c_p), ("nx", _c_i), ("y", _c_p), ("ny", _c_i),
                ("mask", _c_p), ("flowmap", _c_p), ("status", _c_p), ("nsteps", _c_p),
                ("ftle", _c_p), ("eigvec_max", _c_p), ("dx", _c_d), ("dy", _c_d)]


class MlpNet(ctypes.Structure):
    """struct nbx_mlp_net (include/numbacs_amd.h)."""
    _fields_ = [("shift", _c_p), ("scale", _c_p), ("Bf", _c_p), ("half", _c_i), ("width", _c_i),
                ("layers", _c_i), ("WT", _c_p * 4), ("b", _c_p * 4), ("Wout", _c_p),
                ("bout", _c_p), ("out_scale", _c_p)]


_lib = None


def load():
    """Load the shared library (no GPU needed); raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"numbacs_amd: {LIB_PATH} is not built; run __graft_entry__.build() "
                "(make -C numbacs_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().nbx_last_error().decode(errors="replace")
        raise RuntimeError(f"numbacs_amd error {rc}: {msg}")


_torch = None


def torch():
    """torch, with a usable ROCm GPU, or a RuntimeError."""
    global _torch
    if _torch is None:
        import torch as _t

        if not _t.cuda.is_available():
            raise RuntimeError("numbacs_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        _torch = _t
    return _torch


def device():
    t = torch()
    return t.device("cuda", t.cuda.current_device())


def stream_ptr():
    return torch().cuda.current_stream().cuda_stream


def ptr(t):
    """Raw device address of a tensor (None for None)."""
    return None if t is None else t.data_ptr()


def to_dev(a, dtype=np.float64):
    """Contiguous device copy of a host array (or pass a device tensor through)."""
    t = torch()
    if isinstance(a, t.Tensor):
        tt = {np.float64: t.float64, np.uint8: t.uint8, np.int32: t.int32}[dtype]
        return a.to(device=device(), dtype=tt).contiguous()
    arr = np.ascontiguousarray(a, dtype=dtype)
    return t.from_numpy(arr).to(device())


def empty(shape, dtype="f64"):
    t = torch()
    tt = {"f64": t.float64, "i32": t.int32, "u8": t.uint8, "i64": t.int64, "f32": t.float32}[dtype]
    return t.empty(shape, dtype=tt, device=device())


def zeros(shape, dtype="f64"):
    t = torch()
    tt = {"f64": t.float64, "i32": t.int32, "u8": t.uint8, "i64": t.int64, "f32": t.float32}[dtype]
    return t.zeros(shape, dtype=tt, device=device())


def to_host(t):
    return t.detach().cpu().numpy()
