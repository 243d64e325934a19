"""Particle flow maps (mirrors numbacs.integration).

Reference: flowmap_grid_2D (integration.py:123-182) and flowmap
(integration.py:7-61).  Same signatures, defaults, 'ij' layouts and float64
outputs; the per-particle numbalsoda.dop853 loop becomes one GPU launch of
one DOP853 solve per thread.

Deviations (documented): ``method='lsoda'`` raises NotImplementedError (the
reference silently returns zeros for unknown methods, integration.py:165-180);
a ``funcptr`` that is not a :class:`numbacs_amd.flows.FlowHandle` raises
TypeError.  ``max_steps`` and ``return_status`` are extra keyword-only knobs,
outside the reference signature.
"""
import numpy as np

from . import _lib
from .flows import FlowHandle

DEFAULT_MAX_STEPS = 100000


def _check_flow(funcptr):
    if not isinstance(funcptr, FlowHandle):
        raise TypeError("funcptr must be a flow created by numbacs_amd.flows "
                        "(get_predefined_flow / get_flow_2D); raw cfunc addresses cannot run "
                        "on the GPU")


def _check_method(method):
    m = method.lower()
    if m == "lsoda":
        raise NotImplementedError("method='lsoda' is not available on the GPU path; use 'dop853'")
    if m != "dop853":
        raise ValueError(f"unknown method {method!r}")


def _params(params):
    p = np.ascontiguousarray(params, dtype=np.float64).ravel()
    if p.size < 1:
        raise ValueError("params must hold at least the integration direction params[0]")
    return p


def flowmap_grid_2d_device(funcptr, t0, T, xd, yd, params, rtol=1e-6, atol=1e-8, maskd=None,
                           max_steps=DEFAULT_MAX_STEPS, precision=_lib.NBX_FP64, out=None,
                           status=None, nsteps=None):
    """Device-tensor form: xd (nx,), yd (ny,) float64 cuda tensors; returns
    the (nx, ny, 2) flow-map tensor (enqueued on the current stream).
    precision: 0 (NBX_FP64) or 1 (NBX_FP32_DISP: float stages and spline,
    double time, step control and anchor)."""
    _check_flow(funcptr)
    p = _params(params)
    nx, ny = int(xd.shape[0]), int(yd.shape[0])
    if out is None:
        out = _lib.empty((nx, ny, 2))
    _lib.check(_lib.load().nbx_flowmap_grid_2d(
        funcptr.handle, float(t0), float(T), _lib.ptr(xd), nx, _lib.ptr(yd), ny, p.ctypes.data,
        p.size, float(rtol), float(atol), _lib.ptr(maskd), int(max_steps), int(precision),
        _lib.ptr(out), _lib.ptr(status), _lib.ptr(nsteps), _lib.stream_ptr()))
    return out


def flowmap_grid_2D(funcptr, t0, T, x, y, params, method="dop853", rtol=1e-6, atol=1e-8,
                    mask=None, *, max_steps=DEFAULT_MAX_STEPS, return_status=False):
    """Flow map at t0+T of the grid x (nx,), y (ny,): returns (nx, ny, 2)
    float64.  Masked points (mask[i, j] True) are not integrated and stay 0.
    With return_status=True also returns the per-particle status
    (1 ok, 0 masked, -2 step budget exhausted, -3 step-size underflow)."""
    _check_flow(funcptr)
    _check_method(method)
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    nx, ny = len(x), len(y)
    if nx * ny == 0:
        out = np.zeros((nx, ny, 2))
        return (out, np.zeros((nx, ny), np.int32)) if return_status else out
    maskd = None
    if mask is not None:
        m = np.asarray(mask)
        if m.shape != (nx, ny):
            raise ValueError(f"mask must have shape {(nx, ny)}")
        maskd = _lib.to_dev(m.astype(np.uint8), np.uint8)
    xd = _lib.to_dev(x)
    yd = _lib.to_dev(y)
    status = _lib.empty((nx, ny), "i32") if return_status else None
    out = flowmap_grid_2d_device(funcptr, t0, T, xd, yd, params, rtol, atol, maskd, max_steps,
                                 status=status)
    res = _lib.to_host(out)
    if return_status:
        return res, _lib.to_host(status)
    return res


def flowmap(funcptr, t0, T, pts, params, method="dop853", rtol=1e-6, atol=1e-8, mask=None, *,
            max_steps=DEFAULT_MAX_STEPS, return_status=False):
    """Flow map of a point list pts (npts, 2) (integration.py:7-61)."""
    _check_flow(funcptr)
    _check_method(method)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] != 2:
        raise ValueError("the GPU flow map integrates 2-D points: pts must be (npts, 2)")
    n = pts.shape[0]
    if n == 0:
        out = np.zeros((0, 2))
        return (out, np.zeros(0, np.int32)) if return_status else out
    p = _params(params)
    maskd = None
    if mask is not None:
        m = np.asarray(mask).ravel()
        if m.size != n:
            raise ValueError("mask must have one entry per point")
        maskd = _lib.to_dev(m.astype(np.uint8), np.uint8)
    pd = _lib.to_dev(pts)
    out = _lib.empty((n, 2))
    status = _lib.empty((n,), "i32") if return_status else None
    _lib.check(_lib.load().nbx_flowmap_pts(
        funcptr.handle, float(t0), float(T), _lib.ptr(pd), n, p.ctypes.data, p.size,
        float(rtol), float(atol), _lib.ptr(maskd), int(max_steps), _lib.NBX_FP64, _lib.ptr(out),
        _lib.ptr(status), None, _lib.stream_ptr()))
    res = _lib.to_host(out)
    if return_status:
        return res, _lib.to_host(status)
    return res
