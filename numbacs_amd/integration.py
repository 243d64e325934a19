"""Particle flow maps (mirrors numbacs.integration).

Reference: flowmap_grid_2D (integration.py:123-182), flowmap
(integration.py:7-61), flowmap_aux_grid_2D (integration.py:249-464) and the
flow-map composition functions (integration.py:609-737).  Same signatures,
defaults, 'ij' layouts and float64 outputs; the per-particle
numbalsoda.dop853 loop becomes one GPU launch of one DOP853 solve per lane
group, and composition runs as one GPU launch over device-resident maps.

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
    precision: 0 (NBX_FP64) or 1 (NBX_FP32_DISP: float stages, step
    displacements and spline; double position, time and step control)."""
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


def _mask_dev(mask, nx, ny):
    if mask is None:
        return None
    m = np.asarray(mask)
    if m.shape != (nx, ny):
        raise ValueError(f"mask must have shape {(nx, ny)}")
    return _lib.to_dev(m.astype(np.uint8), np.uint8)


def flowmap_aux_grid_2d_device(funcptr, t0, T, xd, yd, params, h=1e-5, eig_main=True,
                               compute_edge=True, rtol=1e-6, atol=1e-8, maskd=None,
                               max_steps=DEFAULT_MAX_STEPS, precision=_lib.NBX_FP64, out=None,
                               status=None):
    """Device-tensor form of flowmap_aux_grid_2D: returns (nx, ny, n_aux, 2)."""
    _check_flow(funcptr)
    p = _params(params)
    nx, ny = int(xd.shape[0]), int(yd.shape[0])
    naux = 5 if eig_main else 4
    if out is None:
        out = _lib.empty((nx, ny, naux, 2))
    _lib.check(_lib.load().nbx_flowmap_aux_grid_2d(
        funcptr.handle, float(t0), float(T), _lib.ptr(xd), nx, _lib.ptr(yd), ny, float(h),
        int(bool(eig_main)), int(bool(compute_edge)), p.ctypes.data, p.size, float(rtol),
        float(atol), _lib.ptr(maskd), int(max_steps), int(precision), _lib.ptr(out),
        _lib.ptr(status), None, _lib.stream_ptr()))
    return out


def flowmap_aux_grid_2D(funcptr, t0, T, x, y, params, h=1e-5, eig_main=True, compute_edge=True,
                        method="dop853", rtol=1e-6, atol=1e-8, mask=None, *,
                        max_steps=DEFAULT_MAX_STEPS):
    """Flow map over the auxiliary grid (x[i], y[j]) + {(h,0), (-h,0), (0,h),
    (0,-h)[, (0,0) with eig_main]}: returns (nx, ny, n_aux, 2) float64.
    compute_edge=True integrates edge points too (only their centre particle
    when eig_main); masked and left-out particles stay 0."""
    _check_flow(funcptr)
    _check_method(method)
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    nx, ny = len(x), len(y)
    naux = 5 if eig_main else 4
    if nx * ny == 0:
        return np.zeros((nx, ny, naux, 2))
    maskd = _mask_dev(mask, nx, ny)
    out = flowmap_aux_grid_2d_device(funcptr, t0, T, _lib.to_dev(x), _lib.to_dev(y), params, h,
                                     eig_main, compute_edge, rtol, atol, maskd, max_steps)
    return _lib.to_host(out)


def _grid_box(grid):
    (x0, x1, nx), (y0, y1, ny) = grid[0], grid[1]
    return float(x0), float(x1), int(nx), float(y0), float(y1), int(ny)


def flowmap_composition_device(maps, grid, nT, first=0, out=None):
    """Compose device flow maps: maps is a (nslots, nx, ny, 2) float64 tensor
    whose chain starts at slot `first` (map k in slot (first+k) % nslots).
    Returns the (nx, ny, 2) tensor."""
    x0, x1, nx, y0, y1, ny = _grid_box(grid)
    if maps.ndim != 4 or tuple(maps.shape[1:]) != (nx, ny, 2) or not maps.is_contiguous():
        raise ValueError(f"flowmaps must be a contiguous (nT, {nx}, {ny}, 2) array")
    if out is None:
        out = _lib.empty((nx, ny, 2))
    _lib.check(_lib.load().nbx_flowmap_compose(_lib.ptr(maps), int(maps.shape[0]), int(first),
                                                int(nT), x0, x1, nx, y0, y1, ny, _lib.ptr(out),
                                                _lib.stream_ptr()))
    return out


def flowmap_composition(flowmaps, grid, nT):
    """Composed flow map from the intermediate flow maps (integration.py:
    609-644): flowmaps (nT, nx, ny, 2) on grid ((x0, x1, nx), (y0, y1, ny)),
    bilinear interpolation, 0 outside the grid box."""
    maps = _lib.to_dev(np.ascontiguousarray(flowmaps, dtype=np.float64))
    return _lib.to_host(flowmap_composition_device(maps, grid, nT))


def _window_times(t0, h, nT):
    # t0 accumulated the way the reference's loop does it (t0 += h)
    ts = []
    for _ in range(nT):
        ts.append(t0)
        t0 += h
    return ts


class FlowmapComposer:
    """Device-resident sliding-window composition: nT flow maps of duration h
    kept in a ring of slots, so each new window (t0 -> t0 + nT h) costs one
    flow map of duration h plus one composition launch.

    initial(t0) integrates the nT maps of the first window in one batched
    work-queue launch (flowmap_composition_initial, integration.py:647-692);
    step(t0) integrates the map starting at t0 into the oldest slot
    (flowmap_composition_step, integration.py:695-737); compose() returns
    the composed (nx, ny, 2) device tensor."""

    def __init__(self, funcptr, x, y, h, nT, params, rtol=1e-6, atol=1e-8, mask=None,
                 precision=_lib.NBX_FP64, max_steps=DEFAULT_MAX_STEPS):
        _check_flow(funcptr)
        self.flow = funcptr
        self.x = _lib.to_dev(x)
        self.y = _lib.to_dev(y)
        nx, ny = int(self.x.shape[0]), int(self.y.shape[0])
        xh, yh = _lib.to_host(self.x), _lib.to_host(self.y)
        self.grid = ((xh[0], xh[-1], nx), (yh[0], yh[-1], ny))
        self.h, self.nT = float(h), int(nT)
        if self.nT < 1:
            raise ValueError("nT must be >= 1")
        self.params = _params(params)
        self.rtol, self.atol, self.max_steps, self.precision = rtol, atol, max_steps, precision
        self.mask = mask
        self.maskd = _mask_dev(mask, nx, ny)
        self.ring = _lib.empty((self.nT, nx, ny, 2))
        self.first = 0

    def initial(self, t0):
        from .batch import FTLEBatch

        reqs = [dict(t0=t, T=self.h, x=self.x, y=self.y, mask=self.mask, flowmap=self.ring[k])
                for k, t in enumerate(_window_times(t0, self.h, self.nT))]
        prec = "fp32" if self.precision == _lib.NBX_FP32_DISP else "fp64"
        b = FTLEBatch(self.flow, reqs, self.params, self.rtol, self.atol, precision=prec,
                      ftle=False, max_steps=self.max_steps)
        b.run()
        self._batch = b  # holds the request table until the launch has run
        self.first = 0
        return self

    def step(self, t0):
        slot = self.ring[self.first]
        flowmap_grid_2d_device(self.flow, t0, self.h, self.x, self.y, self.params, self.rtol,
                               self.atol, self.maskd, self.max_steps, self.precision, out=slot)
        self.first = (self.first + 1) % self.nT
        return self

    def compose(self, out=None):
        return flowmap_composition_device(self.ring, self.grid, self.nT, self.first, out)

    def flowmaps(self):
        """The window's maps in chain order, (nT, nx, ny, 2) device tensor."""
        t = _lib.torch()
        return t.roll(self.ring, -self.first, dims=0).contiguous()


def flowmap_composition_initial(funcptr, t0, T, h, x, y, grid, params, **kwargs):
    """First window of the composition method (integration.py:647-692):
    returns (flowmap0 (nx, ny, 2), flowmaps (nT, nx, ny, 2), nT) with
    nT = |round(T/h)| and flowmaps[k] the flow map from t0 + k h over h.
    kwargs: method, rtol, atol, mask as for flowmap_grid_2D."""
    _check_method(kwargs.pop("method", "dop853"))
    nT = abs(round(T / h))
    c = FlowmapComposer(funcptr, x, y, h, nT, params, **kwargs)
    if tuple(int(v) for v in (c.grid[0][2], c.grid[1][2])) != (int(grid[0][2]), int(grid[1][2])):
        raise ValueError("grid must have the particle grid's point counts")
    c.initial(t0)
    flowmap0 = flowmap_composition_device(c.ring, grid, nT)
    return _lib.to_host(flowmap0), _lib.to_host(c.ring), nT


def flowmap_composition_step(flowmaps, funcptr, t0, h, nT, x, y, grid, params, **kwargs):
    """Next window of the composition method (integration.py:695-737): shifts
    flowmaps (nT, nx, ny, 2) down by one in place, integrates the new last
    map from t0 over h and returns (flowmap_k, flowmaps)."""
    _check_method(kwargs.pop("method", "dop853"))
    mask = kwargs.pop("mask", None)
    flowmaps[:-1] = flowmaps[1:]
    flowmaps[-1] = flowmap_grid_2D(funcptr, t0, h, x, y, params, mask=mask, **kwargs)
    maps = _lib.to_dev(np.ascontiguousarray(flowmaps, dtype=np.float64))
    return _lib.to_host(flowmap_composition_device(maps, grid, nT)), flowmaps
