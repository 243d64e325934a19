"""Velocity fields for the flow-map path (mirrors numbacs.flows).

Reference: get_predefined_flow (flows.py:1104-1295), get_interp_arrays_2D
(flows.py:9-46), get_flow_2D (flows.py:121-258) and get_flow_linear_2D
(flows.py:418-506).  The reference returns
the address of a numba cfunc as ``funcptr``; here ``funcptr`` is a
:class:`FlowHandle` -- an ``int`` (the address of an opaque ``nbx_flow``)
that owns the device-side description of the flow.  A raw cfunc address from
a user-defined numba flow cannot run on the GPU and is rejected by the
integrators with ``TypeError``.
"""
import ctypes
from math import pi, sqrt

import numpy as np

from . import _lib


class FlowHandle(int):
    """``funcptr`` stand-in: int value = address of the native nbx_flow."""

    def __new__(cls, addr, kind, keep=None, grid=None):
        obj = int.__new__(cls, addr)
        obj.kind = kind
        obj._keep = keep
        obj.grid = grid
        obj._alive = True
        return obj

    @property
    def handle(self):
        return ctypes.c_void_p(int(self))

    def close(self):
        if getattr(self, "_alive", False):
            self._alive = False
            try:
                _lib.load().nbx_flow_destroy(ctypes.c_void_p(int(self)))
            except Exception:
                pass

    def __del__(self):
        self.close()

    def __reduce__(self):
        raise TypeError("FlowHandle refers to device memory and cannot be pickled")


def _analytic(kind):
    out = ctypes.c_void_p()
    _lib.check(_lib.load().nbx_flow_create_analytic(kind, ctypes.byref(out)))
    return FlowHandle(out.value, kind)


def _ret(funcptr, default_params, domain, p_str, return_default_params, return_domain,
         parameter_description):
    """Return-tuple arity matrix of flows.py:1279-1295."""
    out = [funcptr]
    if return_default_params:
        out.append(default_params)
    if return_domain:
        out.append(domain)
    if parameter_description:
        out.append(p_str)
    return out[0] if len(out) == 1 else tuple(out)


def get_predefined_flow(flow_str, int_direction=1.0, return_default_params=True,
                        return_domain=True, parameter_description=False):
    """Predefined analytic flows (flows.py:1104-1295): 'double_gyre' and
    'bickley_jet'.  As in the reference, the Bickley jet's default parameters
    carry int_direction = 1.0 whatever is passed (flows.py:1219); set
    params[0] = -1 for backward time."""
    if flow_str == "double_gyre":
        funcptr = _analytic(_lib.NBX_FLOW_DOUBLE_GYRE)
        params = np.array([int_direction, 0.1, 0.25, 0.0, 0.2 * pi, 0.0])
        domain = ((0.0, 2.0), (0.0, 1.0))
        p_str = ("p[0] = int_direction, p[1] = A, p[2] = eps, p[3] = alpha, "
                 + "p[4] = omega, p[5] = psi")
    elif flow_str == "bickley_jet":
        funcptr = _analytic(_lib.NBX_FLOW_BICKLEY_JET)
        r_e = 6371.0e-3
        U0 = 86400 * 62.66e-6
        c2 = 0.205 * U0
        c3 = 0.461 * U0
        c1 = c3 + (sqrt(5) - 1) * (c2 - c3)
        params = np.array([1.0, U0, 1770.0e-3, 0.0075, 0.15, 0.3, 2.0 / r_e, 4.0 / r_e, 6.0 / r_e,
                           c1, c2, c3])
        domain = ((0.0, r_e * pi), (-3.0, 3.0))
        p_str = ("p[0] = int_direction, p[1] = U0, p[2] = L, p[3] = A1, p[4] = A2, "
                 + "p[5] = A3, p[6] = k1, p[7] = k2, p[8] = k3, p[9] = c1, p[10] = c2, "
                 + "p[11] = c3, units: time - days, length - Mm")
    elif flow_str == "abc":
        raise NotImplementedError("the 3-D 'abc' flow is not part of the 2-D FTLE path")
    else:
        raise ValueError(f"unknown predefined flow {flow_str!r}")
    return _ret(funcptr, params, domain, p_str, return_default_params, return_domain,
                parameter_description)


def _grid_tuple(tvals, xvals, yvals, shape):
    nt, nx, ny = shape
    return ((float(tvals[0]), float(tvals[-1]), int(nt)),
            (float(xvals[0]), float(xvals[-1]), int(nx)),
            (float(yvals[0]), float(yvals[-1]), int(ny)))


def prefilter3_device(U):
    """Natural cubic B-spline coefficients of a (nt, nx, ny) device or host
    array, returned as a (nt+2, nx+2, ny+2) float64 device tensor."""
    Ud = _lib.to_dev(U)
    nt, nx, ny = Ud.shape
    C = _lib.empty((nt + 2, nx + 2, ny + 2))
    _lib.check(_lib.load().nbx_prefilter3(_lib.ptr(Ud), nt, nx, ny, _lib.ptr(C),
                                          _lib.stream_ptr()))
    return C


def get_interp_arrays_2D(tvals, xvals, yvals, U, V):
    """Grid tuple and cubic-spline coefficient arrays (flows.py:9-46).

    U, V: (nt, nx, ny) velocity samples on ascending uniform axes.  Returns
    (grid_vel, C_eval_u, C_eval_v) with coefficient arrays of shape
    (nt+2, nx+2, ny+2) (numpy, float64), computed on the GPU."""
    U = np.asarray(U)
    V = np.asarray(V)
    if U.ndim != 3 or U.shape != V.shape:
        raise ValueError("U and V must be (nt, nx, ny) arrays of the same shape")
    grid = _grid_tuple(tvals, xvals, yvals, U.shape)
    Cu = prefilter3_device(U)
    Cv = prefilter3_device(V)
    return grid, _lib.to_host(Cu), _lib.to_host(Cv)


def get_flow_2D(grid_vel, C_eval_u, C_eval_v, spherical=0, extrap_mode="constant", r=6371.0):
    """Tricubic-spline flow over gridded velocity (flows.py:121-258).

    spherical: 0 planar; 1 lon in [-180,180) / 2 lon in [0,360) in degrees
    with velocities in km/h converted by radius r.  extrap_mode: 'constant'
    (zero velocity outside the grid box), 'linear' (first-order continuation
    from the box) or 'nearest' (clamp the query into the box)."""
    if spherical not in (0, 1, 2):
        raise ValueError("spherical must be 0, 1 or 2")
    if extrap_mode not in _lib.EXTRAP:
        raise ValueError(f"unknown extrap_mode {extrap_mode!r}")
    grid9 = np.array([float(v) for ax in grid_vel for v in ax], dtype=np.float64)
    if grid9.size != 9:
        raise ValueError("grid_vel must be ((t0,t1,nt),(x0,x1,nx),(y0,y1,ny))")
    shape = tuple(int(grid9[3 * d + 2]) + 2 for d in range(3))
    Cu = _lib.to_dev(C_eval_u)
    Cv = _lib.to_dev(C_eval_v)
    if tuple(Cu.shape) != shape or tuple(Cv.shape) != shape:
        raise ValueError(f"coefficient arrays must have shape {shape}")
    out = ctypes.c_void_p()
    _lib.check(_lib.load().nbx_flow_create_spline3(
        grid9.ctypes.data, _lib.ptr(Cu), _lib.ptr(Cv), int(spherical), _lib.EXTRAP[extrap_mode],
        float(r), ctypes.byref(out), _lib.stream_ptr()))
    return FlowHandle(out.value, _lib.NBX_FLOW_SPLINE3, grid=grid_vel)


def get_flow_linear_2D(grid_vel, U, V, spherical=0, extrap_mode="constant", r=6371.0):
    """Trilinear flow over gridded velocity samples (flows.py:418-506).

    U, V: (nt, nx, ny) samples on grid_vel ((t0,t1,nt),(x0,x1,nx),(y0,y1,ny)).
    spherical and r as for get_flow_2D.  extrap_mode: 'constant' (zero
    outside the grid box, the interpolation package's CONSTANT), 'linear'
    (the boundary cell continued) or 'nearest' (clamp into the box)."""
    if spherical not in (0, 1, 2):
        raise ValueError("spherical must be 0, 1 or 2")
    if extrap_mode not in _lib.EXTRAP:
        raise ValueError(f"unknown extrap_mode {extrap_mode!r}")
    grid9 = np.array([float(v) for ax in grid_vel for v in ax], dtype=np.float64)
    if grid9.size != 9:
        raise ValueError("grid_vel must be ((t0,t1,nt),(x0,x1,nx),(y0,y1,ny))")
    shape = tuple(int(grid9[3 * d + 2]) for d in range(3))
    Ud = _lib.to_dev(U)
    Vd = _lib.to_dev(V)
    if tuple(Ud.shape) != shape or tuple(Vd.shape) != shape:
        raise ValueError(f"velocity arrays must have shape {shape}")
    out = ctypes.c_void_p()
    _lib.check(_lib.load().nbx_flow_create_linear3(
        grid9.ctypes.data, _lib.ptr(Ud), _lib.ptr(Vd), int(spherical), _lib.EXTRAP[extrap_mode],
        float(r), ctypes.byref(out), _lib.stream_ptr()))
    return FlowHandle(out.value, _lib.NBX_FLOW_LINEAR3, grid=grid_vel)


def ingest_xy(U_xy, scale=1.0):
    """Velocity samples stored 'xy' (nt, ny, nx) -> 'ij' (nt, nx, ny) times
    scale, on the GPU (the np.moveaxis(u, 1, 2) * 3.6 of the reference's
    plot_merra_ftle.py:37-41; scale=3.6 converts m/s to km/h).  Returns a
    device tensor, ready for prefilter3_device / get_flow_linear_2D."""
    Ud = _lib.to_dev(U_xy)
    if Ud.ndim != 3:
        raise ValueError("U_xy must be a (nt, ny, nx) array")
    nt, ny, nx = (int(v) for v in Ud.shape)
    out = _lib.empty((nt, nx, ny))
    _lib.check(_lib.load().nbx_ingest_xy(_lib.ptr(Ud), nt, ny, nx, float(scale), _lib.ptr(out),
                                         _lib.stream_ptr()))
    return out


def flow_velocity(funcptr, t, pts, params):
    """Right-hand side p0 * v(p0 * t, y) of a flow at points pts (n, 2) and
    times t (n,), evaluated on the GPU (the cfunc's output, flows.py:1146)."""
    if not isinstance(funcptr, FlowHandle):
        raise TypeError("funcptr must come from numbacs_amd.flows")
    t = np.ascontiguousarray(t, dtype=np.float64).ravel()
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    params = np.ascontiguousarray(params, dtype=np.float64)
    n = len(t)
    if pts.shape[0] != n:
        raise ValueError("t and pts must have the same length")
    td = _lib.to_dev(t)
    pd = _lib.to_dev(pts)
    out = _lib.empty((n, 2))
    _lib.check(_lib.load().nbx_flow_eval(funcptr.handle, params.ctypes.data, len(params),
                                         _lib.ptr(td), _lib.ptr(pd), n, _lib.ptr(out),
                                         _lib.stream_ptr()))
    return _lib.to_host(out)
