"""CPU oracle for the FTLE hot path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``libnbx_oracle.so`` (built from ``nbx_oracle.c`` by the
Makefile next to it), a plain-C, IEEE-ordered restatement of the reference
algorithm (integration.py:123-182, flows.py:9-46,121-258,1104-1245,
diagnostics.py:21-65,200-269, extraction/ridges.py:9-229).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker / the timed CPU
baseline; the ``numbacs_amd`` product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnbx_oracle.so")

KIND = {"double_gyre": 0, "bickley_jet": 1, "spline": 2, "linear": 3}
EXTRAP = {"constant": 0, "linear": 1, "nearest": 2}

_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_l = ctypes.c_long
_p = ctypes.c_void_p


def build():
    """Compile the oracle library with its Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_flowmap_grid_2d.argtypes = [_i, _p, _p, _p, _p, _i, _i, _d, _d, _d, _p, _i, _p, _i,
                                          _d, _d, _p, _l, _i, _i, _p, _p, _p, _i]
        L.orc_flowmap_grid_2d.restype = None
        L.orc_flowmap_pts.argtypes = [_i, _p, _p, _p, _p, _i, _i, _d, _d, _d, _p, _l, _d, _d,
                                      _l, _p, _p, _i]
        L.orc_flowmap_pts.restype = None
        L.orc_rhs_eval.argtypes = [_i, _p, _p, _p, _p, _i, _i, _d, _p, _p, _l, _p]
        L.orc_rhs_eval.restype = None
        L.orc_prefilter3.argtypes = [_p, _i, _i, _i, _p]
        L.orc_prefilter3.restype = None
        L.orc_prefilter2.argtypes = [_p, _i, _i, _p]
        L.orc_prefilter2.restype = None
        L.orc_spline3_eval_pts.argtypes = [_p, _p, _i, _p, _l, _p]
        L.orc_spline3_eval_pts.restype = None
        L.orc_linear3_eval_pts.argtypes = [_p, _p, _i, _p, _l, _p]
        L.orc_linear3_eval_pts.restype = None
        L.orc_ftle_grid_2d.argtypes = [_p, _i, _i, _d, _d, _d, _p, _p]
        L.orc_ftle_grid_2d.restype = None
        L.orc_C_eig_2d.argtypes = [_p, _i, _i, _d, _d, _p, _p, _p]
        L.orc_C_eig_2d.restype = None
        L.orc_percentile.argtypes = [_p, _l, _d]
        L.orc_percentile.restype = _d
        L.orc_ridge_predicate.argtypes = [_p, _p, _p, _i, _p, _i, _d, _d, _p, _p]
        L.orc_ridge_predicate.restype = _l
        L.orc_label8.argtypes = [_p, _i, _i, _p]
        L.orc_label8.restype = ctypes.c_int32
        L.orc_flowmap_aux_grid_2d.argtypes = [_i, _p, _p, _p, _p, _i, _i, _d, _d, _d, _p, _i,
                                              _p, _i, _d, _i, _i, _d, _d, _p, _l, _p, _i]
        L.orc_flowmap_aux_grid_2d.restype = None
        L.orc_C_tensor_aux_2d.argtypes = [_p, _i, _i, _i, _d, _p, _p]
        L.orc_C_tensor_aux_2d.restype = None
        L.orc_C_eig_aux_2d.argtypes = [_p, _i, _i, _i, _d, _d, _d, _i, _p, _p, _p]
        L.orc_C_eig_aux_2d.restype = None
        L.orc_flowmap_composition.argtypes = [_p, _i, _i, _i, _p, _p]
        L.orc_flowmap_composition.restype = None
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = _i
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Spline:
    """Tricubic spline flow description for the oracle (mirrors get_flow_2D)."""

    def __init__(self, grid, Cu, Cv, spherical=0, extrap_mode="constant", r=6371.0):
        self.grid9 = np.array([v for ax in grid for v in ax], dtype=np.float64)
        self.Cu = _f64(Cu)
        self.Cv = _f64(Cv)
        self.spherical = int(spherical)
        self.extrap = EXTRAP[extrap_mode]
        self.r = float(r)


class Linear(Spline):
    """Trilinear flow description for the oracle (mirrors get_flow_linear_2D,
    flows.py:418-506): U, V are the raw (nt, nx, ny) samples."""


def _flow_args(kind, params, spline):
    params = _f64(params)
    if kind in ("spline", "linear"):
        s = spline
        return (KIND[kind], params, s.grid9, s.Cu, s.Cv, s.spherical, s.extrap, s.r)
    return (KIND[kind], params, None, None, None, 0, 0, 0.0)


def flowmap_grid_2d(kind, params, t0, T, x, y, rtol=1e-6, atol=1e-8, mask=None, spline=None,
                    nmax=100000, rows=None, nthreads=0):
    """flowmap_grid_2D restated (integration.py:123-182).

    Returns (flowmap (nx,ny,2), status (nx,ny) Hairer idid, nsteps (nx,ny)).
    ``rows=(lo, hi)`` integrates only those x-rows (others stay 0)."""
    x = _f64(x)
    y = _f64(y)
    nx, ny = len(x), len(y)
    out = np.zeros((nx, ny, 2))
    status = np.zeros((nx, ny), np.int32)
    nsteps = np.zeros((nx, ny), np.int32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lo, hi = (0, nx) if rows is None else rows
    k, p, g, cu, cv, sph, ex, r = _flow_args(kind, params, spline)
    lib().orc_flowmap_grid_2d(k, _ptr(p), _ptr(g), _ptr(cu), _ptr(cv), sph, ex, r, float(t0),
                              float(T), _ptr(x), nx, _ptr(y), ny, float(rtol), float(atol),
                              _ptr(m), int(nmax), int(lo), int(hi), _ptr(out), _ptr(status),
                              _ptr(nsteps), int(nthreads))
    return out, status, nsteps


def flowmap_pts(kind, params, t0, T, pts, rtol=1e-6, atol=1e-8, spline=None, nmax=100000,
                nthreads=0):
    """flowmap restated (integration.py:7-61). Returns (out (n,2), status (n,))."""
    pts = _f64(pts)
    n = pts.shape[0]
    out = np.zeros((n, 2))
    status = np.zeros(n, np.int32)
    k, p, g, cu, cv, sph, ex, r = _flow_args(kind, params, spline)
    lib().orc_flowmap_pts(k, _ptr(p), _ptr(g), _ptr(cu), _ptr(cv), sph, ex, r, float(t0),
                          float(T), _ptr(pts), n, float(rtol), float(atol), int(nmax), _ptr(out),
                          _ptr(status), int(nthreads))
    return out, status


def flowmap_aux_grid_2d(kind, params, t0, T, x, y, h=1e-5, eig_main=True, compute_edge=True,
                        rtol=1e-6, atol=1e-8, mask=None, spline=None, nmax=100000, nthreads=0):
    """flowmap_aux_grid_2D restated (integration.py:249-464): (nx, ny, n_aux, 2)."""
    x = _f64(x)
    y = _f64(y)
    nx, ny = len(x), len(y)
    naux = 5 if eig_main else 4
    out = np.zeros((nx, ny, naux, 2))
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    k, p, g, cu, cv, sph, ex, r = _flow_args(kind, params, spline)
    lib().orc_flowmap_aux_grid_2d(k, _ptr(p), _ptr(g), _ptr(cu), _ptr(cv), sph, ex, r,
                                  float(t0), float(T), _ptr(x), nx, _ptr(y), ny, float(h),
                                  int(bool(eig_main)), int(bool(compute_edge)), float(rtol),
                                  float(atol), _ptr(m), int(nmax), _ptr(out), int(nthreads))
    return out


def C_tensor_aux_2d(fm_aux, h=1e-5, mask=None):
    """C_tensor_2D restated (diagnostics.py:67-110): (nx, ny, 3)."""
    F = _f64(fm_aux)
    nx, ny, naux = F.shape[:3]
    C = np.zeros((nx, ny, 3))
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_C_tensor_aux_2d(_ptr(F), nx, ny, naux, float(h), _ptr(m), _ptr(C))
    return C


def C_eig_aux_2d(fm_aux, dx, dy, h=1e-5, eig_main=True, mask=None):
    """C_eig_aux_2D restated (diagnostics.py:113-197)."""
    F = _f64(fm_aux)
    nx, ny, naux = F.shape[:3]
    ev = np.zeros((nx, ny, 2))
    V = np.zeros((nx, ny, 2, 2))
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_C_eig_aux_2d(_ptr(F), nx, ny, naux, float(dx), float(dy), float(h),
                           int(bool(eig_main)), _ptr(m), _ptr(ev), _ptr(V))
    return ev, V


def flowmap_composition(flowmaps, grid, nT):
    """flowmap_composition restated (integration.py:609-644); grid =
    ((x0, x1, nx), (y0, y1, ny))."""
    F = _f64(flowmaps)
    nx, ny = int(grid[0][2]), int(grid[1][2])
    g4 = np.array([grid[0][0], grid[0][1], grid[1][0], grid[1][1]], dtype=np.float64)
    out = np.zeros((nx, ny, 2))
    lib().orc_flowmap_composition(_ptr(F), int(nT), nx, ny, _ptr(g4), _ptr(out))
    return out


def rhs_eval(kind, params, t, pts, spline=None):
    """Right-hand side dy = f(t, y) at points (n,2) and times (n,)."""
    t = _f64(t)
    pts = _f64(pts)
    n = len(t)
    out = np.zeros((n, 2))
    k, p, g, cu, cv, sph, ex, r = _flow_args(kind, params, spline)
    lib().orc_rhs_eval(k, _ptr(p), _ptr(g), _ptr(cu), _ptr(cv), sph, ex, r, _ptr(t), _ptr(pts),
                       n, _ptr(out))
    return out


def prefilter3(U):
    U = _f64(U)
    nt, nx, ny = U.shape
    C = np.zeros((nt + 2, nx + 2, ny + 2))
    lib().orc_prefilter3(_ptr(U), nt, nx, ny, _ptr(C))
    return C


def prefilter2(U):
    U = _f64(U)
    nx, ny = U.shape
    C = np.zeros((nx + 2, ny + 2))
    lib().orc_prefilter2(_ptr(U), nx, ny, _ptr(C))
    return C


def spline3_eval(grid, C, pts, extrap_mode="constant"):
    grid9 = np.array([v for ax in grid for v in ax], dtype=np.float64)
    C = _f64(C)
    pts = _f64(pts)
    n = pts.shape[0]
    out = np.zeros(n)
    lib().orc_spline3_eval_pts(_ptr(grid9), _ptr(C), EXTRAP[extrap_mode], _ptr(pts), n, _ptr(out))
    return out


def linear3_eval(grid, V, pts, extrap_mode="constant"):
    """Trilinear interpolant of V (nt, nx, ny) at pts (n, 3) = (t, x, y)."""
    grid9 = np.array([v for ax in grid for v in ax], dtype=np.float64)
    V = _f64(V)
    pts = _f64(pts)
    out = np.zeros(len(pts))
    lib().orc_linear3_eval_pts(_ptr(grid9), _ptr(V), EXTRAP[extrap_mode], _ptr(pts), len(pts),
                               _ptr(out))
    return out


def ftle_grid_2d(fm, T, dx, dy, mask=None):
    fm = _f64(fm)
    nx, ny = fm.shape[:2]
    out = np.zeros((nx, ny))
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_ftle_grid_2d(_ptr(fm), nx, ny, float(T), float(dx), float(dy), _ptr(m), _ptr(out))
    return out


def C_eig_2d(fm, dx, dy, mask=None):
    fm = _f64(fm)
    nx, ny = fm.shape[:2]
    ev = np.zeros((nx, ny, 2))
    V = np.zeros((nx, ny, 2, 2))
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_C_eig_2d(_ptr(fm), nx, ny, float(dx), float(dy), _ptr(m), _ptr(ev), _ptr(V))
    return ev, V


def percentile(f, q):
    f = _f64(f).ravel()
    return lib().orc_percentile(_ptr(f), f.size, float(q))


def ridge_predicate(f, evec, x, y, sdd_thresh=0.0, percentile_=0):
    """_ftle_ridges restated (ridges.py:79-154): (r_pts (nx*ny,2), ridge_bool (nx,ny))."""
    f = _f64(f)
    evec = _f64(evec)
    x = _f64(x)
    y = _f64(y)
    nx, ny = f.shape
    f_min = 0.0 if percentile_ == 0 else percentile(f, percentile_)
    rb = np.zeros((nx, ny), np.uint8)
    pts = np.zeros((nx * ny, 2))
    lib().orc_ridge_predicate(_ptr(f), _ptr(evec), _ptr(x), nx, _ptr(y), ny, float(sdd_thresh),
                              float(f_min), _ptr(rb), _ptr(pts))
    return pts, rb.astype(bool)


def ridge_pts(f, evec, x, y, sdd_thresh=0.0, percentile_=0):
    """ftle_ridge_pts restated (ridges.py:9-76)."""
    pts, rb = ridge_predicate(f, evec, x, y, sdd_thresh, percentile_)
    return pts[rb.ravel()]


def label8(b):
    b = np.ascontiguousarray(b, dtype=np.uint8)
    nx, ny = b.shape
    lab = np.zeros((nx, ny), np.int32)
    n = lib().orc_label8(_ptr(b), nx, ny, _ptr(lab))
    return lab, int(n)


def ridges(f, evec, x, y, sdd_thresh=0.0, percentile_=0, min_ridge_pts=3):
    """ftle_ridges restated (ridges.py:186-229)."""
    pts, rb = ridge_predicate(f, evec, x, y, sdd_thresh, percentile_)
    lab, n = label8(rb)
    flat = lab.ravel()
    out = []
    for i in range(1, n + 1):
        inds = np.nonzero(flat == i)[0]
        if len(inds) >= min_ridge_pts:
            out.append(pts[inds, :])
    return out


def num_threads():
    return lib().orc_num_threads()
