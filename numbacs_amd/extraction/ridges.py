"""FTLE ridge extraction (mirrors numbacs.extraction.ridges).

Reference: ftle_ridge_pts (ridges.py:9-76), ftle_ridges
(ridges.py:186-229) and ftle_ordered_ridges (ridges.py:720-1054).  The percentile threshold, ridge predicate, compaction,
8-connected labelling and grouping all run on the GPU; the host only splits
the grouped point buffer into the list the reference returns.
"""
import ctypes

import numpy as np

from .. import _lib


def _inputs(f, eigvec_max, x, y):
    f = np.asarray(f, dtype=np.float64) if not _is_tensor(f) else f
    nx, ny = int(f.shape[0]), int(f.shape[1])
    if tuple(eigvec_max.shape) != (nx, ny, 2):
        raise ValueError("eigvec_max must have shape (nx, ny, 2)")
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    if len(x) != nx or len(y) != ny:
        raise ValueError("x and y must match the grid of f")
    # the reference's spacing (ridges.py:45-46)
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    return (_lib.to_dev(f), _lib.to_dev(eigvec_max), _lib.to_dev(x), _lib.to_dev(y), nx, ny,
            float(dx), float(dy))


def _is_tensor(a):
    try:
        import torch

        return isinstance(a, torch.Tensor)
    except ImportError:
        return False


def ftle_ridge_pts_device(fd, evd, xd, yd, nx, ny, dx, dy, sdd_thresh=0.0, percentile=0):
    """Device form: returns (pts tensor (capacity nx*ny, 2), count)."""
    pts = _lib.empty((nx * ny, 2))
    n = ctypes.c_int64()
    _lib.check(_lib.load().nbx_ridge_pts(_lib.ptr(fd), _lib.ptr(evd), _lib.ptr(xd), nx,
                                         _lib.ptr(yd), ny, dx, dy, float(sdd_thresh),
                                         float(percentile), _lib.ptr(pts), ctypes.byref(n),
                                         _lib.stream_ptr()))
    return pts, n.value


def ftle_ridges_device(fd, evd, xd, yd, nx, ny, dx, dy, sdd_thresh=0.0, percentile=0,
                       min_ridge_pts=3):
    """Device form: returns (pts tensor, offsets tensor, nridges, npts)."""
    pts = _lib.empty((max(nx * ny, 1), 2))
    offsets = _lib.empty((nx * ny + 1,), "i64")
    nr = ctypes.c_int64()
    npts = ctypes.c_int64()
    _lib.check(_lib.load().nbx_ridges(_lib.ptr(fd), _lib.ptr(evd), _lib.ptr(xd), nx, _lib.ptr(yd),
                                      ny, dx, dy, float(sdd_thresh), float(percentile),
                                      int(min_ridge_pts), _lib.ptr(pts), _lib.ptr(offsets),
                                      ctypes.byref(nr), ctypes.byref(npts), _lib.stream_ptr()))
    return pts, offsets, nr.value, npts.value


def ftle_ridges_device_async(fd, evd, xd, yd, nx, ny, dx, dy, sdd_thresh=0.0, percentile=0,
                             min_ridge_pts=3, pts=None, offsets=None, counts=None):
    """Device form that does not wait on the stream: returns (pts, offsets,
    counts), counts a device int64 tensor (ridges, points kept, ridge points
    before the length filter) valid once the current stream's work is done.
    pts, offsets and counts may be given (capacities nx*ny x 2, nx*ny + 1
    and 3) so that a caller can reuse them from call to call."""
    n = max(nx * ny, 1)
    pts = _lib.empty((n, 2)) if pts is None else pts
    offsets = _lib.empty((nx * ny + 1,), "i64") if offsets is None else offsets
    counts = _lib.empty((3,), "i64") if counts is None else counts
    if pts.numel() < 2 * n or offsets.numel() < nx * ny + 1 or counts.numel() < 3:
        raise ValueError("output buffers too small for the grid")
    t = _lib.torch()
    for buf, dt in ((pts, t.float64), (offsets, t.int64), (counts, t.int64)):
        if buf.dtype != dt or not buf.is_contiguous() or not buf.is_cuda:
            raise ValueError("outputs must be contiguous device tensors: pts float64, "
                             "offsets and counts int64")
    _lib.check(_lib.load().nbx_ridges_async(_lib.ptr(fd), _lib.ptr(evd), _lib.ptr(xd), nx,
                                            _lib.ptr(yd), ny, dx, dy, float(sdd_thresh),
                                            float(percentile), int(min_ridge_pts),
                                            _lib.ptr(pts), _lib.ptr(offsets), _lib.ptr(counts),
                                            _lib.stream_ptr()))
    return pts, offsets, counts


def ftle_ridge_pts(f, eigvec_max, x, y, sdd_thresh=0.0, percentile=0):
    """Sub-pixel FTLE ridge points (k, 2) in raveled grid order: points with
    f > percentile(f) (0 when percentile == 0) where the derivative of f
    along eigvec_max vanishes within the cell and the second directional
    derivative is below -sdd_thresh."""
    fd, evd, xd, yd, nx, ny, dx, dy = _inputs(f, eigvec_max, x, y)
    if nx * ny == 0:
        return np.zeros((0, 2))
    pts, n = ftle_ridge_pts_device(fd, evd, xd, yd, nx, ny, dx, dy, sdd_thresh, percentile)
    return _lib.to_host(pts[:n])


def ftle_ridges(f, eigvec_max, x, y, sdd_thresh=0.0, percentile=0, min_ridge_pts=3):
    """Ridge points grouped into 8-connected ridges: a list of (m, 2) arrays
    in label order (raster order of each ridge's first point), points in
    raveled order, ridges with fewer than min_ridge_pts points dropped."""
    fd, evd, xd, yd, nx, ny, dx, dy = _inputs(f, eigvec_max, x, y)
    if nx * ny == 0:
        return []
    pts, offsets, nr, npts = ftle_ridges_device(fd, evd, xd, yd, nx, ny, dx, dy, sdd_thresh,
                                                percentile, min_ridge_pts)
    P = _lib.to_host(pts[:npts])
    off = _lib.to_host(offsets[:nr + 1])
    return [P[off[k]:off[k + 1]] for k in range(nr)]


def label8(b):
    """8-connected labels of a boolean grid (scipy.ndimage.label with a full
    3x3 structure): (labels (nx, ny) int32, count)."""
    b = np.asarray(b)
    nx, ny = b.shape
    bd = _lib.to_dev(b.astype(np.uint8), np.uint8)
    lab = _lib.empty((nx, ny), "i32")
    n = ctypes.c_int32()
    _lib.check(_lib.load().nbx_label8(_lib.ptr(bd), nx, ny, _lib.ptr(lab), ctypes.byref(n),
                                      _lib.stream_ptr()))
    return _lib.to_host(lab), n.value


def ftle_ordered_ridges(f, eigvec_max, x, y, dist_tol, ep_tan_ang=np.pi / 4, min_ridge_pts=5,
                        sdd_thresh=0.0, percentile=0, c=1.0):
    """Ordered FTLE ridge curves (ridges.py:720-1054): ridge points linked
    along the ridge tangent from the strongest point outward, then curves
    joined where their endpoints lie within dist_tol and their tangents agree
    within ep_tan_ang; curves shorter than min_ridge_pts are dropped.
    Returns a list of (n_k, 2) arrays of ordered points.  The ridge predicate
    and percentile run on the GPU, the serial linking in native host code."""
    fd, evd, xd, yd, nx, ny, _, _ = _inputs(f, eigvec_max, x, y)
    n = nx * ny
    if n == 0:
        return []
    pts = np.zeros((n, 2))
    offsets = np.zeros(n + 1, np.int64)
    nr, npt = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.load().nbx_ordered_ridges(
        _lib.ptr(fd), _lib.ptr(evd), _lib.ptr(xd), nx, _lib.ptr(yd), ny, float(dist_tol),
        float(ep_tan_ang), int(min_ridge_pts), float(sdd_thresh), float(percentile), float(c),
        pts.ctypes.data, n, offsets.ctypes.data, n + 1, ctypes.byref(nr), ctypes.byref(npt),
        _lib.stream_ptr()))
    return [pts[offsets[k]:offsets[k + 1]].copy() for k in range(nr.value)]
