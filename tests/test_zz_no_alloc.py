"""nbx_ridges_noalloc (the ridge pipeline on the library's fixed device
scratch array, no allocation) against nbx_ridges through the C ABI: two calls
in a row, the second on a smaller grid, bit-equal points, offsets and
counts."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(nx, ny, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 3, nx)
    y = np.linspace(-1, 1, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = np.zeros((nx, ny))
    for _ in range(12):
        a, b, c = rng.uniform(0.5, 6, 3)
        f += np.exp(-((np.sin(a * X + c) - np.cos(b * Y)) ** 2) * 8)
    th = rng.uniform(0, np.pi, (nx, ny))
    ev = np.stack((np.cos(th), np.sin(th)), -1)
    return f, ev, x, y


def _call(name, f, ev, x, y, percentile, min_pts=3):
    from numbacs_amd import _lib

    lib = _lib.load()
    nx, ny = f.shape
    fd, evd, xd, yd = (_lib.to_dev(a) for a in (f, ev, x, y))
    pts = _lib.zeros((nx * ny, 2))
    off = _lib.zeros((nx * ny + 1,), "i64")
    nr, npts = ctypes.c_int64(-1), ctypes.c_int64(-1)
    _lib.check(getattr(lib, name)(_lib.ptr(fd), _lib.ptr(evd), _lib.ptr(xd), nx, _lib.ptr(yd),
                                  ny, float(x[1] - x[0]), float(y[1] - y[0]), 0.0,
                                  float(percentile), min_pts, _lib.ptr(pts), _lib.ptr(off),
                                  ctypes.byref(nr), ctypes.byref(npts), _lib.stream_ptr()))
    P = _lib.to_host(pts[:npts.value])
    O = _lib.to_host(off[:nr.value + 1])
    return nr.value, npts.value, P, O


@pytest.mark.parametrize("order", ["noalloc_first", "existing_first"])
def test_noalloc_ridges_match_existing_entry_point(order):
    from numbacs_amd import _lib

    _lib.torch()
    cases = [(_field(676, 251, 11), 90), (_field(311, 173, 7), 0)]
    assert cases[1][0][0].size < cases[0][0][0].size
    for (f, ev, x, y), q in cases:
        names = ["nbx_ridges_noalloc", "nbx_ridges"]
        if order == "existing_first":
            names.reverse()
        got = {n: _call(n, f, ev, x, y, q) for n in names}
        a, b = got["nbx_ridges_noalloc"], got["nbx_ridges"]
        assert a[0] == b[0] and a[1] == b[1]
        assert a[0] > 3 and a[1] > 100
        assert np.array_equal(a[2], b[2])
        assert np.array_equal(a[3], b[3])
