"""Ordered FTLE ridges: the Python restatement (oracle/ordered.py) pinned by
ridges whose answer is known in closed form, and the library's native
linking (nbx_ordered_ridges_link, host code, no GPU) against it bit for bit.

The reference's ordered_ridges.pkl (converted to data by
tests/golden/make_pickle_fixtures.py) pins the exact sequence on the fixture
world: two curves of 6 and 7 points at dist_tol 0.1, in the reference's
order and direction.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import ordered as O


def _grid(nx=61, ny=41, lx=3.0, ly=2.0):
    return np.linspace(0, lx, nx), np.linspace(0, ly, ny)


def _straight(x, y, y0=1.03, w=0.15, gap=None):
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = np.exp(-((Y - y0) / w) ** 2)
    if gap is not None:
        f[(X > gap[0]) & (X < gap[1])] = 0.0
    ev = np.zeros(X.shape + (2,))
    ev[..., 1] = 1.0
    return f, ev


def _circle(x, y, c=(1.5, 1.0), R=0.6, w=0.12):
    X, Y = np.meshgrid(x, y, indexing="ij")
    dx, dy = X - c[0], Y - c[1]
    r = np.hypot(dx, dy)
    f = np.exp(-((r - R) / w) ** 2)
    ev = np.stack([dx / np.maximum(r, 1e-12), dy / np.maximum(r, 1e-12)], -1)
    return f, ev


def product(f, ev, x, y, dist_tol, ep_tan_ang=math.pi / 4, min_ridge_pts=5, sdd_thresh=0.0,
            percentile=0, c=1.0):
    """The library's host linking on the oracle's dense predicate buffers."""
    from numbacs_amd import _lib

    nx, ny = f.shape
    pts, vecs, sdd, h = O.ridge_connect_fields(f, ev, x, y, sdd_thresh, percentile)
    n = nx * ny
    out = np.zeros((n, 2))
    off = np.zeros(n + 1, np.int64)
    nr, npt = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.load().nbx_ordered_ridges_link(
        pts.ctypes.data, vecs.ctypes.data, sdd.ctypes.data, nx, ny, float(h), float(dist_tol),
        float(ep_tan_ang), int(min_ridge_pts), float(sdd_thresh), float(c), out.ctypes.data, n,
        off.ctypes.data, n + 1, ctypes.byref(nr), ctypes.byref(npt)))
    return [out[off[k]:off[k + 1]] for k in range(nr.value)]


def _monotone(a):
    d = np.diff(a)
    return bool((d > 0).all() or (d < 0).all())


def test_straight_ridge_is_one_ordered_line():
    x, y = _grid()
    f, ev = _straight(x, y)
    r = O.ordered_ridges(f, ev, x, y, 0.1)
    assert len(r) == 1
    assert len(r[0]) == len(x) - 4                 # every interior column
    assert _monotone(r[0][:, 0])                   # ordered along the ridge
    assert np.abs(r[0][:, 1] - 1.03).max() < 0.01  # on the ridge line (subpixel)


@pytest.mark.parametrize("tol,n_expected", [(0.05, 2), (0.5, 1)])
def test_gap_joined_within_dist_tol(tol, n_expected):
    """A ridge cut by a gap of ~0.3: two curves, joined into one ordered
    curve when dist_tol exceeds the gap."""
    x, y = _grid()
    f, ev = _straight(x, y, gap=(1.2, 1.45))
    r = O.ordered_ridges(f, ev, x, y, tol)
    assert len(r) == n_expected
    if n_expected == 1:
        assert _monotone(r[0][:, 0])
        steps = np.abs(np.diff(r[0][:, 0]))
        assert steps.max() > 0.25                  # the joint crosses the gap


def test_min_ridge_pts_and_angle():
    """A short piece is dropped by min_ridge_pts; a perpendicular piece within
    dist_tol is joined only when ep_tan_ang allows its angle."""
    x, y = _grid()
    f, ev = _straight(x, y, gap=(0.2, 2.6))      # pieces of 2 and 5 columns survive
    assert len(O.ordered_ridges(f, ev, x, y, 0.01, min_ridge_pts=5)) == 1
    assert len(O.ordered_ridges(f, ev, x, y, 0.01, min_ridge_pts=2)) == 2
    f2, ev2 = _straight(x, y, gap=(1.5, 3.1))
    X, Y = np.meshgrid(x, y, indexing="ij")
    g = np.exp(-((X - 2.2) / 0.15) ** 2) * (Y < 0.8)
    ev_g = np.zeros_like(ev2)
    ev_g[..., 0] = 1.0
    use = g > f2
    f3 = np.where(use, g, f2)
    ev3 = np.where(use[..., None], ev_g, ev2)
    r = O.ordered_ridges(f3, ev3, x, y, 5.0, ep_tan_ang=math.pi / 8)
    assert len(r) == 2
    # with no angle limit the same two pieces join into one curve
    assert len(O.ordered_ridges(f3, ev3, x, y, 5.0, ep_tan_ang=math.pi)) == 1


def test_circle_is_one_closed_curve():
    x, y = _grid(91, 61)
    f, ev = _circle(x, y)
    r = O.ordered_ridges(f, ev, x, y, 0.05)
    assert len(r) == 1
    p = r[0]
    rad = np.hypot(p[:, 0] - 1.5, p[:, 1] - 1.0)
    assert np.abs(rad - 0.6).max() < 0.02
    step = np.hypot(*np.diff(p, axis=0).T)
    assert step.max() < 2.5 * (x[1] - x[0])      # consecutive points are neighbours
    ang = np.unwrap(np.arctan2(p[:, 1] - 1.0, p[:, 0] - 1.5))
    assert _monotone(ang) and abs(ang[-1] - ang[0]) > 1.8 * math.pi


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(u, v) for u, v in zip(a, b))


def test_oracle_matches_reference_fixture(coords_dg, ftle_data, C_eig_data,
                                          ordered_ridge_data):
    """tests/test_extraction.py:22-28 of the reference: ftle_ordered_ridges
    on the fixture world at dist_tol 0.1 (defaults otherwise) gives the two
    curves of ordered_ridges.pkl, point for point and in the same direction
    (within 2e-9: the pickle predates the float32 storage of its inputs);
    the native linking gives the same."""
    x, y = coords_dg
    f = ftle_data.astype(np.float64)
    ev = C_eig_data[1][:, :, :, 1].astype(np.float64)
    want = ordered_ridge_data
    assert [len(r) for r in want] == [6, 7]
    for got in (O.ordered_ridges(f, ev, x, y, 0.1), product(f, ev, x, y, 0.1)):
        assert [len(r) for r in got] == [6, 7]
        for g, w in zip(got, want):
            assert np.abs(g - w).max() < 2e-9


@pytest.mark.parametrize("case", ["straight", "gap", "circle", "fixture", "random"])
def test_native_linking_matches_oracle(case, coords_dg, ftle_data, C_eig_data):
    x, y = _grid(91, 61)
    kw = {}
    if case == "straight":
        x, y = _grid()
        f, ev = _straight(x, y)
    elif case == "gap":
        x, y = _grid()
        f, ev = _straight(x, y, gap=(1.2, 1.45))
        kw = dict(dist_tol=0.5)
    elif case == "circle":
        f, ev = _circle(x, y)
    elif case == "fixture":
        x, y = coords_dg
        f = ftle_data.astype(np.float64)
        ev = C_eig_data[1][:, :, :, 1].astype(np.float64)
    else:
        # six crossing straight ridges of random direction; the normal at a
        # point is that of the strongest ridge there
        rng = np.random.default_rng(11)
        X, Y = np.meshgrid(x, y, indexing="ij")
        ang = rng.uniform(0, np.pi, 6)
        cx, cy = rng.uniform(0.5, 2.5, 6), rng.uniform(0.3, 1.7, 6)
        amp = rng.uniform(0.3, 1.0, 6)
        ridges = np.stack([amp[k] * np.exp(-((np.sin(ang[k]) * (X - cx[k]) -
                                               np.cos(ang[k]) * (Y - cy[k])) / 0.1) ** 2)
                           for k in range(6)])
        f = ridges.sum(0)
        top = np.argmax(ridges, axis=0)
        ev = np.stack([np.sin(ang)[top], -np.cos(ang)[top]], -1)
        kw = dict(percentile=50, sdd_thresh=1.0, min_ridge_pts=3)
    dist_tol = kw.pop("dist_tol", 0.1)
    want = O.ordered_ridges(f, ev, x, y, dist_tol, **kw)
    got = product(f, ev, x, y, dist_tol, **kw)
    assert _same(got, want)
    assert len(want) >= (4 if case == "random" else 1)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("tol", [0.01, 0.07, 0.3, 1.5])
def test_native_joining_across_tolerances(seed, tol):
    """Crossing ridges of random direction at tolerances from below the grid
    step to most of the domain: the library's endpoint joining (a uniform
    grid of endpoints) takes the restatement's decisions (a scan over all
    remaining endpoints) exactly."""
    rng = np.random.default_rng(100 + seed)
    x, y = _grid(121, 81)
    X, Y = np.meshgrid(x, y, indexing="ij")
    nr = int(rng.integers(4, 14))
    ang = rng.uniform(0, np.pi, nr)
    cx, cy = rng.uniform(0.3, 2.7, nr), rng.uniform(0.2, 1.8, nr)
    amp = rng.uniform(0.3, 1.0, nr)
    ridges = np.stack([amp[k] * np.exp(-((np.sin(ang[k]) * (X - cx[k]) -
                                           np.cos(ang[k]) * (Y - cy[k])) / 0.08) ** 2)
                       for k in range(nr)])
    f = ridges.sum(0)
    top = np.argmax(ridges, axis=0)
    ev = np.stack([np.sin(ang)[top], -np.cos(ang)[top]], -1)
    for kw in (dict(), dict(percentile=60, min_ridge_pts=3),
               dict(ep_tan_ang=math.pi / 2, min_ridge_pts=2)):
        want = O.ordered_ridges(f, ev, x, y, tol, **kw)
        assert _same(product(f, ev, x, y, tol, **kw), want)
