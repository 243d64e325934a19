"""ftle_ordered_ridges on the GPU path (ridge predicate and percentile on the
device, linking on the host) against the Python restatement.

The device predicate evaluates the reference's stencil in source order
without FMA contraction, so ridge points and second directional
derivatives equal the restatement's bit for bit, and with them every
linking decision: the curves are compared exactly.
"""
import math

import numpy as np
import pytest

from oracle import ordered as O
from test_ordered_ridges import _circle, _grid, _straight

pytestmark = pytest.mark.gpu


def _check(got, want):
    assert [len(a) for a in got] == [len(a) for a in want]
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["fixture", "straight", "gap", "circle", "percentile"])
def test_ordered_ridges_gpu_vs_oracle(case, coords_dg, ftle_data, C_eig_data):
    from numbacs_amd.extraction import ftle_ordered_ridges

    kw = {}
    if case == "fixture":
        x, y = coords_dg
        f = ftle_data.astype(np.float64)
        ev = C_eig_data[1][:, :, :, 1].astype(np.float64)
        tol = 0.1
    elif case in ("straight", "gap"):
        x, y = _grid()
        f, ev = _straight(x, y, gap=(1.2, 1.45) if case == "gap" else None)
        tol = 0.5
    else:
        x, y = _grid(91, 61)
        f, ev = _circle(x, y)
        tol = 0.05
        if case == "percentile":
            kw = dict(percentile=80, sdd_thresh=0.5, min_ridge_pts=3)
    got = ftle_ordered_ridges(f, ev, x, y, tol, **kw)
    want = O.ordered_ridges(f, ev, x, y, tol, **kw)
    assert len(want) >= 1
    _check(got, want)


def test_ordered_ridges_reference_fixture(coords_dg, ftle_data, C_eig_data, ordered_ridge_data):
    """tests/test_extraction.py:22-28 of the reference on the GPU path: two
    curves of 6 and 7 points equal to ordered_ridges.pkl in order and
    direction (within 2e-9: the pickle predates the float32 storage of its
    inputs)."""
    from numbacs_amd.extraction import ftle_ordered_ridges

    x, y = coords_dg
    f = ftle_data.astype(np.float64)
    ev = C_eig_data[1][:, :, :, 1].astype(np.float64)
    got = ftle_ordered_ridges(f, ev, x, y, 0.1)
    assert [len(a) for a in got] == [len(a) for a in ordered_ridge_data] == [6, 7]
    for a, b in zip(got, ordered_ridge_data):
        assert np.abs(a - b).max() < 2e-9


def test_ordered_ridges_gridded_workload_structure():
    """On the gridded workload's FTLE field: every curve is an ordered chain
    of neighbouring ridge points or joined pieces within dist_tol, the curves
    hold distinct points, and they equal the restatement's on the same field."""
    import torch

    from numbacs_amd import _lib
    from numbacs_amd.diagnostics import ftle_eig_2d_device
    from numbacs_amd.extraction import ftle_ordered_ridges
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import flowmap_grid_2d_device
    from numbacs_amd.workloads import particle_axes, synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity(nt=40)
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    f = get_flow_2D(grid, Cu, Cv, spherical=1)
    x, y = particle_axes(0.4)
    xd, yd = _lib.to_dev(x), _lib.to_dev(y)
    fm = flowmap_grid_2d_device(f, 36.0, -24.0, xd, yd, [-1.0])
    ftle, evec = ftle_eig_2d_device(fm, -24.0, x[1] - x[0], y[1] - y[0])
    torch.cuda.synchronize()
    tol = 3 * (x[1] - x[0])
    curves = ftle_ordered_ridges(ftle, evec, x, y, tol, percentile=90, min_ridge_pts=5)
    assert len(curves) > 5
    h = max(x[1] - x[0], y[1] - y[0])
    for cv in curves:
        steps = np.hypot(*np.diff(cv, axis=0).T)
        assert (steps < max(tol, 2 * math.sqrt(2) * h)).all()
    allp = np.concatenate(curves)
    assert len(np.unique(allp, axis=0)) == len(allp)
    want = O.ordered_ridges(lib_host(ftle), lib_host(evec), x, y, tol, percentile=90,
                            min_ridge_pts=5)
    _check(curves, want)


def lib_host(t):
    from numbacs_amd import _lib

    return _lib.to_host(t)
