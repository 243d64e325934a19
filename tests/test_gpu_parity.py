"""Parity of the HIP path (through the C ABI) with the CPU oracle and the
reference's own fixtures.  Tolerances: flow maps 1e-7 relative unless the
trajectory's own conditioning (measured by a 1-ulp perturbation of the
initial condition in the oracle) is looser; FTLE 1e-6 relative; integer and
index results (ridge flags, labels, orderings) exact; percentile bit-exact.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle
from conftest import FLOWMAP_RTOL, apply_mask, flowmap_agreement, ic_sensitivity, reconstruct

pytestmark = pytest.mark.gpu

T_FIX = 8.0


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def nb():
    import numbacs_amd
    from numbacs_amd import _lib

    _lib.torch()
    return numbacs_amd


# ---------------------------------------------------------------- flows --


def test_predefined_velocity_literals(nb, literals):
    """Velocity of the double gyre and Bickley jet at the 27 points of the
    reference test (tests/test_flows.py:307-357)."""
    from numbacs_amd.flows import flow_velocity, get_predefined_flow

    ti = np.array([0.05, 0.12, 0.18])
    xi = np.array([0.1, 0.4, 0.7])
    Ti, Xi, Yi = np.meshgrid(ti, xi, xi, indexing="ij")
    pts = np.column_stack((Xi.ravel(), Yi.ravel()))
    f, p, dom = get_predefined_flow("double_gyre")
    v = flow_velocity(f, Ti.ravel(), pts, p)
    assert np.allclose(v.T, literals["dg_vel"])
    assert np.allclose(dom, literals["dg_domain"])
    f, p, _ = get_predefined_flow("bickley_jet")
    v = flow_velocity(f, Ti.ravel(), pts, p)
    assert np.allclose(v.T, literals["bickley_vel"])
    # against the oracle to machine precision
    assert relerr(v, oracle.rhs_eval("bickley_jet", p, Ti.ravel(), pts)) < 1e-14


def test_prefilter_literals(nb, literals):
    """get_interp_arrays_2D coefficients (tests/test_flows.py:68-126)."""
    from numbacs_amd.flows import get_interp_arrays_2D

    t = np.array([0.0, 0.1, 0.2])
    x = np.array([0.0, 0.5, 1.0])
    T_, X, Y = np.meshgrid(t, x, x, indexing="ij")
    u = np.sin(X) * np.cos(Y) + np.sin(T_)
    v = np.cos(X) * np.sin(Y) + np.cos(T_)
    grid, Cu, Cv = get_interp_arrays_2D(t, x, x, u, v)
    assert np.allclose(grid, ((0.0, 0.2, 3), (0.0, 1.0, 3), (0.0, 1.0, 3)))
    assert np.allclose(Cu, literals["prefilter3_Cu"])
    assert np.allclose(Cv, literals["prefilter3_Cv"])


def test_prefilter_vs_oracle_long_axes(nb):
    from numbacs_amd.flows import get_interp_arrays_2D

    rng = np.random.default_rng(3)
    U = rng.standard_normal((70, 45, 33))
    V = rng.standard_normal((70, 45, 33))
    t, x, y = np.arange(70.0), np.linspace(0, 1, 45), np.linspace(0, 2, 33)
    _, Cu, Cv = get_interp_arrays_2D(t, x, y, U, V)
    assert relerr(Cu, oracle.prefilter3(U)) < 1e-13
    assert relerr(Cv, oracle.prefilter3(V)) < 1e-13


def test_spline_velocity_literals(nb, literals):
    """Tricubic evaluation at 27 in-grid points (tests/test_flows.py:208-258)."""
    from numbacs_amd.flows import flow_velocity, get_flow_2D, get_interp_arrays_2D

    t = np.array([0.0, 0.1, 0.2])
    x = np.array([0.0, 0.5, 1.0])
    T_, X, Y = np.meshgrid(t, x, x, indexing="ij")
    u = np.sin(X) * np.cos(Y) + np.sin(T_)
    v = np.cos(X) * np.sin(Y) + np.cos(T_)
    grid, Cu, Cv = get_interp_arrays_2D(t, x, x, u, v)
    f = get_flow_2D(grid, Cu, Cv)
    ti = np.array([0.05, 0.12, 0.18])
    xi = np.array([0.1, 0.4, 0.7])
    Ti, Xi, Yi = np.meshgrid(ti, xi, xi, indexing="ij")
    vel = flow_velocity(f, Ti.ravel(), np.column_stack((Xi.ravel(), Yi.ravel())), [1.0])
    assert np.allclose(vel.T, literals["spline3_vel"])


@pytest.mark.parametrize("extrap", ["constant", "linear", "nearest"])
def test_spline_velocity_out_of_grid_vs_oracle(nb, extrap):
    """Extrapolated values follow the oracle's semantics (constant: zero
    outside the box, as the dependency's eval_linear; linear and nearest are
    unpinned by the reference)."""
    from numbacs_amd.flows import flow_velocity, get_flow_2D, get_interp_arrays_2D

    rng = np.random.default_rng(5)
    t, x, y = np.arange(6.0), np.linspace(0, 1, 9), np.linspace(-1, 1, 7)
    U = rng.standard_normal((6, 9, 7))
    V = rng.standard_normal((6, 9, 7))
    grid, Cu, Cv = get_interp_arrays_2D(t, x, y, U, V)
    f = get_flow_2D(grid, Cu, Cv, extrap_mode=extrap)
    n = 500
    tq = rng.uniform(-1, 6, n)
    pq = np.column_stack((rng.uniform(-0.3, 1.3, n), rng.uniform(-1.5, 1.5, n)))
    vel = flow_velocity(f, tq, pq, [1.0])
    ref = oracle.rhs_eval("spline", np.array([1.0]), tq, pq,
                          spline=oracle.Spline(grid, Cu, Cv, extrap_mode=extrap))
    assert relerr(vel, ref) < 1e-12


# ------------------------------------------------------------- flow maps --


def test_flowmap_grid_2D_fixture(nb, coords_dg, fm_data, mask_dg):
    """tests/test_integration.py:55-64 of the reference."""
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2D

    f, p, _ = get_predefined_flow("double_gyre")
    x, y = coords_dg
    fm = flowmap_grid_2D(f, 0.0, T_FIX, x, y, p)
    assert np.allclose(fm.astype(np.float32), fm_data)
    ref, _, _ = oracle.flowmap_grid_2d("double_gyre", p, 0.0, T_FIX, x, y)
    assert relerr(fm, ref) < 1e-10
    fmm = flowmap_grid_2D(f, 0.0, T_FIX, x, y, p, mask=mask_dg)
    assert np.allclose(fmm.astype(np.float32), apply_mask(fm_data, mask_dg))
    assert (fmm[mask_dg] == 0).all()


def test_flowmap_points_fixture(nb, coords_dg, fm_data):
    """tests/test_integration.py:38-42 of the reference."""
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap

    f, p, _ = get_predefined_flow("double_gyre")
    x, y = coords_dg
    X, Y = np.meshgrid(x, y, indexing="ij")
    pts = np.column_stack((X.ravel(), Y.ravel()))
    fm = flowmap(f, 0.0, T_FIX, pts, p).reshape(21, 11, 2)
    assert np.allclose(fm.astype(np.float32), fm_data)


@pytest.mark.parametrize("case", ["dg_fwd8", "dg_c1_sub", "dg_t20_sub"])
def test_flowmap_vs_scipy_dop853(nb, scipy_golden, case):
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2D

    G = scipy_golden
    f = get_predefined_flow("double_gyre", return_default_params=False, return_domain=False)
    t0, T = G[case + "_t0T"]
    fm = flowmap_grid_2D(f, t0, T, G[case + "_x"], G[case + "_y"], G[case + "_params"])
    assert relerr(fm, G[case + "_fm"]) < FLOWMAP_RTOL


def test_flowmap_double_gyre_full_vs_oracle(nb):
    """Double gyre 401x201, t0=0, T=-10 (the CPU-runnable configuration)."""
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2D

    f, p, _ = get_predefined_flow("double_gyre", int_direction=-1.0)
    x = np.linspace(0, 2, 401)
    y = np.linspace(0, 1, 201)
    fm, st = flowmap_grid_2D(f, 0.0, -10.0, x, y, p, return_status=True)
    ref, st_ref, _ = oracle.flowmap_grid_2d("double_gyre", p, 0.0, -10.0, x, y)
    assert (st == 1).all() and (st_ref == 1).all()
    assert relerr(fm, ref) < 1e-7
    from numbacs_amd.diagnostics import ftle_grid_2D

    ftle = ftle_grid_2D(fm, -10.0, 0.005, 0.005)
    ftle_ref = oracle.ftle_grid_2d(ref, -10.0, 0.005, 0.005)
    assert relerr(ftle, ftle_ref) < 1e-6


def test_flowmap_bickley_vs_oracle(nb):
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2D

    f, p, _ = get_predefined_flow("bickley_jet")
    p = p.copy()
    p[0] = -1.0
    x = np.linspace(0, math.pi * 6.371, 2001)[::20]
    y = np.linspace(-3, 3, 801)[::20]
    fm, st = flowmap_grid_2D(f, 0.0, -40.0, x, y, p, return_status=True)
    ref, sens = ic_sensitivity(
        lambda xx, yy: oracle.flowmap_grid_2d("bickley_jet", p, 0.0, -40.0, xx, yy)[0], x, y)
    assert (st == 1).all()
    ok, frac, err = flowmap_agreement(fm, ref, sens)
    assert ok, float(err.max())
    assert frac > 0.8, frac


def test_flowmap_bickley_vs_scipy(nb, scipy_golden):
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2D

    """Every particle of this 40-day case is chaotic (a 1e-9 relative shift
    of the start moves it by 6e-11 .. 4e-3 of the domain), so rounding-level
    differences from scipy (device math, contraction) are amplified and the
    per-particle error is bimodal; agreement is judged against the oracle's
    measured sensitivity, as for the oracle comparison."""
    G = scipy_golden
    case = "bickley_t40_sub"
    x, y, p = G[case + "_x"], G[case + "_y"], G[case + "_params"]
    f = get_predefined_flow("bickley_jet", return_default_params=False, return_domain=False)
    fm = flowmap_grid_2D(f, 0.0, -40.0, x, y, p)
    ref = G[case + "_fm"]
    _, sens = ic_sensitivity(
        lambda xx, yy: oracle.flowmap_grid_2d("bickley_jet", p, 0.0, -40.0, xx, yy)[0], x, y)
    ok, frac, err = flowmap_agreement(fm, ref, sens)
    assert ok, float(err.max())
    assert frac > 0.8, frac


@pytest.fixture(scope="module")
def c2_small():
    """The gridded-velocity configuration's field with every 4th particle."""
    from numbacs_amd.workloads import particle_axes, synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity()
    xp, yp = particle_axes()
    return t, lon, lat, U, V, xp[::4], yp[::4]


def test_flowmap_spline_spherical_vs_oracle(nb, c2_small):
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import flowmap_grid_2D

    t, lon, lat, U, V, xp, yp = c2_small
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    assert relerr(Cu, oracle.prefilter3(U)) < 1e-13
    f = get_flow_2D(grid, Cu, Cv, spherical=1)
    fm, st = flowmap_grid_2D(f, 72.0, -72.0, xp, yp, np.array([-1.0]), return_status=True)
    S = oracle.Spline(grid, Cu, Cv, spherical=1)
    ref, sens = ic_sensitivity(
        lambda xx, yy: oracle.flowmap_grid_2d("spline", np.array([-1.0]), 72.0, -72.0, xx, yy,
                                              spline=S)[0], xp, yp)
    assert (st == 1).all()
    ok, frac, err = flowmap_agreement(fm, ref, sens)
    assert ok, float(err.max())
    assert frac > 0.99, frac


def test_flowmap_status_budget(nb):
    """A step budget of 1 stops every solve with status -2 (Hairer's NMAX)."""
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2D

    f, p, _ = get_predefined_flow("double_gyre")
    x, y = np.linspace(0, 2, 9), np.linspace(0, 1, 5)
    mask = np.zeros((9, 5), bool)
    mask[0, 0] = True
    fm, st = flowmap_grid_2D(f, 0.0, 8.0, x, y, p, mask=mask, max_steps=1, return_status=True)
    ref, st_ref, _ = oracle.flowmap_grid_2d("double_gyre", p, 0.0, 8.0, x, y, mask=mask, nmax=1)
    assert st[0, 0] == 0 and (st.ravel()[1:] == -2).all()
    assert (st == st_ref).all()
    # two steps of a truncated solve: the intermediate state depends on the
    # step sizes, which carry the rounding of the cancelling error estimate
    assert relerr(fm, ref) < FLOWMAP_RTOL


def test_flowmap_empty(nb):
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap, flowmap_grid_2D

    f, p, _ = get_predefined_flow("double_gyre")
    assert flowmap_grid_2D(f, 0.0, 1.0, np.zeros(0), np.linspace(0, 1, 4), p).shape == (0, 4, 2)
    assert flowmap(f, 0.0, 1.0, np.zeros((0, 2)), p).shape == (0, 2)


# ------------------------------------------------------------ FTLE / eig --


def test_ftle_fixture(nb, coords_dg, fm_data, ftle_data, mask_dg):
    """tests/test_diagnostics.py:82-96 of the reference."""
    from numbacs_amd.diagnostics import ftle_grid_2D

    x, y = coords_dg
    ftle = ftle_grid_2D(fm_data, T_FIX, x[1], y[1])
    assert np.allclose(ftle.astype(np.float32), ftle_data)
    assert relerr(ftle, oracle.ftle_grid_2d(fm_data, T_FIX, x[1], y[1])) < 1e-13
    ftle_m = ftle_grid_2D(fm_data, T_FIX, x[1], y[1], mask=mask_dg)
    assert np.allclose(ftle_m.astype(np.float32), apply_mask(ftle_data, mask_dg))


def test_C_eig_fixture(nb, coords_dg, fm_data, C_eig_data):
    """tests/test_diagnostics.py:133-150 of the reference (reconstruction)."""
    from numbacs_amd.diagnostics import C_eig_2D

    x, y = coords_dg
    ev, V = C_eig_2D(fm_data, x[1], y[1])
    ev_ref, V_ref = C_eig_data
    assert np.allclose(ev.astype(np.float32), ev_ref)
    assert np.allclose(reconstruct(ev, V).astype(np.float32),
                       reconstruct(ev_ref.astype(np.float64), V_ref.astype(np.float64)).astype(
                           np.float32))
    ev_o, V_o = oracle.C_eig_2d(fm_data, x[1], y[1])
    assert relerr(ev, ev_o) < 1e-13
    assert relerr(V, V_o) < 1e-12


def test_ftle_eig_fused_matches_separate(nb):
    from numbacs_amd import _lib
    from numbacs_amd.diagnostics import C_eig_2D, ftle_eig_2d_device, ftle_from_eig

    rng = np.random.default_rng(11)
    fm = np.cumsum(rng.standard_normal((67, 45, 2)), axis=0)
    ev, V = C_eig_2D(fm, 0.3, 0.2)
    fd, ed = ftle_eig_2d_device(_lib.to_dev(fm), -5.0, 0.3, 0.2)
    assert relerr(_lib.to_host(fd), ftle_from_eig(ev[:, :, 1], -5.0)) < 1e-14
    assert np.array_equal(_lib.to_host(ed), V[:, :, :, 1])


# ---------------------------------------------------------------- ridges --


def test_ridge_pts_fixture(nb, coords_dg, ftle_data, C_eig_data, ridge_pts_data):
    """tests/test_extraction.py:6-12 of the reference."""
    from numbacs_amd.extraction import ftle_ridge_pts

    x, y = coords_dg
    _, evecs = C_eig_data
    pts = ftle_ridge_pts(ftle_data, evecs[:, :, :, 1], x, y)
    assert pts.shape == ridge_pts_data.shape
    assert np.allclose(pts, ridge_pts_data)
    ref = oracle.ridge_pts(ftle_data, evecs[:, :, :, 1], x, y)
    assert np.array_equal(pts, ref)  # the predicate runs uncontracted in source order


def test_ridges_fixture_grouping(nb, coords_dg, ftle_data, C_eig_data, ridge_data):
    """ftle_ridges on the fixture world: grouping equals scipy.ndimage.label's
    (the reference's dependency) applied to the oracle's ridge flags, and the
    ridges equal the reference's ridges.pkl (tests/test_extraction.py:14-19;
    within 2e-9: the pickle predates the float32 storage of its inputs)."""
    from numbacs_amd.extraction import ftle_ridges

    x, y = coords_dg
    _, evecs = C_eig_data
    ridges = ftle_ridges(ftle_data, evecs[:, :, :, 1], x, y)
    pts, rb = oracle.ridge_predicate(ftle_data, evecs[:, :, :, 1], x, y)
    lab, n = ndimage.label(rb, structure=ndimage.generate_binary_structure(2, 2))
    expect = []
    for i in range(1, n + 1):
        inds = np.nonzero(lab.ravel() == i)[0]
        if len(inds) >= 3:
            expect.append(pts[inds])
    assert [len(r) for r in ridges] == [len(r) for r in expect] == [6, 9, 3]
    assert all(np.array_equal(a, b) for a, b in zip(ridges, expect))
    assert [len(r) for r in ridge_data] == [6, 9, 3]
    assert all(np.abs(a - b).max() < 2e-9 for a, b in zip(ridges, ridge_data))


@pytest.mark.parametrize("q", [0, 50, 90, 99.5, 100])
def test_percentile_bit_exact(nb, q):
    from numbacs_amd import _lib

    rng = np.random.default_rng(int(q * 10))
    f = rng.standard_normal(100003) * 3
    f[::7] = 0.0  # ties, like the zero border of an FTLE field
    f[5] = -0.0
    fd = _lib.to_dev(f)
    out = _lib.empty((1,))
    _lib.check(_lib.load().nbx_percentile(_lib.ptr(fd), f.size, float(q), _lib.ptr(out),
                                          _lib.stream_ptr()))
    got = _lib.to_host(out)[0]
    assert got == np.percentile(f, q)
    assert got == oracle.percentile(f, q)


def test_percentile_nan(nb):
    from numbacs_amd import _lib

    f = np.arange(1000, dtype=np.float64)
    f[17] = np.nan
    fd = _lib.to_dev(f)
    out = _lib.empty((1,))
    _lib.check(_lib.load().nbx_percentile(_lib.ptr(fd), f.size, 90.0, _lib.ptr(out),
                                          _lib.stream_ptr()))
    assert np.isnan(_lib.to_host(out)[0]) and np.isnan(np.percentile(f, 90))


@pytest.mark.parametrize("shape,density", [((1, 1), 1.0), ((3, 200), 0.5), ((257, 131), 0.35),
                                           ((1000, 37), 0.6)])
def test_label8_vs_scipy(nb, shape, density):
    from numbacs_amd.extraction import label8

    rng = np.random.default_rng(shape[0])
    b = rng.random(shape) < density
    lab, n = label8(b)
    lab_ref, n_ref = ndimage.label(b, structure=ndimage.generate_binary_structure(2, 2))
    assert n == n_ref
    assert np.array_equal(lab, lab_ref)


def _field_and_evec(nx, ny, seed):
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
    f[0, :] = f[-1, :] = f[:, 0] = f[:, -1] = 0
    return f, ev, x, y


@pytest.mark.parametrize("percentile", [0, 90])
def test_ridges_vs_oracle_random(nb, percentile):
    from numbacs_amd.extraction import ftle_ridge_pts, ftle_ridges

    f, ev, x, y = _field_and_evec(311, 173, 7)
    pts = ftle_ridge_pts(f, ev, x, y, percentile=percentile)
    ref = oracle.ridge_pts(f, ev, x, y, percentile_=percentile)
    assert pts.shape == ref.shape and len(ref) > 100
    assert np.array_equal(pts, ref)
    ridges = ftle_ridges(f, ev, x, y, percentile=percentile, min_ridge_pts=3)
    ridges_ref = oracle.ridges(f, ev, x, y, percentile_=percentile, min_ridge_pts=3)
    assert len(ridges) == len(ridges_ref) > 3
    for a, b in zip(ridges, ridges_ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("percentile", [20, 97.5, 100])
def test_ridges_percentile_sized_grouping(nb, percentile):
    """With a percentile threshold the grouping stage is sized by the bound
    on the ridge-point count (values above the percentile) instead of the
    count read back: far above the count (20), close to it (97.5) and zero
    (100, no point can exceed the maximum); against the oracle."""
    from numbacs_amd.extraction import ftle_ridges

    f, ev, x, y = _field_and_evec(311, 173, 7)
    ridges = ftle_ridges(f, ev, x, y, percentile=percentile, min_ridge_pts=3)
    ridges_ref = oracle.ridges(f, ev, x, y, percentile_=percentile, min_ridge_pts=3)
    assert len(ridges) == len(ridges_ref)
    assert (len(ridges) == 0) == (percentile == 100)
    for a, b in zip(ridges, ridges_ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("percentile", [0, 50])
def test_ridges_grouping_beyond_one_workgroup(nb, percentile):
    """28,120 ridge points on 95 wavy ridge lines: the grouping sorts 32,768
    keys (the count read back, percentile 0) or 262,144 (sized by the
    median's bound of 180,000, mostly padding), past the 16,384 of the
    workloads above, so the bitonic sort runs global steps between its LDS
    chunks; against the oracle."""
    from numbacs_amd.extraction import ftle_ridges

    nx, ny = 1200, 300
    x, y = np.linspace(0, 3, nx), np.linspace(-1, 1, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = 2 + np.cos(200 * X + 0.3 * np.sin(3 * Y)) + 0.1 * np.cos(7 * Y)
    ev = np.zeros((nx, ny, 2))
    ev[..., 0] = 1.0
    ridges = ftle_ridges(f, ev, x, y, percentile=percentile, min_ridge_pts=3)
    ridges_ref = oracle.ridges(f, ev, x, y, percentile_=percentile, min_ridge_pts=3)
    assert sum(len(r) for r in ridges_ref) > 16384
    assert len(ridges) == len(ridges_ref) == 95
    for a, b in zip(ridges, ridges_ref):
        assert np.array_equal(a, b)


def test_ridges_grouping_many_segments(nb):
    """More than 1024 components (one scan tile of segments): 95 ridge lines
    cut into random dashes by zeroed rows, about 8,000 components of which
    min_ridge_pts drops the short ones, so kept and dropped segments span
    many segment-scan tiles; against the oracle."""
    from numbacs_amd.extraction import ftle_ridges

    nx, ny = 1200, 300
    x, y = np.linspace(0, 3, nx), np.linspace(-1, 1, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = 2 + np.cos(200 * X + 0.3 * np.sin(3 * Y)) + 0.1 * np.cos(7 * Y)
    gap = np.random.default_rng(5).random(ny) < 0.3
    f[:, gap] = 0.0
    ev = np.zeros((nx, ny, 2))
    ev[..., 0] = 1.0
    ridges = ftle_ridges(f, ev, x, y, min_ridge_pts=3)
    ridges_ref = oracle.ridges(f, ev, x, y, min_ridge_pts=3)
    _, rb = oracle.ridge_predicate(f, ev, x, y)
    _, ncomp = oracle.label8(rb)
    assert ncomp > 4 * 1024 and 1024 < len(ridges_ref) < ncomp
    assert len(ridges) == len(ridges_ref)
    for a, b in zip(ridges, ridges_ref):
        assert np.array_equal(a, b)


def test_ridges_two_streams_concurrently(nb):
    """ftle_ridges / percentile calls on two streams from two host threads
    at once (each stream has its own workspace) give the sequential results."""
    import threading

    import torch

    from numbacs_amd.extraction import ftle_ridges

    cases = [(_field_and_evec(301, 211, 1), 90), (_field_and_evec(257, 389, 2), 0)]
    seq = [ftle_ridges(f, ev, x, y, percentile=q) for (f, ev, x, y), q in cases]
    assert all(len(r) > 3 for r in seq)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    got = [[], []]
    errors = []

    def work(k):
        try:
            (f, ev, x, y), q = cases[k]
            with torch.cuda.stream(streams[k]):
                for _ in range(6):
                    got[k].append(ftle_ridges(f, ev, x, y, percentile=q))
        except Exception as e:  # reported by the main thread
            errors.append(e)

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k in range(2):
        assert len(got[k]) == 6
        for r in got[k]:
            assert len(r) == len(seq[k]) and all(np.array_equal(a, b) for a, b in zip(r, seq[k]))


def _split_async(pts, off, cnt):
    from numbacs_amd import _lib

    nr, nkept, nall = (int(v) for v in _lib.to_host(cnt))
    P, O = _lib.to_host(pts[:nkept]), _lib.to_host(off[:nr + 1])
    return [P[O[k]:O[k + 1]] for k in range(nr)], nall


@pytest.mark.parametrize("percentile", [0, 90, 100])
def test_ridges_async_matches_sync(nb, percentile):
    """nbx_ridges_async (counts left in device memory, nothing waits)
    against ftle_ridges and the oracle: the same ridges point for point, and
    counts[2] is the oracle's ridge-point count.  Without a percentile its
    grouping is sized by the grid (53,803 points: 65,536 keys sorted) rather
    than by the count read back; at 100 nothing passes."""
    import torch

    from numbacs_amd import _lib
    from numbacs_amd.extraction import ftle_ridges
    from numbacs_amd.extraction.ridges import ftle_ridges_device_async

    f, ev, x, y = _field_and_evec(311, 173, 7)
    nx, ny = f.shape
    fd, evd, xd, yd = (_lib.to_dev(a) for a in (f, ev, x, y))
    cnt = _lib.zeros((3,), "i64") - 7  # stale values the call must overwrite
    out = ftle_ridges_device_async(fd, evd, xd, yd, nx, ny, float(x[1] - x[0]),
                                   float(y[1] - y[0]), 0.0, percentile, 3, counts=cnt)
    torch.cuda.synchronize()
    got, nall = _split_async(*out)
    want = ftle_ridges(f, ev, x, y, percentile=percentile, min_ridge_pts=3)
    ref = oracle.ridges(f, ev, x, y, percentile_=percentile, min_ridge_pts=3)
    assert nall == len(oracle.ridge_pts(f, ev, x, y, percentile_=percentile))
    assert len(got) == len(want) == len(ref)
    assert (len(got) == 0) == (percentile == 100)
    for a, b, c in zip(got, want, ref):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_ridges_async_pipelined_on_a_second_stream(nb):
    """A field pipeline: each field is written on the main stream and its
    ridges queued on a second stream behind an event, while the main stream
    goes on to the next field (the field buffers reused only after the
    previous field's ridges are done); every field's ridges equal
    ftle_ridges' on that field."""
    import torch

    from numbacs_amd import _lib
    from numbacs_amd.extraction import ftle_ridges
    from numbacs_amd.extraction.ridges import ftle_ridges_device_async

    fields = [_field_and_evec(301, 211, s) for s in (1, 2, 3, 4)]
    x, y = fields[0][2], fields[0][3]
    nx, ny = len(x), len(y)
    want = [ftle_ridges(f, ev, x, y, percentile=90) for f, ev, _, _ in fields]
    assert all(len(w) > 3 for w in want)
    xd, yd = _lib.to_dev(x), _lib.to_dev(y)
    fbuf, evbuf = _lib.empty((nx, ny)), _lib.empty((nx, ny, 2))
    outs = [(_lib.empty((nx * ny, 2)), _lib.empty((nx * ny + 1,), "i64"),
             _lib.empty((3,), "i64")) for _ in fields]
    main, side = torch.cuda.current_stream(), torch.cuda.Stream()
    done = None
    for k, (f, ev, _, _) in enumerate(fields):
        if done is not None:
            main.wait_event(done)
        fbuf.copy_(torch.from_numpy(f))
        evbuf.copy_(torch.from_numpy(ev))
        ready = main.record_event()
        with torch.cuda.stream(side):
            side.wait_event(ready)
            ftle_ridges_device_async(fbuf, evbuf, xd, yd, nx, ny, float(x[1] - x[0]),
                                     float(y[1] - y[0]), 0.0, 90, 3, *outs[k])
            done = side.record_event()
    torch.cuda.synchronize()
    for k in range(len(fields)):
        got, _ = _split_async(*outs[k])
        assert len(got) == len(want[k])
        assert all(np.array_equal(a, b) for a, b in zip(got, want[k]))


def test_ridges_empty_and_degenerate(nb):
    from numbacs_amd.extraction import ftle_ridge_pts, ftle_ridges

    f = np.zeros((6, 5))
    ev = np.zeros((6, 5, 2))
    x, y = np.arange(6.0), np.arange(5.0)
    assert ftle_ridge_pts(f, ev, x, y).shape == (0, 2)
    assert ftle_ridges(f, ev, x, y) == []
    assert ftle_ridges(f, ev, x, y, percentile=90) == []
    # the asynchronous form: a flat field and an empty grid leave zero counts
    # (stale values in the counts buffer are overwritten, not kept)
    import torch

    from numbacs_amd import _lib
    from numbacs_amd.extraction.ridges import ftle_ridges_device_async

    for nx in (6, 0):
        fd = _lib.to_dev(f[:nx]) if nx else _lib.empty((0, 5))
        evd = _lib.to_dev(ev[:nx]) if nx else _lib.empty((0, 5, 2))
        cnt = _lib.zeros((3,), "i64") - 7
        ftle_ridges_device_async(fd, evd, _lib.to_dev(x), _lib.to_dev(y), nx, 5, 1.0, 1.0, 0.0,
                                 90, 3, counts=cnt)
        torch.cuda.synchronize()
        assert _lib.to_host(cnt).tolist() == [0, 0, 0]


# ----------------------------------------------------- trilinear / ingest --
def test_trilinear_flow_literals_and_oracle(nb, literals):
    """get_flow_linear_2D at the reference test's 27 points
    (tests/test_flows.py:261-285) and against the oracle bit for bit (both
    evaluate the interpolant in one source order without contraction)."""
    from numbacs_amd.flows import flow_velocity, get_flow_linear_2D

    t, x, y = np.array([0.0, 0.1, 0.2]), np.array([0.0, 0.5, 1.0]), np.array([0.0, 0.5, 1.0])
    T_, X, Y = np.meshgrid(t, x, y, indexing="ij")
    u = np.sin(X) * np.cos(Y) + np.sin(T_)
    v = np.cos(X) * np.sin(Y) + np.cos(T_)
    grid = ((0.0, 0.2, 3), (0.0, 1.0, 3), (0.0, 1.0, 3))
    ti, xi = np.array([0.05, 0.12, 0.18]), np.array([0.1, 0.4, 0.7])
    Ti, Xi, Yi = np.meshgrid(ti, xi, xi, indexing="ij")
    f = get_flow_linear_2D(grid, u, v)
    vel = flow_velocity(f, Ti.ravel(), np.column_stack((Xi.ravel(), Yi.ravel())), [1.0])
    assert np.allclose(vel.T, literals["linear3_vel"])
    rng = np.random.default_rng(8)
    for extrap in ("constant", "linear", "nearest"):
        f = get_flow_linear_2D(grid, u, v, extrap_mode=extrap)
        tq = rng.uniform(-0.05, 0.25, 400)
        pq = np.column_stack((rng.uniform(-0.2, 1.2, 400), rng.uniform(-0.2, 1.2, 400)))
        got = flow_velocity(f, tq, pq, [-1.0])
        ref = oracle.rhs_eval("linear", np.array([-1.0]), tq, pq,
                              spline=oracle.Linear(grid, u, v, extrap_mode=extrap))
        assert np.array_equal(got, ref), extrap


def test_trilinear_flowmap_vs_oracle(nb):
    """A spherical trilinear flow over the synthetic field: GPU flow maps
    (both precision forms) against the FP64 oracle."""
    from numbacs_amd import _lib
    from numbacs_amd.flows import get_flow_linear_2D
    from numbacs_amd.integration import flowmap_grid_2D, flowmap_grid_2d_device
    from numbacs_amd.workloads import synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity(nt=30)
    grid = ((t[0], t[-1], len(t)), (lon[0], lon[-1], len(lon)), (lat[0], lat[-1], len(lat)))
    f = get_flow_linear_2D(grid, U, V, spherical=1)
    x, y = np.linspace(-90, 20, 56), np.linspace(0, 40, 41)
    fm = flowmap_grid_2D(f, 20.0, -12.0, x, y, [-1.0])
    L = oracle.Linear(grid, U, V, spherical=1)
    _, st, _ = oracle.flowmap_grid_2d("linear", np.array([-1.0]), 20.0, -12.0, x, y, spline=L)
    assert (st == 1).all()
    # the trilinear RHS has kinks at cell faces, where the step controller's
    # accept/reject decisions follow rounding: the conditioning bar applies
    ref, sens = ic_sensitivity(lambda xx, yy: oracle.flowmap_grid_2d(
        "linear", np.array([-1.0]), 20.0, -12.0, xx, yy, spline=L)[0], x, y)
    ok, frac, err = flowmap_agreement(fm, ref, sens)
    assert ok, float(err.max())
    assert frac > 0.95, frac  # measured 97.5% within the plain bar
    f32 = _lib.to_host(flowmap_grid_2d_device(
        f, 20.0, -12.0, _lib.to_dev(x), _lib.to_dev(y), [-1.0], precision=_lib.NBX_FP32_DISP))
    assert np.mean(np.abs(f32 - ref).max(-1) <= 1e-3) > 0.95


@pytest.mark.parametrize("shape", [(5, 7, 11), (1, 1, 1), (2, 32, 64), (3, 70, 45),
                                   (4, 33, 97)])
def test_ingest_xy(nb, shape):
    """'xy' -> 'ij' transpose with the m/s -> km/h factor, exactly as
    np.moveaxis(u, 1, 2) * 3.6 (plot_merra_ftle.py:37-41), on shapes below,
    at and across the kernel's 32 x 32 tiles."""
    from numbacs_amd import _lib
    from numbacs_amd.flows import ingest_xy

    rng = np.random.default_rng(1)
    u = rng.standard_normal(shape)
    got = _lib.to_host(ingest_xy(u, 3.6))
    assert np.array_equal(got, np.moveaxis(u, 1, 2) * 3.6)


@pytest.mark.parametrize("precision", [0, 1])
def test_particles_stop_at_grid_edge(nb, precision):
    """Zero velocity outside the grid ('constant' extrapolation): in a
    uniform westward field every particle runs to x = 0 and stops there,
    x(T) = max(x0 - 0.5 T, 0) up to the step control's resolution of the
    velocity jump, with a bounded number of steps in both precision forms
    (a float displacement from the start point once stalled here, its
    increments falling below float resolution as the steps shrank)."""
    from numbacs_amd import _lib
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import flowmap_grid_2d_device

    t, x, y = np.linspace(0, 48, 25), np.linspace(0, 10, 21), np.linspace(0, 5, 11)
    U = np.full((25, 21, 11), -0.5)
    V = np.zeros((25, 21, 11))
    grid, Cu, Cv = get_interp_arrays_2D(t, x, y, U, V)
    f = get_flow_2D(grid, Cu, Cv)
    x0, y0 = np.linspace(0.3, 9.7, 95), np.linspace(0.5, 4.5, 9)
    st = _lib.empty((95, 9), "i32")
    ns = _lib.empty((95, 9), "i32")
    fm = _lib.to_host(flowmap_grid_2d_device(f, 0.0, 40.0, _lib.to_dev(x0), _lib.to_dev(y0),
                                             [1.0], precision=precision, status=st, nsteps=ns))
    assert (_lib.to_host(st) == 1).all()
    assert _lib.to_host(ns).max() < 1000
    want = np.maximum(x0 - 0.5 * 40.0, 0.0)
    assert np.abs(fm[..., 0] - want[:, None]).max() < 1e-3
    assert np.abs(fm[..., 1] - y0[None, :]).max() < 1e-12


# --------------------------------------------- spherical conversion --


@pytest.mark.parametrize("spherical", [1, 2])
@pytest.mark.parametrize("a", [0.0, 0.5])
@pytest.mark.parametrize("T", [24.0, -48.0])
@pytest.mark.parametrize("precision", [0, 1])
def test_spherical_conversion_closed_form(nb, spherical, a, T, precision):
    """flows.py:154-239 on the GPU: the trajectories of tests/spherical_cases.py
    (rhumb lines of a uniform field; a zonal profile U0 (1 + a sin lon)
    crossing the +-180 / 0-360 wrap) against their closed form.  FP64 to the
    oracle's own bar (rtol 1e-10); the FP32-displacement form (float spline
    and conversion; float stages cannot resolve rtol 1e-10, so it runs at the
    default rtol 1e-6) to 1e-5 relative in the lon integral (1e-4 deg for the
    uniform field) and 1e-5 deg in lat."""
    import spherical_cases as SC
    from numbacs_amd import _lib
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import flowmap_grid_2d_device

    V0 = 20.0
    t, lon, lat, U, V = SC.field(spherical, a, 40.0, V0)
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    f = get_flow_2D(grid, Cu, Cv, spherical=spherical)
    lon0, lat0 = SC.particles(spherical)
    st = _lib.empty((len(lon0), len(lat0)), "i32")
    tol = dict(rtol=1e-10, atol=1e-12) if precision == 0 else dict(rtol=1e-6, atol=1e-8)
    out = flowmap_grid_2d_device(f, SC.T0, T, _lib.to_dev(lon0), _lib.to_dev(lat0),
                                 np.array([math.copysign(1.0, T)]), precision=precision,
                                 status=st, **tol)
    fm = _lib.to_host(out)
    assert (_lib.to_host(st) == 1).all()
    if precision == 0:
        SC.check(fm, spherical, a, T, 40.0, V0, 1e-8, 1e-9)
    else:
        SC.check(fm, spherical, a, T, 40.0, V0, 1e-5 if a else 1e-4, 1e-5)
