"""Parity of the HIP path with the CPU oracle at the benchmark workloads'
own sizes.

* gridded velocity (676x251 particles on the 64-mode field, spherical=1,
  T=-72 h, FTLE and p90 ridges): flow map, FTLE and ridge sets;
* FTLE time series (48 uneven requests over the 480-frame field in one
  batched launch), FP64 and FP32-displacement;
* precision study (double gyre 2001x1001 at T=-10/-20, Bickley 2001x801 at
  T=-40 d) in the FP32-displacement form against the FP64 oracle;
* surrogate with integrator fallback on the 2701x1001 grid.

Bars: flow maps within the looser of 1e-6 relative and the trajectory's own
conditioning (conftest.flowmap_agreement); FTLE within 1e-6 relative (of the
field's maximum) or within the bound the flow-map difference implies
(conftest.ftle_bound); ridge points bit-equal when the inputs are equal.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import flowmap_agreement, ftle_bound, ic_sensitivity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from numbacs_amd import _lib

    _lib.torch()
    return _lib


def _flush(msg):
    print(msg, flush=True)


# -------------------------------------------------- gridded velocity (C2) --


@pytest.fixture(scope="module")
def gridded(lib):
    """Full gridded workload on the GPU (FP64) and in the oracle."""
    from numbacs_amd.diagnostics import ftle_eig_2d_device, ftle_grid_2D
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import flowmap_grid_2d_device
    from numbacs_amd.workloads import particle_axes, synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity()
    xp, yp = particle_axes()
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    flow = get_flow_2D(grid, Cu, Cv, spherical=1)
    p = np.array([-1.0])
    st = lib.empty((len(xp), len(yp)), "i32")
    fmd = flowmap_grid_2d_device(flow, 72.0, -72.0, lib.to_dev(xp), lib.to_dev(yp), p,
                                 status=st)
    dx, dy = xp[1] - xp[0], yp[1] - yp[0]
    fd, evd = ftle_eig_2d_device(fmd, -72.0, dx, dy)
    S = oracle.Spline(grid, Cu, Cv, spherical=1)
    run = lambda xx, yy: oracle.flowmap_grid_2d("spline", p, 72.0, -72.0, xx, yy, spline=S)[0]
    ref, sens = ic_sensitivity(run, xp, yp)
    fm = lib.to_host(fmd)
    return dict(xp=xp, yp=yp, dx=dx, dy=dy, fm=fm, st=lib.to_host(st), ref=ref, sens=sens,
                ftle=lib.to_host(fd), evec=lib.to_host(evd),
                ftle_api=ftle_grid_2D(fm, -72.0, dx, dy))


def test_gridded_flowmap_full(gridded):
    g = gridded
    assert (g["st"] == 1).all()
    ok, frac, err = flowmap_agreement(g["fm"], g["ref"], g["sens"])
    _flush(f"gridded flow map: {frac:.5f} within 1e-6, max {err.max():.3e}")
    assert ok, float(err.max())
    assert frac > 0.999, frac


def test_gridded_ftle_full(gridded):
    """FTLE at every point within 1e-6 of the field's maximum, or within
    what the measured flow-map difference implies there; the fused eigen
    kernel and ftle_grid_2D give the same field; on the oracle's flow map
    the GPU stencil equals the oracle's to rounding."""
    g = gridded
    T, dx, dy = -72.0, g["dx"], g["dy"]
    fref = oracle.ftle_grid_2d(g["ref"], T, dx, dy)
    scale = np.abs(fref).max()
    err = np.abs(g["ftle"] - fref)
    bound = ftle_bound(g["fm"], g["ref"], T, dx, dy)
    within = err <= 1e-6 * scale
    _flush(f"gridded FTLE: {within.mean():.5f} within 1e-6, max rel {err.max() / scale:.3e}, "
           f"{(~within).sum()} points held to the conditioning bound")
    assert np.all(within | (err <= bound))
    assert within.mean() > 0.99
    assert np.abs(g["ftle_api"] - g["ftle"]).max() <= 1e-12 * scale
    from numbacs_amd.diagnostics import ftle_grid_2D

    on_ref = ftle_grid_2D(g["ref"], T, dx, dy)
    assert np.abs(on_ref - fref).max() <= 1e-12 * scale


def test_gridded_ridges_full(gridded):
    """p90 ridges: on the oracle's FTLE field and eigenvectors the GPU ridge
    points and grouping equal the oracle's bit for bit; on the GPU's own
    field the flagged grid points differ from the oracle's in at most 1% of
    the ridge points, and each differing point is marginal: its FTLE lies
    within 1e-3 relative of the percentile threshold, its eigenvector
    differs from the oracle's, or the FTLE of a point within two grid steps
    differs from the oracle's by more than 1e-6 relative."""
    from numbacs_amd.extraction import ftle_ridge_pts, ftle_ridges

    g = gridded
    T, dx, dy, xp, yp = -72.0, g["dx"], g["dy"], g["xp"], g["yp"]
    ev_ref, V_ref = oracle.C_eig_2d(g["ref"], dx, dy)
    fref = oracle.ftle_grid_2d(g["ref"], T, dx, dy)
    eref = np.ascontiguousarray(V_ref[:, :, :, 1])
    want = oracle.ridge_pts(fref, eref, xp, yp, percentile_=90)
    got = ftle_ridge_pts(fref, eref, xp, yp, percentile=90)
    assert np.array_equal(got, want)
    rw = oracle.ridges(fref, eref, xp, yp, percentile_=90)
    rg = ftle_ridges(fref, eref, xp, yp, percentile=90)
    assert len(rg) == len(rw) and all(np.array_equal(a, b) for a, b in zip(rg, rw))

    _, flag_ref = oracle.ridge_predicate(fref, eref, xp, yp, percentile_=90)
    _, flag_gpu = oracle.ridge_predicate(g["ftle"], g["evec"], xp, yp, percentile_=90)
    own = ftle_ridge_pts(g["ftle"], g["evec"], xp, yp, percentile=90)
    assert len(own) == int(flag_gpu.sum())
    diff = flag_ref ^ flag_gpu
    nd = int(diff.sum())
    _flush(f"gridded ridges: {int(flag_ref.sum())} oracle / {int(flag_gpu.sum())} GPU ridge "
           f"points, {nd} differ; {len(rw)} oracle ridges, "
           f"{len(ftle_ridges(g['ftle'], g['evec'], xp, yp, percentile=90))} GPU")
    assert nd <= 0.01 * flag_ref.sum()
    scale = np.abs(fref).max()
    thr = oracle.percentile(fref, 90)
    ferr = np.abs(g["ftle"] - fref) > 1e-6 * scale
    near = np.zeros_like(ferr)
    for di in range(-2, 3):
        for dj in range(-2, 3):
            near |= np.roll(np.roll(ferr, di, 0), dj, 1)
    ecos = np.abs((g["evec"] * eref).sum(-1))
    marginal = near | (np.abs(fref - thr) <= 1e-3 * scale) | (ecos < 1 - 1e-12)
    assert np.all(marginal[diff])


# ----------------------------------------------------- time series (C3) --


@pytest.fixture(scope="module")
def timeseries(lib):
    """The 48 time-series requests over the 480-frame field, batched in FP64
    and in the FP32-displacement form."""
    from numbacs_amd.batch import FTLEBatch
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.workloads import synthetic_velocity, time_series_requests

    t, lon, lat, U, V = synthetic_velocity(480)
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    flow = get_flow_2D(grid, Cu, Cv, spherical=1)
    reqs = time_series_requests()
    res = {}
    for prec in ("fp64", "fp32"):
        b = FTLEBatch(flow, reqs, [-1.0], precision=prec).run()
        res[prec] = b.results()
        b.close()
    S = oracle.Spline(grid, Cu, Cv, spherical=1)
    return dict(reqs=reqs, res=res, S=S)


def test_timeseries_fp64_every_request_vs_oracle(timeseries):
    """Every request's FP64 flow map against the oracle on every 32nd
    particle row, under the flow-map bar (the trajectory's conditioning
    measured by the oracle itself)."""
    reqs, res, S = timeseries["reqs"], timeseries["res"], timeseries["S"]
    p = np.array([-1.0])
    fracs = []
    for k, q in enumerate(reqs):
        rows = np.arange(k % 32, len(q["x"]), 32)
        run = lambda xx, yy: oracle.flowmap_grid_2d("spline", p, q["t0"], q["T"], xx, yy,
                                                    spline=S)[0]
        ref, sens = ic_sensitivity(run, q["x"][rows], q["y"])
        assert (res["fp64"][k]["status"] == 1).all()
        ok, frac, err = flowmap_agreement(res["fp64"][k]["flowmap"][rows], ref, sens)
        assert ok, (k, float(err.max()))
        fracs.append(frac)
    _flush(f"time series FP64: requests' fraction within 1e-6: min {min(fracs):.5f}, "
           f"median {np.median(fracs):.5f}")
    assert min(fracs) > 0.995


def test_timeseries_fp32_error_per_request(timeseries):
    """The FP32-displacement form against the batched FP64 result, FTLE
    relative to the request's largest FP64 FTLE.  Particles that end on the
    other side of the velocity box's edge than in FP64 (zero velocity
    outside: a rounding-level difference decides which side) are counted
    apart; away from them every request keeps 95% of its points within
    1e-3 and an RMS error below 2e-2 (measured: see the printed figures)."""
    reqs, res = timeseries["reqs"], timeseries["res"]
    rms, within, split = [], [], 0
    for k, q in enumerate(reqs):
        a, b = res["fp32"][k], res["fp64"][k]
        d = np.abs(a["flowmap"] - b["flowmap"]).max(-1)
        split += int((d > 1.0).sum())
        e = np.abs(a["ftle"] - b["ftle"]) / np.abs(b["ftle"]).max()
        rms.append(float(np.sqrt(np.mean(e ** 2))))
        within.append(float(np.mean(e <= 1e-3)))
    _flush(f"time series FP32 vs FP64: FTLE RMS median {np.median(rms):.2e} max {max(rms):.2e}; "
           f"within 1e-3 min {min(within):.4f} median {np.median(within):.4f}; "
           f"{split} particles split at the grid edge")
    assert max(rms) < 2e-2
    assert min(within) > 0.95


# ------------------------------------------------ precision study (C4) --


def _ridge_flags(f, ev, x, y):
    _, flag = oracle.ridge_predicate(f, ev, x, y, percentile_=90)
    return flag


@pytest.mark.parametrize("case", ["dg10", "dg20", "bickley"])
def test_precision_study_fp32_vs_fp64(lib, case):
    """FP32-displacement form (FP32 stages and velocity, FP64 position, step
    control, Cauchy-Green and eigenvalue) against the FP64 path: the FP64
    flow map matches the oracle on sampled rows, the FP32 FTLE error is held
    to its measured level, and the p90 ridge sets (grid points passing the
    ridge predicate) agree but for the stated fraction."""
    from numbacs_amd.diagnostics import ftle_eig_2d_device
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2d_device
    from numbacs_amd.workloads import bickley_grid, bickley_params, double_gyre_params

    if case == "bickley":
        name, T = "bickley_jet", -40.0
        p = bickley_params(-1.0)
        x, y = bickley_grid(2001, 801)
        # chaotic over 40 days: rounding is amplified as any perturbation
        # (measured: RMS 8.6e-3, 82% within 1e-3, ridge points differing 6.6%)
        bars = dict(rms=2e-2, within=0.75, ridge=0.1)
    else:
        name, T = "double_gyre", (-10.0 if case == "dg10" else -20.0)
        p = double_gyre_params(-1.0)
        x, y = np.linspace(0, 2, 2001), np.linspace(0, 1, 1001)
        # measured: T=-10 RMS 1.0e-4, 99.95% within 1e-3, ridge points
        # differing 4.5% (many sit on the sub-pixel test's edge at this
        # horizon); T=-20 RMS 2.3e-3, 99.76%, 0.55%
        bars = (dict(rms=2e-4, within=0.999, ridge=0.06) if case == "dg10" else
                dict(rms=5e-3, within=0.995, ridge=0.01))
    f = get_predefined_flow(name, return_default_params=False, return_domain=False)
    xd, yd = lib.to_dev(x), lib.to_dev(y)
    dx, dy = x[1] - x[0], y[1] - y[0]
    out = {}
    for prec in (0, 1):
        fm = flowmap_grid_2d_device(f, 0.0, T, xd, yd, p, precision=prec)
        ft, ev = ftle_eig_2d_device(fm, T, dx, dy)
        out[prec] = (lib.to_host(fm), lib.to_host(ft), lib.to_host(ev))
    rows = np.arange(7, len(x), 100)
    run = lambda xx, yy: oracle.flowmap_grid_2d(name, p, 0.0, T, xx, yy)[0]
    ref, sens = ic_sensitivity(run, x[rows], y)
    ok, frac, err = flowmap_agreement(out[0][0][rows], ref, sens)
    assert ok, float(err.max())
    e = np.abs(out[1][1] - out[0][1]) / np.abs(out[0][1]).max()
    rms, within = float(np.sqrt(np.mean(e ** 2))), float(np.mean(e <= 1e-3))
    f64, f32 = _ridge_flags(out[0][1], out[0][2], x, y), _ridge_flags(out[1][1], out[1][2], x, y)
    sym = int((f64 ^ f32).sum()) / max(int(f64.sum()), 1)
    _flush(f"precision {case}: FP64 vs oracle {frac:.5f} within 1e-6; FP32 FTLE max "
           f"{e.max():.2e} RMS {rms:.2e} within 1e-3 {within:.5f}; ridge points FP64 "
           f"{int(f64.sum())} FP32 {int(f32.sum())}, symmetric difference {sym:.4f}")
    assert rms < bars["rms"] and within > bars["within"] and sym < bars["ridge"]


# ----------------------------------------------- surrogate with fallback (C5) --


def test_surrogate_fallback_on_the_full_grid(lib):
    """flowmap_with_fallback on the 2701x1001 grid at one t0 (T = -72 h):
    with the 0.05 dx tolerance every tile whose probes miss falls back and
    those tiles equal the plain DOP853 flow map bit for bit; with a loose
    tolerance (2 degrees) the accepted tiles' probe particles lie within it
    of the FP64 oracle."""
    import torch

    from numbacs_amd import surrogate as S
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import flowmap_grid_2d_device
    from numbacs_amd.workloads import synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity()
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    flow = get_flow_2D(grid, Cu, Cv, spherical=1)
    box = ((-100.0, 35.0), (-5.0, 45.0))
    reqs = S.sample_requests(100, 100, box, (72.0, 95.0), (-72.0, -72.0), seed=5)
    inp, tgt = S.integrate_samples(flow, [-1.0], reqs)
    net = S.MLPSurrogate([-32.5, 20.0, 83.5, -72.0], [67.5, 25.0, 11.5, 1.0], [15.0, 15.0],
                         sigma=2.0, seed=7)
    S.train(net, inp, tgt, epochs=2, batch=8192, lr=2e-3)
    nx, ny = 2701, 1001
    x, y = np.linspace(-100, 35, nx), np.linspace(-5, 45, ny)
    dx = x[1] - x[0]
    t0 = 80.0
    plain = lib.to_host(flowmap_grid_2d_device(flow, t0, -72.0, lib.to_dev(x), lib.to_dev(y),
                                               [-1.0]))
    fm, info = S.flowmap_with_fallback(net, flow, [-1.0], x, y, t0, -72.0, 0.05 * dx)
    fm = lib.to_host(fm)
    _flush(f"surrogate 0.05 dx: {info['fallback_tiles']}/{info['tiles']} tiles fall back, "
           f"probe error median {info['probe_err_median']:.3g} deg")
    tiles = [(i0, j0) for i0 in range(0, nx, 64) for j0 in range(0, ny, 64)]
    bad = set(info["fallback"])
    for i0, j0 in tiles:
        a, b = fm[i0:i0 + 64, j0:j0 + 64], plain[i0:i0 + 64, j0:j0 + 64]
        if (i0, j0) in bad:
            assert np.array_equal(a, b)
    tol = 2.0
    fm2, info2 = S.flowmap_with_fallback(net, flow, [-1.0], x, y, t0, -72.0, tol)
    fm2 = lib.to_host(fm2)
    acc = [tl for tl in tiles if tl not in set(info2["fallback"])]
    _flush(f"surrogate 2 deg: {info2['fallback_tiles']}/{info2['tiles']} tiles fall back")
    assert 0 < len(acc) < len(tiles)
    P = S.PROBES
    pts = np.array([(x[min(i0 + pi, nx - 1)], y[min(j0 + pj, ny - 1)])
                    for i0, j0 in acc for pi, pj in P])
    Sp = oracle.Spline(grid, Cu, Cv, spherical=1)
    ref, st = oracle.flowmap_pts("spline", np.array([-1.0]), t0, -72.0, pts, spline=Sp)
    got = np.array([fm2[min(i0 + pi, nx - 1), min(j0 + pj, ny - 1)] for i0, j0 in acc
                    for pi, pj in P])
    # the probes were accepted against the GPU integrator; the oracle agrees
    # with it to the flow-map bar, far below the tolerance
    assert np.abs(got - ref).max() <= tol * (1 + 1e-6)
    torch.cuda.synchronize()
