"""Batched time series (persistent tile-queue launch, batched FTLE, HIP
graph replay), the persistent single-grid launch and the FP32-displacement
form, all through the C ABI.

The queue kernel runs the same per-particle solver as the one-group-per-
particle kernel -- a tile handed back after a chunk of steps resumes from its
stored solver states exactly -- so with the same lanes-per-particle setting
its flow maps must be bit-identical to flowmap_grid_2D's; FTLE fields from
the batch must be bit-identical to ftle_grid_2D / the fused eigen stencil on
those maps.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from numbacs_amd import _lib

    _lib.torch()
    return _lib


@pytest.fixture
def restore_config(lib):
    from numbacs_amd import autotune

    saved = {(k, p): autotune.get(k, p) for k in (0, 1, 2) for p in (0, 1)}
    orders = {p: autotune.get_order(p) for p in (0, 1)}
    yield autotune
    for (k, p), cfg in saved.items():
        autotune.set_config(k, p, cfg)
    for p, o in orders.items():
        autotune.set_order(p, o)


def _dg_requests():
    return [dict(t0=0.0, T=-10.0, x=np.linspace(0, 2, 101), y=np.linspace(0, 1, 51)),
            dict(t0=2.5, T=-5.0, x=np.linspace(0, 2, 37), y=np.linspace(0, 1, 73)),
            dict(t0=1.0, T=8.0, x=np.linspace(0.2, 1.8, 64), y=np.linspace(0.1, 0.9, 1)),
            dict(t0=0.0, T=-3.0, x=np.linspace(0, 2, 5), y=np.linspace(0, 1, 3))]


@pytest.mark.parametrize("block,persist", [(64, 2), (256, 8)])
def test_batch_matches_single_grid_double_gyre(lib, restore_config, block, persist):
    from numbacs_amd.batch import FTLEBatch
    from numbacs_amd.diagnostics import ftle_grid_2D
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2D

    restore_config.set_config(0, 0, (1, block, persist))
    f, p, _ = get_predefined_flow("double_gyre", int_direction=-1.0)
    reqs = _dg_requests()
    b = FTLEBatch(f, reqs, p).run()
    res = b.results()
    assert b.particles == sum(len(q["x"]) * len(q["y"]) for q in reqs)
    for q, r in zip(reqs, res):
        fm, st = flowmap_grid_2D(f, q["t0"], q["T"], q["x"], q["y"], p, return_status=True)
        assert np.array_equal(r["flowmap"], fm)
        assert np.array_equal(r["status"], st)
        dx = q["x"][1] - q["x"][0]
        dy = q["y"][1] - q["y"][0] if len(q["y"]) > 1 else 1.0
        assert np.array_equal(r["ftle"], ftle_grid_2D(fm, q["T"], dx, dy))
    b.close()


def test_batch_matches_oracle_double_gyre(lib):
    from numbacs_amd.batch import FTLEBatch
    from numbacs_amd.flows import get_predefined_flow

    f, p, _ = get_predefined_flow("double_gyre", int_direction=-1.0)
    reqs = _dg_requests()[:2]
    res = FTLEBatch(f, reqs, p).run().results()
    for q, r in zip(reqs, res):
        ref, _, _ = oracle.flowmap_grid_2d("double_gyre", p, q["t0"], q["T"], q["x"], q["y"])
        assert np.abs(r["flowmap"] - ref).max() / np.abs(ref).max() < 1e-6


@pytest.fixture(scope="module")
def small_field(lib):
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.workloads import synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity(nt=40)
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    return grid, Cu, Cv, get_flow_2D(grid, Cu, Cv, spherical=1)


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_batch_matches_single_grid_spline(lib, restore_config, small_field, lanes):
    from numbacs_amd.batch import FTLEBatch
    from numbacs_amd.diagnostics import C_eig_2D, ftle_from_eig
    from numbacs_amd.integration import flowmap_grid_2D

    restore_config.set_config(2, 0, (lanes, 64, 4))
    grid, Cu, Cv, f = small_field
    reqs = [dict(t0=30.0 + k, T=(-6.0, -12.0, -24.0)[k % 3], x=np.linspace(-100, 35, 60 + 7 * k),
                 y=np.linspace(-5, 45, 25 + 3 * k)) for k in range(5)]
    reqs[1]["mask"] = np.zeros((len(reqs[1]["x"]), len(reqs[1]["y"])), bool)
    reqs[1]["mask"][::3, ::2] = True
    b = FTLEBatch(f, reqs, [-1.0], eigvec=True).run()
    for q, r in zip(reqs, b.results()):
        fm = flowmap_grid_2D(f, q["t0"], q["T"], q["x"], q["y"], [-1.0], mask=q.get("mask"))
        assert np.array_equal(r["flowmap"], fm)
        dx, dy = q["x"][1] - q["x"][0], q["y"][1] - q["y"][0]
        ev, V = C_eig_2D(fm, dx, dy, mask=q.get("mask"))
        assert np.array_equal(r["eigvec_max"], V[:, :, :, 1])
        assert np.allclose(r["ftle"], ftle_from_eig(ev[:, :, 1], q["T"]), rtol=1e-14, atol=0)
    b.close()


@pytest.mark.parametrize("cfg", [(4, 64, 0), (4, 64, 4)])
def test_graph_replays_match_eager(lib, restore_config, small_field, cfg):
    """The captured graph replayed several times, with eager work and
    synchronisation between replays, reproduces the eager batch exactly:
    the tile launch, and the tile queue (whose ring is reset by a kernel in
    the graph and whose tiles go round in chunks)."""
    import torch

    from numbacs_amd.batch import FTLEBatch
    from numbacs_amd.integration import flowmap_grid_2D

    restore_config.set_config(2, 0, cfg)
    grid, Cu, Cv, f = small_field
    reqs = [dict(t0=20.0 + 2 * k, T=-12.0, x=np.linspace(-100, 35, 50), y=np.linspace(-5, 45, 31))
            for k in range(4)]
    b = FTLEBatch(f, reqs, [-1.0]).run()
    eager = b.results()
    b.capture()
    for rep in range(3):
        for d in b.requests:
            d["flowmap"].fill_(0.0)
            d["ftle"].fill_(-1.0)
        torch.cuda.synchronize()
        b.run_graph()
        torch.cuda.synchronize()
        got = b.results()
        for e, g in zip(eager, got):
            assert np.array_equal(e["flowmap"], g["flowmap"]), rep
            assert np.array_equal(e["ftle"], g["ftle"]), rep
        # unrelated eager work between replays
        flowmap_grid_2D(f, 25.0, -6.0, np.linspace(-90, 20, 17), np.linspace(0, 40, 9), [-1.0])
    b.close()


@pytest.mark.parametrize("lanes,block,persist", [(2, 64, 4), (4, 128, 2), (1, 256, 8)])
def test_persistent_single_grid_matches(lib, restore_config, small_field, lanes, block, persist):
    from numbacs_amd.integration import flowmap_grid_2D

    grid, Cu, Cv, f = small_field
    x, y = np.linspace(-100, 35, 140), np.linspace(-5, 45, 61)
    restore_config.set_config(2, 0, (lanes, block, 0))
    a, sa = flowmap_grid_2D(f, 35.0, -30.0, x, y, [-1.0], return_status=True)
    restore_config.set_config(2, 0, (lanes, block, persist))
    b, sb = flowmap_grid_2D(f, 35.0, -30.0, x, y, [-1.0], return_status=True)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)


def test_fp32_displacement_error_double_gyre(lib):
    """FP32-displacement form against the FP64 oracle on the double gyre
    (401x201, T=-10).  Measured: 99.2% of endpoints within 1e-4, FTLE RMS
    relative error 4.4e-5 and 99.96% of points within 1e-3; the largest
    errors sit on the chaotic saddle where float rounding is amplified."""
    from numbacs_amd.diagnostics import ftle_grid_2d_device
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2d_device

    f, p, _ = get_predefined_flow("double_gyre", int_direction=-1.0)
    x, y = np.linspace(0, 2, 401), np.linspace(0, 1, 201)
    xd, yd = lib.to_dev(x), lib.to_dev(y)
    fm = flowmap_grid_2d_device(f, 0.0, -10.0, xd, yd, p, precision=lib.NBX_FP32_DISP)
    ft = lib.to_host(ftle_grid_2d_device(fm, -10.0, 0.005, 0.005))
    ref, _, _ = oracle.flowmap_grid_2d("double_gyre", p, 0.0, -10.0, x, y)
    ftr = oracle.ftle_grid_2d(ref, -10.0, 0.005, 0.005)
    perr = np.abs(lib.to_host(fm) - ref).max(-1)
    assert np.mean(perr <= 1e-4) > 0.985
    err = np.abs(ft - ftr) / np.abs(ftr).max()
    assert np.sqrt(np.mean(err ** 2)) < 1e-4
    assert np.mean(err <= 1e-3) > 0.999


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["gridded", "wrap_nearest_mask"])
def test_staged_and_queue_launches_equal_grid_launch(lib, restore_config, precision, case):
    """The grid launch in edge-first tile order (tiles nearest the velocity
    box's edge launched first) and in plain tile order, the LDS-staged
    launch (coefficient boxes copied to LDS per step, persistent -1) and the
    persistent tile queue (tiles handed back every 16 steps with their
    solver states stored, persistent n > 0) run the same arithmetic: equal
    flow maps, status and step counts, on the gridded workload and on a
    global field whose boxes straddle the longitude wrap, with nearest
    extrapolation and a mask."""
    from numbacs_amd import autotune
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.workloads import particle_axes, synthetic_velocity

    if case == "gridded":
        t, lon, lat, U, V = synthetic_velocity()
        grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
        f = get_flow_2D(grid, Cu, Cv, spherical=1)
        x, y = particle_axes()
        mask, t0, T = None, 72.0, -72.0
    else:
        t = np.linspace(0.0, 48.0, 25)
        lon = np.linspace(-180.0, 179.0, 360)
        lat = np.linspace(-60.0, 60.0, 121)
        TT, LON, LAT = np.meshgrid(t, lon, lat, indexing="ij")
        U = 60.0 + 20.0 * np.sin(np.deg2rad(3 * LON) + 0.1 * TT) * np.cos(np.deg2rad(2 * LAT))
        V = 15.0 * np.cos(np.deg2rad(2 * LON) - 0.05 * TT)
        grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
        f = get_flow_2D(grid, Cu, Cv, spherical=1, extrap_mode="nearest")
        x, y = np.linspace(150.0, 200.0, 121), np.linspace(-50.0, 50.0, 83)
        mask = np.zeros((len(x), len(y)), bool)
        mask[::7, ::5] = True
        t0, T = 40.0, -36.0
    xd, yd = lib.to_dev(x), lib.to_dev(y)
    md = None if mask is None else lib.to_dev(mask.astype(np.uint8), np.uint8)
    # launches that must agree bit for bit: the 4-lane forms, then the
    # 2-lane form (its own summation order) in both tile orders
    groups = [(((4, 64, 0), 0), ((4, 64, 0), 1), ((4, 256, 0), 1), ((4, 64, -1), 1),
               ((4, 64, 4), 1), ((4, 256, 2), 1)),
              (((2, 128, 0), 0), ((2, 128, 0), 1))]
    for group in groups:
        _equal_launches(lib, autotune, precision, group, f, t0, T, xd, yd, md, x, y)


def _equal_launches(lib, autotune, precision, group, f, t0, T, xd, yd, md, x, y):
    from numbacs_amd.integration import flowmap_grid_2d_device

    res = []
    for cfg, edge_first in group:
        autotune.set_config(lib.NBX_FLOW_SPLINE3, precision, cfg)
        autotune.set_order(precision, edge_first)
        st = lib.empty((len(x), len(y)), "i32")
        ns = lib.empty((len(x), len(y)), "i32")
        fm = flowmap_grid_2d_device(f, t0, T, xd, yd, [-1.0], maskd=md, precision=precision,
                                    status=st, nsteps=ns)
        res.append((lib.to_host(fm), lib.to_host(st), lib.to_host(ns)))
    assert (res[0][1] != 0).any()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert np.array_equal(a, b)
