"""Aux-grid flow map, its Cauchy-Green stencils and flow-map composition on
the GPU, against the reference fixtures and the oracle.

Tolerances: the double gyre over T = 8 (or 2) is not chaotic at these grid
sizes, so GPU and oracle flow maps agree to the integrator's rounding
(relative 1e-10); the Cauchy-Green stencils are a few flops on identical
inputs (relative 1e-12); the composition kernel evaluates the bilinear
interpolation in the oracle's source order with contraction off, so it is
compared bit for bit.
"""
import numpy as np
import pytest

import oracle
from conftest import apply_mask, reconstruct

pytestmark = pytest.mark.gpu

DG = "double_gyre"


def dg_params(direction=1.0):
    return np.array([direction, 0.1, 0.25, 0.0, 0.2 * np.pi, 0.0])


@pytest.fixture(scope="module")
def lib():
    from numbacs_amd import _lib

    _lib.torch()
    return _lib


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_flowmap_aux_fixture(lib, coords_dg, fm_aux_data, mask_dg):
    """tests/test_integration.py:66-75 of the reference."""
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_aux_grid_2D

    f, p, _ = get_predefined_flow("double_gyre")
    x, y = coords_dg
    fa = flowmap_aux_grid_2D(f, 0.0, 8.0, x, y, p)
    assert fa.shape == (21, 11, 5, 2)
    assert np.allclose(fa.astype(np.float32), fm_aux_data)
    fam = flowmap_aux_grid_2D(f, 0.0, 8.0, x, y, p, mask=mask_dg)
    assert np.allclose(fam.astype(np.float32), apply_mask(fm_aux_data, mask_dg))
    ref = oracle.flowmap_aux_grid_2d(DG, dg_params(), 0.0, 8.0, x, y)
    assert _rel(fa, ref) < 1e-10


@pytest.mark.parametrize("eig_main", [True, False])
@pytest.mark.parametrize("compute_edge", [True, False])
def test_flowmap_aux_edge_rules_vs_oracle(lib, eig_main, compute_edge):
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_aux_grid_2D

    f, p, _ = get_predefined_flow("double_gyre", int_direction=-1.0)
    x, y = np.linspace(0, 2, 33), np.linspace(0, 1, 17)
    mask = np.zeros((33, 17), bool)
    mask[5:9, 3:6] = True
    fa = flowmap_aux_grid_2D(f, 1.0, -2.0, x, y, p, h=1e-3, eig_main=eig_main,
                             compute_edge=compute_edge, mask=mask)
    ref = oracle.flowmap_aux_grid_2d(DG, p, 1.0, -2.0, x, y, h=1e-3, eig_main=eig_main,
                                     compute_edge=compute_edge, mask=mask)
    assert fa.shape == ref.shape == (33, 17, 5 if eig_main else 4, 2)
    assert np.array_equal(fa == 0, ref == 0)
    assert _rel(fa, ref) < 1e-10


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_flowmap_aux_spline_lane_groups(lib, lanes):
    """The aux-grid particle source under every lane-group shape of the
    spline gathers: the point-list flow map of the same starts, to rounding.
    The two entry points are separate kernel instantiations whose FMA
    contraction may differ, so a few trajectories differ in the last bits
    (measured 1.7e-13 / 2.7e-13 of a scale of 88 for 2 / 4 lanes); a
    misplaced start would differ by the offset h = 0.05 and more."""
    from numbacs_amd import autotune
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import flowmap, flowmap_aux_grid_2D
    from numbacs_amd.workloads import synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity(nt=30)
    grid, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    f = get_flow_2D(grid, Cu, Cv, spherical=1)
    saved = autotune.get(2, 0)
    try:
        autotune.set_config(2, 0, (lanes, 64, 0))
        x, y = np.linspace(-90, 20, 23), np.linspace(0, 40, 19)
        h = 0.05
        fa = flowmap_aux_grid_2D(f, 20.0, -12.0, x, y, [-1.0], h=h)
        off = np.array([[h, 0], [-h, 0], [0, h], [0, -h], [0, 0]])
        X, Y = np.meshgrid(x, y, indexing="ij")
        starts = np.stack([X, Y], -1)[:, :, None, :] + off[None, None]
        inner = starts[1:-1, 1:-1].reshape(-1, 2)
        pts = flowmap(f, 20.0, -12.0, inner, [-1.0])
        got = fa[1:-1, 1:-1].reshape(-1, 2)
        assert np.abs(got - pts).max() <= 1e-12 * np.abs(pts).max()
        assert np.mean(np.all(got == pts, axis=-1)) > 0.99
        assert not fa[0, :, :4].any() and fa[0, :, 4].all()
    finally:
        autotune.set_config(2, 0, saved)


def test_C_tensor_fixture(lib, coords_dg, fm_aux_data, C_aux_data, mask_dg):
    """tests/test_diagnostics.py:99-110 of the reference."""
    from numbacs_amd.diagnostics import C_tensor_2D

    x, y = coords_dg
    C = C_tensor_2D(fm_aux_data, x[1], y[1])
    assert np.allclose(C.astype(np.float32), C_aux_data)
    Cm = C_tensor_2D(fm_aux_data, x[1], y[1], mask=mask_dg)
    assert np.allclose(Cm.astype(np.float32), apply_mask(C_aux_data, mask_dg))
    assert _rel(C, oracle.C_tensor_aux_2d(fm_aux_data)) < 1e-12


def test_C_eig_aux_fixture(lib, coords_dg, fm_aux_data, C_eig_aux_data, mask_dg):
    """tests/test_diagnostics.py:112-131 of the reference (compared as
    V diag(l) V^T), and against the oracle for both eig_main modes."""
    from numbacs_amd.diagnostics import C_eig_aux_2D

    x, y = coords_dg
    ev, V = C_eig_aux_2D(fm_aux_data, x[1], y[1])
    want = reconstruct(*C_eig_aux_data)
    assert np.allclose(reconstruct(ev.astype(np.float32), V.astype(np.float32)), want)
    evm, Vm = C_eig_aux_2D(fm_aux_data, x[1], y[1], mask=mask_dg)
    assert np.allclose(reconstruct(evm.astype(np.float32), Vm.astype(np.float32)),
                       apply_mask(want, mask_dg))
    for eig_main in (True, False):
        F = fm_aux_data if eig_main else fm_aux_data[:, :, :4].copy()
        ev, V = C_eig_aux_2D(F, x[1], y[1], eig_main=eig_main)
        er, Vr = oracle.C_eig_aux_2d(F, x[1], y[1], eig_main=eig_main)
        assert _rel(ev, er) < 1e-12
        assert np.abs(np.abs(V) - np.abs(Vr)).max() < 1e-10


def test_composition_kernel_bit_exact(lib, coords_dg, composition_data):
    """The composition launch against the oracle on the reference's maps and
    on maps that leave the box, every ring rotation, nT = 1 .. 8."""
    from numbacs_amd.integration import flowmap_composition, flowmap_composition_device

    x, y = coords_dg
    grid = ((0.0, 2.0, 21), (0.0, 1.0, 11))
    fms = np.stack([oracle.flowmap_grid_2d(DG, dg_params(), float(k), 1.0, x, y)[0]
                    for k in range(8)])
    rng = np.random.default_rng(3)
    noisy = fms + rng.normal(0, 0.05, fms.shape)
    for maps in (fms, noisy):
        for nT in (1, 2, 3, 8):
            got = flowmap_composition(maps[:nT], grid, nT)
            assert np.array_equal(got, oracle.flowmap_composition(maps[:nT], grid, nT)), nT
        dev = lib.to_dev(maps)
        for first in range(8):
            chain = np.roll(maps, -first, axis=0)
            got = lib.to_host(flowmap_composition_device(dev, grid, 8, first))
            assert np.array_equal(got, oracle.flowmap_composition(chain, grid, 8)), first


def test_composition_fixtures(lib, coords_dg, composition_data):
    """tests/test_integration.py:92-114 of the reference, through the
    drop-in functions."""
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_composition_initial, flowmap_composition_step

    f, p, _ = get_predefined_flow("double_gyre")
    x, y = coords_dg
    grid = ((x[0], x[-1], 21), (y[0], y[-1], 11))
    flowmap0, flowmaps, nT = flowmap_composition_initial(f, 0.0, 8.0, 1.0, x, y, grid, p)
    assert nT == 8
    assert np.allclose(flowmap0.astype(np.float32), composition_data["fm_ci"])
    assert np.allclose(flowmaps.astype(np.float32), composition_data["fms_ci"])
    fms = composition_data["fms_ci"].astype(np.float64)
    flowmap_k, fms2 = flowmap_composition_step(fms, f, 8.0, 1.0, 8, x, y, grid, p)
    assert fms2 is fms
    assert np.allclose(flowmap_k.astype(np.float32), composition_data["fm_cs"])
    assert np.allclose(fms2.astype(np.float32), composition_data["fms_cs"])


def test_composer_ring_matches_shifting(lib):
    """Five steps of the device ring equal the reference's shift-and-append
    window on a spline flow (same maps, same composition)."""
    from numbacs_amd.flows import get_flow_2D, get_interp_arrays_2D
    from numbacs_amd.integration import FlowmapComposer, flowmap_composition, flowmap_grid_2D
    from numbacs_amd.workloads import synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity(nt=40)
    grid_v, Cu, Cv = get_interp_arrays_2D(t, lon, lat, U, V)
    f = get_flow_2D(grid_v, Cu, Cv, spherical=1)
    x, y = np.linspace(-90, 20, 45), np.linspace(0, 40, 33)
    grid = ((x[0], x[-1], 45), (y[0], y[-1], 33))
    c = FlowmapComposer(f, x, y, -3.0, 4, [-1.0]).initial(30.0)
    maps = np.stack([flowmap_grid_2D(f, 30.0 - 3.0 * k, -3.0, x, y, [-1.0]) for k in range(4)])
    assert np.array_equal(lib.to_host(c.flowmaps()), maps)
    for s in range(5):
        t0 = 30.0 - 3.0 * (4 + s)
        c.step(t0)
        maps = np.concatenate([maps[1:], flowmap_grid_2D(f, t0, -3.0, x, y, [-1.0])[None]])
        assert np.array_equal(lib.to_host(c.flowmaps()), maps)
        assert np.array_equal(lib.to_host(c.compose()), flowmap_composition(maps, grid, 4))
