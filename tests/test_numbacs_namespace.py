"""The ``numbacs`` import path: the reference's own import lines resolve to
the GPU implementations (CPU test), and the double-gyre FTLE example
(examples/ftle/plot_dg_ftle.py:14-48 of the reference, plotting left out)
runs through them on the GPU against the oracle."""
import math

import numpy as np
import pytest

import oracle


def test_reference_import_lines_resolve():
    import numbacs_amd
    from numbacs.diagnostics import C_eig_2D, ftle_from_eig, ftle_grid_2D
    from numbacs.extraction import ftle_ordered_ridges, ftle_ridge_pts, ftle_ridges
    from numbacs.flows import get_flow_2D, get_interp_arrays_2D, get_predefined_flow
    from numbacs.integration import flowmap, flowmap_grid_2D

    assert flowmap_grid_2D is numbacs_amd.integration.flowmap_grid_2D
    assert flowmap is numbacs_amd.integration.flowmap
    assert ftle_grid_2D is numbacs_amd.diagnostics.ftle_grid_2D
    assert C_eig_2D is numbacs_amd.diagnostics.C_eig_2D
    assert ftle_from_eig is numbacs_amd.diagnostics.ftle_from_eig
    assert get_predefined_flow is numbacs_amd.flows.get_predefined_flow
    assert get_interp_arrays_2D is numbacs_amd.flows.get_interp_arrays_2D
    assert get_flow_2D is numbacs_amd.flows.get_flow_2D
    assert ftle_ridges is numbacs_amd.extraction.ftle_ridges
    assert ftle_ridge_pts is numbacs_amd.extraction.ftle_ridge_pts
    assert ftle_ordered_ridges is numbacs_amd.extraction.ftle_ordered_ridges


@pytest.mark.parametrize("stmt", ["from numbacs.flows import get_predefined_callable",
                                  "from numbacs.integration import flowmap_n",
                                  "from numbacs.diagnostics import lavd_grid_2D",
                                  "from numbacs.extraction import hyperbolic_lcs"])
def test_names_outside_the_path_raise_import_error(stmt):
    with pytest.raises(ImportError, match="outside the FTLE path"):
        exec(stmt, {})


@pytest.mark.gpu
def test_double_gyre_example_through_numbacs_imports():
    from numbacs.diagnostics import ftle_grid_2D
    from numbacs.flows import get_predefined_flow
    from numbacs.integration import flowmap_grid_2D

    t0 = 0.0
    T = -10.0
    int_direction = math.copysign(1, T)
    funcptr, params, domain = get_predefined_flow("double_gyre", int_direction=int_direction)
    nx, ny = 401, 201
    x = np.linspace(domain[0][0], domain[0][1], nx)
    y = np.linspace(domain[1][0], domain[1][1], ny)
    dx = x[1] - x[0]
    dy = y[1] - y[0]
    flowmap = flowmap_grid_2D(funcptr, t0, T, x, y, params)
    ftle = ftle_grid_2D(flowmap, T, dx, dy)
    assert flowmap.shape == (nx, ny, 2) and ftle.shape == (nx, ny)
    assert flowmap.dtype == np.float64 and ftle.dtype == np.float64
    ref, st, _ = oracle.flowmap_grid_2d("double_gyre", params, t0, T, x, y)
    assert (st == 1).all()
    assert np.abs(flowmap - ref).max() <= 1e-7 * np.abs(ref).max()
    fref = oracle.ftle_grid_2d(ref, T, dx, dy)
    assert np.abs(ftle - fref).max() <= 1e-6 * np.abs(fref).max()
