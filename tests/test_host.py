"""Host-side behaviour of the Python entry points that needs no GPU work:
return-tuple arity, parameter defaults and the documented errors."""
import math

import numpy as np
import pytest

from numbacs_amd.flows import FlowHandle, get_predefined_flow
from numbacs_amd.integration import flowmap, flowmap_grid_2D


def test_predefined_flow_arity_matrix():
    """flows.py:1279-1295 return-tuple matrix."""
    for rp in (False, True):
        for rd in (False, True):
            for pd in (False, True):
                r = get_predefined_flow("double_gyre", 1.0, rp, rd, pd)
                n = 1 + rp + rd + pd
                if n == 1:
                    assert isinstance(r, FlowHandle)
                else:
                    assert len(r) == n and isinstance(r[0], FlowHandle)


def test_predefined_defaults():
    f, p, dom = get_predefined_flow("double_gyre", int_direction=-1.0)
    assert np.allclose(p, [-1.0, 0.1, 0.25, 0.0, 0.2 * math.pi, 0.0])
    assert dom == ((0.0, 2.0), (0.0, 1.0))
    f, p, dom = get_predefined_flow("bickley_jet", int_direction=-1.0)
    assert p[0] == 1.0 and len(p) == 12  # flows.py:1219 quirk kept
    assert np.isclose(dom[0][1], math.pi * 6.371)


def test_unknown_and_3d_flows():
    with pytest.raises(ValueError):
        get_predefined_flow("nope")
    with pytest.raises(NotImplementedError):
        get_predefined_flow("abc")


def test_raw_cfunc_address_rejected():
    with pytest.raises(TypeError):
        flowmap_grid_2D(123456, 0.0, 1.0, np.zeros(2), np.zeros(2), np.ones(6))
    with pytest.raises(TypeError):
        flowmap(123456, 0.0, 1.0, np.zeros((2, 2)), np.ones(6))


def test_lsoda_not_available():
    f, p, _ = get_predefined_flow("double_gyre")
    with pytest.raises(NotImplementedError):
        flowmap_grid_2D(f, 0.0, 1.0, np.zeros(2), np.zeros(2), p, method="lsoda")
    with pytest.raises(ValueError):
        flowmap_grid_2D(f, 0.0, 1.0, np.zeros(2), np.zeros(2), p, method="rk4")


def test_flow_handle_not_picklable():
    import pickle

    f = get_predefined_flow("double_gyre", return_default_params=False, return_domain=False)
    with pytest.raises(TypeError):
        pickle.dumps(f)


def test_pmc_dispatch_mean_parses_counter_csv(tmp_path):
    """bench.py's reading of a rocprofv3 counter-collection CSV (the flow
    map's FETCH_SIZE / WRITE_SIZE per dispatch): rows of one dispatch are
    summed (a counter collected per XCD or channel), dispatches averaged,
    other kernels ignored; no match gives None."""
    import importlib.util
    import os

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("nbx_bench", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    rows = [("1", "void nbx::k_flowmap<nbx::Spline3<double, 4, false, false>, double, 4>", 100.0),
            ("1", "void nbx::k_flowmap<nbx::Spline3<double, 4, false, false>, double, 4>", 20.0),
            ("2", "void nbx::k_ftle_strip<true>(...)", 999.0),
            ("3", "void nbx::k_flowmap<nbx::Spline3<double, 4, false, false>, double, 4>", 60.0)]
    path = tmp_path / "run_counter_collection.csv"
    with open(path, "w") as fh:
        fh.write("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n")
        for d, k, v in rows:
            fh.write(f'{d},"{k}",FETCH_SIZE,{v}\n')
    assert bench.pmc_dispatch_mean(str(path), ("k_flowmap", "Spline3<double")) == (90.0, 2)
    assert bench.pmc_dispatch_mean(str(path), ("copyBuffer",)) is None
