"""The CPU oracle against the reference's fixtures and literals, scipy's
DOP853 (Hairer's Fortran code), numpy.percentile and scipy.ndimage.label --
the checks that pin it before it is trusted as the checker of the HIP path.
Planted mistakes (in the oracle's inputs, on the CPU) show the checks bite.
"""
import math
import os

import numpy as np
import pytest
from scipy import ndimage

import oracle
from conftest import apply_mask, flowmap_agreement, ic_sensitivity, reconstruct

DG = "double_gyre"


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.lib()


def dg_params(direction=1.0):
    return np.array([direction, 0.1, 0.25, 0.0, 0.2 * math.pi, 0.0])


def test_flowmap_fixture_bit_exact_f32(coords_dg, fm_data, mask_dg):
    """fm.npy (tests/test_integration.py:55-64): all 462 values equal after
    the float32 cast."""
    x, y = coords_dg
    fm, st, ns = oracle.flowmap_grid_2d(DG, dg_params(), 0.0, 8.0, x, y)
    assert (st == 1).all()
    assert (fm.astype(np.float32) == fm_data).all()
    fmm, stm, _ = oracle.flowmap_grid_2d(DG, dg_params(), 0.0, 8.0, x, y, mask=mask_dg)
    assert np.allclose(fmm.astype(np.float32), apply_mask(fm_data, mask_dg))
    assert (stm[mask_dg] == 0).all()


@pytest.mark.parametrize("case", ["dg_fwd8", "dg_c1_sub", "dg_t20_sub"])
def test_flowmap_vs_scipy_dop853(scipy_golden, case):
    G = scipy_golden
    t0, T = G[case + "_t0T"]
    fm, _, _ = oracle.flowmap_grid_2d(DG, G[case + "_params"], t0, T, G[case + "_x"],
                                      G[case + "_y"])
    ref = G[case + "_fm"]
    assert np.abs(fm - ref).max() / np.abs(ref).max() < 1e-11


def test_flowmap_bickley_vs_scipy_dop853(scipy_golden):
    """Bickley jet over 40 days is chaotic: most trajectories agree to
    rounding, the rest within the spread a 1-ulp change of the initial
    condition produces."""
    G = scipy_golden
    c = "bickley_t40_sub"
    x, y, p = G[c + "_x"], G[c + "_y"], G[c + "_params"]
    ref = G[c + "_fm"]

    def run(xx, yy):
        return oracle.flowmap_grid_2d("bickley_jet", p, 0.0, -40.0, xx, yy)[0]

    fm, sens = ic_sensitivity(run, x, y)
    ok, frac, err = flowmap_agreement(fm, ref, sens)
    assert ok
    assert frac > 0.8
    assert np.median(err) < 1e-10 * np.abs(ref).max()


def test_planted_tableau_error_is_caught(scipy_golden):
    """A wrong direction sign (planted in the oracle's input) breaks the
    agreement with scipy by orders of magnitude."""
    G = scipy_golden
    c = "dg_c1_sub"
    p = G[c + "_params"].copy()
    p[4] *= 1.0 + 1e-9   # perturb omega at the 1e-9 level
    fm, _, _ = oracle.flowmap_grid_2d(DG, p, 0.0, -10.0, G[c + "_x"], G[c + "_y"])
    assert np.abs(fm - G[c + "_fm"]).max() > 1e-9


def test_velocity_literals(literals):
    ti = np.array([0.05, 0.12, 0.18])
    xi = np.array([0.1, 0.4, 0.7])
    Ti, Xi, Yi = np.meshgrid(ti, xi, xi, indexing="ij")
    pts = np.column_stack((Xi.ravel(), Yi.ravel()))
    v = oracle.rhs_eval(DG, dg_params(), Ti.ravel(), pts)
    assert np.allclose(v.T, literals["dg_vel"])
    r_e = 6371.0e-3
    U0 = 86400 * 62.66e-6
    c2, c3 = 0.205 * U0, 0.461 * U0
    pb = np.array([1.0, U0, 1.77, 0.0075, 0.15, 0.3, 2 / r_e, 4 / r_e, 6 / r_e,
                   c3 + (math.sqrt(5) - 1) * (c2 - c3), c2, c3])
    v = oracle.rhs_eval("bickley_jet", pb, Ti.ravel(), pts)
    assert np.allclose(v.T, literals["bickley_vel"])


def _literal_field():
    t = np.array([0.0, 0.1, 0.2])
    x = np.array([0.0, 0.5, 1.0])
    T_, X, Y = np.meshgrid(t, x, x, indexing="ij")
    return t, x, np.sin(X) * np.cos(Y) + np.sin(T_), np.cos(X) * np.sin(Y) + np.cos(T_)


def test_prefilter_literals(literals):
    t, x, u, v = _literal_field()
    assert np.allclose(oracle.prefilter3(u), literals["prefilter3_Cu"])
    assert np.allclose(oracle.prefilter3(v), literals["prefilter3_Cv"])
    X, Y = np.meshgrid(x, x, indexing="ij")
    assert np.allclose(oracle.prefilter2(np.sin(X) * np.cos(Y)), literals["prefilter2_Cu"])
    assert np.allclose(oracle.prefilter2(np.cos(X) * np.sin(Y)), literals["prefilter2_Cv"])


def test_prefilter_interpolates_and_is_natural():
    """The spline reproduces the samples at the nodes and has zero second
    derivative at the ends of each axis."""
    rng = np.random.default_rng(1)
    U = rng.standard_normal((9, 13, 11))
    C = oracle.prefilter3(U)
    grid = ((0.0, 8.0, 9), (0.0, 1.2, 13), (-1.0, 1.0, 11))
    T_, X, Y = np.meshgrid(np.linspace(0, 8, 9), np.linspace(0, 1.2, 13), np.linspace(-1, 1, 11),
                           indexing="ij")
    vals = oracle.spline3_eval(grid, C, np.column_stack((T_.ravel(), X.ravel(), Y.ravel())))
    assert np.abs(vals - U.ravel()).max() < 1e-12
    d2 = C[:, :, 0] - 2 * C[:, :, 1] + C[:, :, 2]
    assert np.abs(d2).max() < 1e-12


def test_spline_eval_literals(literals):
    t, x, u, v = _literal_field()
    grid = ((0.0, 0.2, 3), (0.0, 1.0, 3), (0.0, 1.0, 3))
    ti = np.array([0.05, 0.12, 0.18])
    xi = np.array([0.1, 0.4, 0.7])
    Ti, Xi, Yi = np.meshgrid(ti, xi, xi, indexing="ij")
    pts = np.column_stack((Ti.ravel(), Xi.ravel(), Yi.ravel()))
    assert np.allclose(oracle.spline3_eval(grid, oracle.prefilter3(u), pts),
                       literals["spline3_vel"][0])
    assert np.allclose(oracle.spline3_eval(grid, oracle.prefilter3(v), pts),
                       literals["spline3_vel"][1])


def test_spline_reproduces_cubics():
    """A cubic polynomial in each variable is reproduced exactly in-grid
    only if it has zero second derivative at the ends -- a linear function
    is, to rounding, everywhere in the box."""
    t, x, y = np.linspace(0, 3, 7), np.linspace(-1, 1, 9), np.linspace(0, 2, 8)
    T_, X, Y = np.meshgrid(t, x, y, indexing="ij")
    F = 0.3 * T_ - 1.7 * X + 2.5 * Y + 0.25
    C = oracle.prefilter3(F)
    rng = np.random.default_rng(2)
    q = np.column_stack((rng.uniform(0, 3, 300), rng.uniform(-1, 1, 300), rng.uniform(0, 2, 300)))
    grid = ((0.0, 3.0, 7), (-1.0, 1.0, 9), (0.0, 2.0, 8))
    exact = 0.3 * q[:, 0] - 1.7 * q[:, 1] + 2.5 * q[:, 2] + 0.25
    assert np.abs(oracle.spline3_eval(grid, C, q) - exact).max() < 1e-12
    # out of the box: constant is zero, nearest clamps, linear continues the
    # plane
    qo = q + np.array([4.0, 0.0, 0.0])
    assert np.array_equal(oracle.spline3_eval(grid, C, qo, "constant"), np.zeros(len(q)))
    clamp = 0.3 * 3.0 - 1.7 * q[:, 1] + 2.5 * q[:, 2] + 0.25
    assert np.abs(oracle.spline3_eval(grid, C, qo, "nearest") - clamp).max() < 1e-12
    lin = 0.3 * qo[:, 0] - 1.7 * q[:, 1] + 2.5 * q[:, 2] + 0.25
    assert np.abs(oracle.spline3_eval(grid, C, qo, "linear") - lin).max() < 1e-11
    # the box is closed: its faces are inside for every mode
    qb = q.copy()
    qb[:, 0] = 3.0
    for mode in ("constant", "nearest", "linear"):
        assert np.abs(oracle.spline3_eval(grid, C, qb, mode) - clamp).max() < 1e-12


def test_ftle_fixture(coords_dg, fm_data, ftle_data, mask_dg):
    x, y = coords_dg
    f = oracle.ftle_grid_2d(fm_data, 8.0, x[1], y[1])
    assert np.allclose(f.astype(np.float32), ftle_data)
    fmk = oracle.ftle_grid_2d(fm_data, 8.0, x[1], y[1], mask=mask_dg)
    assert np.allclose(fmk.astype(np.float32), apply_mask(ftle_data, mask_dg))


def test_gradient_and_eig_literals(coords_dg, fm_data, literals):
    """gradF_stencil_2D at (5, 5) and eigvalsh_max_2D literals
    (tests/test_utils.py:74-88,133-141): an identity flow map shifted by a
    linear map A has FTLE log(max eig(A^T A))/(2|T|)."""
    B = np.array([[1.3, -2.4], [0.7, 5.2]])
    x = np.linspace(0, 1, 5)
    X, Y = np.meshgrid(x, x, indexing="ij")
    fm = np.stack((B[0, 0] * X + B[0, 1] * Y, B[1, 0] * X + B[1, 1] * Y), -1)
    ev, V = oracle.C_eig_2d(fm, x[1], x[1])
    assert np.isclose(ev[2, 2, 1], literals["eigvalsh_max"])
    ftle = oracle.ftle_grid_2d(fm, 2.0, x[1], x[1])
    assert np.isclose(ftle[2, 2], math.log(float(literals["eigvalsh_max"])) / 4)
    xs, ys = coords_dg
    g = literals["gradF_5_5"]
    dxdx = (fm_data[6, 5, 0] - fm_data[4, 5, 0]) / (2 * xs[1])
    assert np.isclose(dxdx, g[0])
    C = np.array([[g[0] ** 2 + g[2] ** 2, g[0] * g[1] + g[2] * g[3]],
                  [g[0] * g[1] + g[2] * g[3], g[1] ** 2 + g[3] ** 2]])
    f = oracle.ftle_grid_2d(fm_data, 8.0, xs[1], ys[1])
    assert np.isclose(f[5, 5], math.log(np.linalg.eigvalsh(C)[1]) / 16, rtol=1e-6)


def test_C_eig_fixture(coords_dg, fm_data, C_eig_data):
    x, y = coords_dg
    ev, V = oracle.C_eig_2d(fm_data, x[1], y[1])
    ev_ref, V_ref = C_eig_data
    assert np.allclose(ev.astype(np.float32), ev_ref)
    R = reconstruct(ev, V)
    assert np.allclose(R.astype(np.float32), reconstruct(ev_ref.astype(np.float64),
                                                         V_ref.astype(np.float64)).astype(
                                                             np.float32))


def test_eig_vs_numpy_eigh():
    rng = np.random.default_rng(4)
    fm = rng.standard_normal((40, 30, 2)).cumsum(0).cumsum(1)
    ev, V = oracle.C_eig_2d(fm, 0.1, 0.2)
    F = fm
    gx = (F[2:, 1:-1] - F[:-2, 1:-1]) / 0.2
    gy = (F[1:-1, 2:] - F[1:-1, :-2]) / 0.4
    J = np.stack((np.stack((gx[..., 0], gy[..., 0]), -1), np.stack((gx[..., 1], gy[..., 1]), -1)),
                 -2)
    C = np.swapaxes(J, -1, -2) @ J
    w, _ = np.linalg.eigh(C)
    assert np.allclose(ev[1:-1, 1:-1], w, rtol=1e-10, atol=1e-12)
    Cr = reconstruct(ev[1:-1, 1:-1], V[1:-1, 1:-1])
    assert np.allclose(Cr, C, rtol=1e-10, atol=1e-10)


def test_ridge_pts_fixture(coords_dg, ftle_data, C_eig_data, ridge_pts_data):
    x, y = coords_dg
    pts = oracle.ridge_pts(ftle_data, C_eig_data[1][:, :, :, 1], x, y)
    assert pts.shape == ridge_pts_data.shape and np.allclose(pts, ridge_pts_data)


def test_ridges_fixture(coords_dg, ftle_data, C_eig_data, ridge_data):
    """ftle_ridges on the fixture world equals ridges.pkl: same ridges, same
    order, same points (tests/test_extraction.py:14-19 of the reference).
    The pickle holds the points the reference computed before its inputs were
    stored as float32, so they agree to 5.5e-10 here, inside the reference's
    own np.allclose bar (atol 1e-8)."""
    x, y = coords_dg
    r = oracle.ridges(ftle_data, C_eig_data[1][:, :, :, 1], x, y)
    assert [len(a) for a in r] == [len(a) for a in ridge_data] == [6, 9, 3]
    for got, want in zip(r, ridge_data):
        assert np.abs(got - want).max() < 2e-9


def test_pickle_fixture_reader_refuses_foreign_globals(tmp_path):
    """The fixture converter accepts numpy's array reconstruct globals only:
    a stream naming any other callable is rejected before anything runs."""
    import importlib.util
    import pickle

    spec = importlib.util.spec_from_file_location(
        "make_pickle_fixtures", os.path.join(os.path.dirname(__file__), "golden",
                                             "make_pickle_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    good = tmp_path / "good.pkl"
    arrs = [np.arange(6.0).reshape(3, 2), np.array([[1.5, -2.0]])]
    good.write_bytes(pickle.dumps(arrs, protocol=4))
    back = mod.read_array_list(str(good))
    assert all(np.array_equal(a, b) for a, b in zip(arrs, back))
    bad = tmp_path / "bad.pkl"
    bad.write_bytes(pickle.dumps([os.getcwd], protocol=4))
    with pytest.raises(ValueError):
        mod.read_array_list(str(bad))


@pytest.mark.parametrize("q", [0, 10, 50, 90, 99.9, 100])
def test_percentile_vs_numpy(q):
    rng = np.random.default_rng(7)
    f = rng.standard_normal(10007)
    f[::5] = 0.0
    assert oracle.percentile(f, q) == np.percentile(f, q)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_label8_vs_scipy(seed):
    rng = np.random.default_rng(seed)
    b = rng.random((120, 77)) < 0.45
    lab, n = oracle.label8(b)
    ref, nref = ndimage.label(b, structure=ndimage.generate_binary_structure(2, 2))
    assert n == nref and np.array_equal(lab, ref)


# ------------------------------------------------- aux grid, composition --
def test_flowmap_aux_fixture(coords_dg, fm_aux_data, mask_dg):
    """fm_aux.npy (tests/test_integration.py:66-75): every value equal after
    the float32 cast; masked points 0."""
    x, y = coords_dg
    fa = oracle.flowmap_aux_grid_2d(DG, dg_params(), 0.0, 8.0, x, y)
    assert np.array_equal(fa.astype(np.float32), fm_aux_data)
    fam = oracle.flowmap_aux_grid_2d(DG, dg_params(), 0.0, 8.0, x, y, mask=mask_dg)
    assert np.allclose(fam.astype(np.float32), apply_mask(fm_aux_data, mask_dg))


def test_flowmap_aux_edge_rules(coords_dg):
    """Edge points carry only the centre particle (eig_main) or all four
    (not eig_main); compute_edge=False leaves the border at 0.  Every
    integrated particle equals the point-list flow map of its start."""
    x, y = coords_dg
    h = 1e-3
    for eig_main in (True, False):
        for edge in (True, False):
            fa = oracle.flowmap_aux_grid_2d(DG, dg_params(), 0.0, 2.0, x, y, h=h,
                                            eig_main=eig_main, compute_edge=edge)
            naux = 5 if eig_main else 4
            assert fa.shape == (len(x), len(y), naux, 2)
            off = np.array([[h, 0], [-h, 0], [0, h], [0, -h], [0, 0]])[:naux]
            X, Y = np.meshgrid(x, y, indexing="ij")
            starts = np.stack([X, Y], -1)[:, :, None, :] + off[None, None]
            want = np.ones((len(x), len(y), naux), bool)
            border = np.zeros((len(x), len(y)), bool)
            border[[0, -1], :] = True
            border[:, [0, -1]] = True
            if not edge:
                want[border] = False
            elif eig_main:
                want[border, :4] = False
            ref, _ = oracle.flowmap_pts(DG, dg_params(), 0.0, 2.0, starts[want])
            assert np.array_equal(fa[want], ref)
            assert not fa[~want].any()


def test_C_tensor_aux_fixture(coords_dg, fm_aux_data, C_aux_data, mask_dg):
    """C.npy (tests/test_diagnostics.py:99-110)."""
    C = oracle.C_tensor_aux_2d(fm_aux_data)
    assert np.allclose(C.astype(np.float32), C_aux_data)
    Cm = oracle.C_tensor_aux_2d(fm_aux_data, mask=mask_dg)
    assert np.allclose(Cm.astype(np.float32), apply_mask(C_aux_data, mask_dg))


def test_C_eig_aux_fixture(coords_dg, fm_aux_data, C_eig_aux_data, mask_dg):
    """Cevals_aux.npy / Cevecs_aux.npy compared as V diag(l) V^T
    (tests/test_diagnostics.py:112-131)."""
    x, y = coords_dg
    ev, V = oracle.C_eig_aux_2d(fm_aux_data, x[1], y[1])
    want = reconstruct(*C_eig_aux_data)
    got = reconstruct(ev.astype(np.float32), V.astype(np.float32))
    assert np.allclose(got, want)
    assert np.allclose(ev.astype(np.float32), C_eig_aux_data[0])
    evm, Vm = oracle.C_eig_aux_2d(fm_aux_data, x[1], y[1], mask=mask_dg)
    assert np.allclose(reconstruct(evm.astype(np.float32), Vm.astype(np.float32)),
                       apply_mask(want, mask_dg))


def test_C_eig_aux_without_main_matches_numpy(fm_aux_data, coords_dg):
    """eig_main=False: both from the aux tensor, on [1, n-2]; eigenvalues
    against numpy.linalg.eigvalsh of the tensor C_tensor_2D restates."""
    x, y = coords_dg
    F = fm_aux_data.astype(np.float64)[:, :, :4]
    ev, V = oracle.C_eig_aux_2d(F, x[1], y[1], eig_main=False)
    h = 1e-5
    g = np.stack([(F[..., 0, 0] - F[..., 1, 0]) / (2 * h), (F[..., 2, 0] - F[..., 3, 0]) / (2 * h),
                  (F[..., 0, 1] - F[..., 1, 1]) / (2 * h), (F[..., 2, 1] - F[..., 3, 1]) / (2 * h)],
                 -1)
    C = np.empty(F.shape[:2] + (2, 2))
    C[..., 0, 0] = g[..., 0] ** 2 + g[..., 2] ** 2
    C[..., 0, 1] = C[..., 1, 0] = g[..., 0] * g[..., 1] + g[..., 2] * g[..., 3]
    C[..., 1, 1] = g[..., 1] ** 2 + g[..., 3] ** 2
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(ev[inner], np.linalg.eigvalsh(C[inner]), rtol=1e-10, atol=1e-12)
    assert not ev[0].any() and not ev[:, 0].any()


def test_composition_fixtures(coords_dg, composition_data):
    """flowmap_composition_initial / _step (tests/test_integration.py:92-114):
    eight 1-unit flow maps of the double gyre from t0 = 0 composed over
    T = 8; the step shifts the window by one.  The CONSTANT extrapolation of
    the interpolation is zero outside the grid box: five trajectories end
    1.8e-15 beyond x = 2, and the fixture holds 0 there."""
    x, y = coords_dg
    grid = ((0.0, 2.0, len(x)), (0.0, 1.0, len(y)))
    fms = np.stack([oracle.flowmap_grid_2d(DG, dg_params(), float(k), 1.0, x, y)[0]
                    for k in range(8)])
    assert np.allclose(fms.astype(np.float32), composition_data["fms_ci"])
    c0 = oracle.flowmap_composition(fms, grid, 8)
    assert np.allclose(c0.astype(np.float32), composition_data["fm_ci"])
    assert (fms[..., 0] > 2.0).sum() == 5 and not c0[-1, 1:].any()
    # the step starts from the float32 fixture, as the reference test does
    fcs = composition_data["fms_ci"].astype(np.float64)
    fcs = np.concatenate([fcs[1:], oracle.flowmap_grid_2d(DG, dg_params(), 8.0, 1.0, x, y)[0][None]])
    assert np.allclose(fcs.astype(np.float32), composition_data["fms_cs"])
    assert np.allclose(oracle.flowmap_composition(fcs, grid, 8).astype(np.float32),
                       composition_data["fm_cs"])


def test_composition_of_linear_maps_is_exact():
    """Affine maps that keep the box are reproduced exactly by bilinear
    interpolation, so composing them equals their product; nT = 1 applies
    the single map twice (the reference's loop bounds)."""
    nx, ny = 9, 7
    x, y = np.linspace(0, 2, nx), np.linspace(0, 1, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    grid = ((0.0, 2.0, nx), (0.0, 1.0, ny))
    maps = [(0.9, 0.05, 0.02, 0.01), (0.8, 0.1, 0.1, 0.05), (0.95, 0.0, 0.03, 0.02)]
    F = np.stack([np.stack([a * X + c, b * Y + d], -1) for a, c, b, d in
                  [(m[0], m[2], m[0], m[3]) for m in maps]])
    px, py = X, Y
    for a, c, b, d in [(m[0], m[2], m[0], m[3]) for m in maps]:
        px, py = a * px + c, b * py + d
    got = oracle.flowmap_composition(F, grid, 3)
    assert np.abs(got - np.stack([px, py], -1)).max() < 1e-13
    one = oracle.flowmap_composition(F[:1], grid, 1)
    a, c, b, d = maps[0][0], maps[0][2], maps[0][0], maps[0][3]
    assert np.abs(one - np.stack([a * (a * X + c) + c, b * (b * Y + d) + d], -1)).max() < 1e-13


# ------------------------------------------------------- trilinear flows --
def _coords_2d():
    t, x, y = np.array([0.0, 0.1, 0.2]), np.array([0.0, 0.5, 1.0]), np.array([0.0, 0.5, 1.0])
    T_, X, Y = np.meshgrid(t, x, y, indexing="ij")
    u = np.sin(X) * np.cos(Y) + np.sin(T_)
    v = np.cos(X) * np.sin(Y) + np.cos(T_)
    ti, xi = np.array([0.05, 0.12, 0.18]), np.array([0.1, 0.4, 0.7])
    Ti, Xi, Yi = np.meshgrid(ti, xi, xi, indexing="ij")
    pts = np.column_stack((Ti.ravel(), Xi.ravel(), Yi.ravel()))
    return ((0.0, 0.2, 3), (0.0, 1.0, 3), (0.0, 1.0, 3)), u, v, pts


def test_trilinear_literals(literals):
    """tests/test_flows.py:261-285 of the reference (get_callable_linear_2D
    at 27 in-grid points), through the interpolant and the flow RHS."""
    grid, u, v, pts = _coords_2d()
    want = literals["linear3_vel"]
    assert np.allclose(oracle.linear3_eval(grid, u, pts), want[0])
    assert np.allclose(oracle.linear3_eval(grid, v, pts), want[1])
    rhs = oracle.rhs_eval("linear", np.array([1.0]), pts[:, 0], pts[:, 1:],
                          spline=oracle.Linear(grid, u, v))
    assert np.allclose(rhs.T, want)


def test_trilinear_exact_and_extrapolation():
    """Fields linear in each variable are reproduced to rounding in the box;
    outside it constant gives 0, nearest the clamped value, linear the
    field itself (a multilinear field continues its boundary cell)."""
    t, x, y = np.linspace(0, 3, 7), np.linspace(-1, 1, 9), np.linspace(0, 2, 8)
    T_, X, Y = np.meshgrid(t, x, y, indexing="ij")

    def f(a, b, c):
        return 0.3 * a - 1.7 * b + 2.5 * c + 0.25 + 0.5 * a * b - 0.2 * b * c

    F = f(T_, X, Y)
    grid = ((0.0, 3.0, 7), (-1.0, 1.0, 9), (0.0, 2.0, 8))
    rng = np.random.default_rng(4)
    q = np.column_stack((rng.uniform(0, 3, 300), rng.uniform(-1, 1, 300), rng.uniform(0, 2, 300)))
    assert np.abs(oracle.linear3_eval(grid, F, q) - f(*q.T)).max() < 1e-12
    qo = q + np.array([0.0, 0.0, 2.5])
    assert np.array_equal(oracle.linear3_eval(grid, F, qo, "constant"), np.zeros(len(q)))
    assert np.abs(oracle.linear3_eval(grid, F, qo, "nearest") - f(q[:, 0], q[:, 1], 2.0)).max() < 1e-12
    assert np.abs(oracle.linear3_eval(grid, F, qo, "linear") - f(*qo.T)).max() < 1e-11


# ------------------------------------------------ spherical conversion --

SPH_TOL = dict(lon_tol_uniform=1e-8, lon_tol_sin=1e-8, lat_tol=1e-9)


def _sph_oracle(spherical, a, T, U0, V0, r=6371.0, field_spherical=None):
    import spherical_cases as SC

    t, lon, lat, U, V = SC.field(field_spherical or spherical, a, U0, V0)
    grid = ((t[0], t[-1], len(t)), (lon[0], lon[-1], len(lon)), (lat[0], lat[-1], len(lat)))
    S = oracle.Spline(grid, oracle.prefilter3(U), oracle.prefilter3(V), spherical=spherical, r=r)
    lon0, lat0 = SC.particles(spherical)
    fm, st, _ = oracle.flowmap_grid_2d("spline", np.array([math.copysign(1.0, T)]), SC.T0, T,
                                       lon0, lat0, rtol=1e-10, atol=1e-12, spline=S)
    assert (st == 1).all()
    return fm


@pytest.mark.parametrize("spherical", [1, 2])
@pytest.mark.parametrize("a", [0.0, 0.5])
@pytest.mark.parametrize("V0", [0.0, 20.0])
@pytest.mark.parametrize("T", [24.0, -48.0])
def test_spherical_conversion_closed_form(spherical, a, V0, T):
    """flows.py:154-239: km/h -> deg/h with 180/(pi r cos lat) and 180/(pi r),
    the +-180 (spherical=1) and 0/360 (spherical=2) wraps of the lookup, for
    both time directions: every trajectory matches the closed form of
    tests/spherical_cases.py (rhumb lines; a separable zonal profile)."""
    import spherical_cases as SC

    fm = _sph_oracle(spherical, a, T, 40.0, V0)
    SC.check(fm, spherical, a, T, 40.0, V0, SPH_TOL["lon_tol_sin" if a else "lon_tol_uniform"],
             SPH_TOL["lat_tol"])


@pytest.mark.parametrize("plant", ["radius", "wrap"])
def test_spherical_planted_mistakes_are_caught(plant):
    """A radius off by 1e-4, or the other wrap convention, fails the check."""
    import spherical_cases as SC

    if plant == "radius":
        fm = _sph_oracle(1, 0.5, -48.0, 40.0, 20.0, r=6371.0 * (1 + 1e-4))
    else:
        fm = _sph_oracle(1, 0.5, -48.0, 40.0, 20.0, field_spherical=2)
    with pytest.raises(AssertionError):
        SC.check(fm, 1, 0.5, -48.0, 40.0, 20.0, SPH_TOL["lon_tol_sin"], SPH_TOL["lat_tol"])


def test_ftle_bound_holds_and_is_tight():
    """conftest.ftle_bound: the FTLE change caused by a flow-map perturbation
    stays within the bound, and the bound is within 10x of the change where
    the change is large (it is a first-order bound, not a blanket)."""
    from conftest import ftle_bound

    x, y = np.linspace(0, 2, 101), np.linspace(0, 1, 51)
    fm, _, _ = oracle.flowmap_grid_2d(DG, dg_params(-1.0), 0.0, -10.0, x, y)
    dx, dy = x[1] - x[0], y[1] - y[0]
    rng = np.random.default_rng(3)
    fm2 = fm + 1e-7 * rng.standard_normal(fm.shape)
    f1 = oracle.ftle_grid_2d(fm, -10.0, dx, dy)
    f2 = oracle.ftle_grid_2d(fm2, -10.0, dx, dy)
    b = ftle_bound(fm, fm2, -10.0, dx, dy)
    d = np.abs(f1 - f2)
    assert (d <= b).all()
    big = d > 1e-3 * d.max()
    assert np.median(b[big] / d[big]) < 10
