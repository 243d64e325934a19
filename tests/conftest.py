import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# the reference's fixture world (tests/conftest.py:9-22 of the reference)
NX, NY = 21, 11


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP library)")


def golden(name):
    return os.path.join(GOLDEN, name)


@pytest.fixture
def coords_dg():
    return np.linspace(0, 2, NX), np.linspace(0, 1, NY)


@pytest.fixture
def mask_dg():
    mask = np.zeros((NX, NY), np.bool_)
    mask[::2, ::2] = True
    return mask


@pytest.fixture
def fm_data():
    return np.load(golden("fm.npy"))


@pytest.fixture
def ftle_data():
    return np.load(golden("ftle.npy"))


@pytest.fixture
def C_eig_data():
    return np.load(golden("Cevals.npy")), np.load(golden("Cevecs.npy"))


@pytest.fixture
def ridge_pts_data():
    return np.load(golden("ridge_pts.npy"))


@pytest.fixture
def ridge_data():
    """The reference's ridges.pkl (tests/conftest.py:134-138 of the
    reference) as data: three ridges of 6, 9 and 3 points.  Converted by
    tests/golden/make_pickle_fixtures.py without unpickling."""
    d = np.load(golden("ridges.npz"))
    return [d[f"r{k}"] for k in range(len(d.files))]


@pytest.fixture
def ordered_ridge_data():
    """The reference's ordered_ridges.pkl (tests/conftest.py:140-144 of the
    reference): two ordered curves of 6 and 7 points at dist_tol 0.1."""
    d = np.load(golden("ordered_ridges.npz"))
    return [d[f"r{k}"] for k in range(len(d.files))]


@pytest.fixture
def fm_aux_data():
    return np.load(golden("fm_aux.npy"))


@pytest.fixture
def C_aux_data():
    return np.load(golden("C.npy"))


@pytest.fixture
def C_eig_aux_data():
    return np.load(golden("Cevals_aux.npy")), np.load(golden("Cevecs_aux.npy"))


@pytest.fixture
def composition_data():
    """fm_ci, fms_ci, fm_cs, fms_cs (tests/test_integration.py:92-114)."""
    return {k: np.load(golden(k + ".npy")) for k in ("fm_ci", "fms_ci", "fm_cs", "fms_cs")}


@pytest.fixture
def literals():
    return np.load(golden("reference_literals.npz"))


@pytest.fixture
def scipy_golden():
    return np.load(golden("scipy_dop853.npz"))


def apply_mask(arr, mask):
    out = arr.copy()
    out[mask] = 0.0
    return out


def reconstruct(eigvals, eigvecs):
    """V diag(l) V^T per grid point (the reference's check,
    tests/test_diagnostics.py:18-47)."""
    D = np.zeros_like(eigvecs)
    D[..., 0, 0] = eigvals[..., 0]
    D[..., 1, 1] = eigvals[..., 1]
    return eigvecs @ D @ np.swapaxes(eigvecs, -1, -2)


def ic_sensitivity(run, x, y, rel=1e-9):
    """Per-particle response (max over components) of a flow-map function
    run(x, y) -> (nx, ny, 2) to shifting the initial grid by rel x its extent
    in x and, separately, in y.  Returns (base, response).

    An adaptive integrator's accept/reject decisions are discontinuous in
    the data: two correct FP64 implementations that round differently (FMA
    contraction, libm vs device math) occasionally take a different decision
    and so inject a local error of order rtol, which a chaotic trajectory then
    amplifies.  1e4 x this response (an initial perturbation of 1e-5 relative)
    bounds that effect; trajectories that are not sensitive get the plain
    tolerance instead."""
    base = run(x, y)
    dx = rel * max(np.ptp(x), 1e-300)
    dy = rel * max(np.ptp(y), 1e-300)
    s1 = np.abs(run(x + dx, y) - base).max(-1)
    s2 = np.abs(run(x, y + dy) - base).max(-1)
    return base, np.maximum(s1, s2)


FLOWMAP_RTOL = 1e-6  # the looser of 1e-7 and the integrator's rtol=1e-6


def flowmap_agreement(fm, ref, sens, rtol=FLOWMAP_RTOL, amplification=1e4):
    """(all particles within max(rtol*scale, amplification*sens),
    fraction within rtol*scale, per-particle error)."""
    scale = np.abs(ref).max()
    err = np.abs(fm - ref).max(-1)
    ok = bool(np.all(err <= np.maximum(rtol * scale, amplification * sens)))
    return ok, float(np.mean(err <= rtol * scale)), err


def _grad(fm, dx, dy):
    """Central-difference flow-map gradient on the interior (utils.py:9-46
    of the reference): (nx-2, ny-2, 2, 2), G[..., a, b] = dF_a / dX_b."""
    G = np.empty((fm.shape[0] - 2, fm.shape[1] - 2, 2, 2))
    G[..., :, 0] = (fm[2:, 1:-1] - fm[:-2, 1:-1]) / (2 * dx)
    G[..., :, 1] = (fm[1:-1, 2:] - fm[1:-1, :-2]) / (2 * dy)
    return G


def ftle_bound(fm_a, fm_b, T, dx, dy, slack=1e-12):
    """Per-point bound on |FTLE(fm_a) - FTLE(fm_b)| that the difference of the
    two flow maps implies (interior points; 0 on the border).

    FTLE = max(log sigma, 0) / |T| with sigma the largest singular value of
    the gradient G.  By Weyl's inequality |sigma_a - sigma_b| <= ||G_a -
    G_b||_2 <= ||G_a - G_b||_F, and |log a - log b| <= |a - b| / min(a, b);
    max(., 0) is 1-Lipschitz.  slack covers the stencil's own rounding
    (relative to the FTLE)."""
    Ga, Gb = _grad(fm_a, dx, dy), _grad(fm_b, dx, dy)
    dG = np.sqrt(((Ga - Gb) ** 2).sum(axis=(-1, -2)))
    sa = np.linalg.svd(Ga, compute_uv=False)[..., 0]
    sb = np.linalg.svd(Gb, compute_uv=False)[..., 0]
    out = np.zeros(fm_a.shape[:2])
    lo = np.minimum(sa, sb)
    b = np.where(lo > 0, dG / np.maximum(lo, 1e-300), np.inf) / abs(T)
    b += slack * np.maximum(np.log(np.maximum(sa, 1.0)), np.log(np.maximum(sb, 1.0))) / abs(T)
    out[1:-1, 1:-1] = b
    return out
