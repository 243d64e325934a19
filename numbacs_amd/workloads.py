"""Seeded synthetic workloads for the FTLE path (the configurations the
benchmark and the parity tests run).

* double gyre, 401x201 particles on [0,2]x[0,1], t0=0, T=-10
  (examples/ftle/plot_dg_ftle.py:24-48 of the reference);
* gridded velocity: a streamfunction that is a sum of 64 random Fourier modes
  (integer wavenumbers with |k| in 1..12 cycles per domain, amplitude k^-2,
  random phase, each mode drifting in time), u = -dpsi/dlat, v = dpsi/dlon,
  scaled to an RMS speed of 30 km/h, sampled on a 0.5 degree lon/lat grid
  (lon -100..35, lat -5..45, 271x101) at hourly frames; particles on a
  0.2 degree grid (676x251), t0=72 h, T=-72 h, spherical=1
  (the analogue of examples/ftle/plot_merra_ftle.py:33-76);
* Bickley jet, 2001x801 particles, T=-40 days (plot_bickley_ftle.py).
"""
import hashlib
import math

import numpy as np

FIELD_SEED = 20261202


def double_gyre_grid(nx=401, ny=201):
    x = np.linspace(0, 2, nx)
    y = np.linspace(0, 1, ny)
    return x, y


def double_gyre_params(direction=-1.0):
    return np.array([direction, 0.1, 0.25, 0.0, 0.2 * math.pi, 0.0])


def bickley_params(direction=-1.0):
    r_e = 6371.0e-3
    U0 = 86400 * 62.66e-6
    c2 = 0.205 * U0
    c3 = 0.461 * U0
    c1 = c3 + (math.sqrt(5) - 1) * (c2 - c3)
    return np.array([direction, U0, 1770.0e-3, 0.0075, 0.15, 0.3, 2.0 / r_e, 4.0 / r_e,
                     6.0 / r_e, c1, c2, c3])


def bickley_grid(nx=2001, ny=801):
    r_e = 6371.0e-3
    return np.linspace(0, math.pi * r_e, nx), np.linspace(-3, 3, ny)


def field_axes(nt=96):
    t = np.arange(nt, dtype=np.float64)           # hours
    lon = np.linspace(-100.0, 35.0, 271)          # 0.5 degree
    lat = np.linspace(-5.0, 45.0, 101)
    return t, lon, lat


def particle_axes(spacing=0.2):
    nx = int(round(135.0 / spacing)) + 1
    ny = int(round(50.0 / spacing)) + 1
    return np.linspace(-100.0, 35.0, nx), np.linspace(-5.0, 45.0, ny)


def fourier_modes(seed=FIELD_SEED, nmodes=64):
    rng = np.random.default_rng(seed)
    modes = []
    while len(modes) < nmodes:
        kx = int(rng.integers(-12, 13))
        ky = int(rng.integers(0, 13))
        k = math.hypot(kx, ky)
        if k < 1 or k > 12:
            continue
        amp = k ** -2
        phase = rng.uniform(0, 2 * math.pi)
        omega = 2 * math.pi * rng.uniform(-1.0, 1.0) / 24.0  # up to one cycle per day
        modes.append((kx, ky, amp, phase, omega))
    return modes


def synthetic_velocity(nt=96, seed=FIELD_SEED, rms_kmh=30.0):
    """(t, lon, lat, U, V): U, V of shape (nt, 271, 101) in km/h, 'ij' order.

    psi = sum_m amp_m cos(B_m(x, y) - omega_m t) with B_m = 2 pi (kx x + ky y)
    + phase_m; u = -dpsi/dlat and v = dpsi/dlon (per degree) are sums of
    amp_m k_m sin(B_m - omega_m t) = amp_m k_m (sin B_m cos omega_m t -
    cos B_m sin omega_m t), formed as two (nt x modes) @ (modes x points)
    products."""
    t, lon, lat = field_axes(nt)
    Lx, Ly = lon[-1] - lon[0], lat[-1] - lat[0]
    X, Y = np.meshgrid((lon - lon[0]) / Lx, (lat - lat[0]) / Ly, indexing="ij")
    modes = fourier_modes(seed)
    B = np.stack([2 * math.pi * (kx * X + ky * Y) + phase for kx, ky, _, phase, _ in modes])
    B = B.reshape(len(modes), -1)
    SB, CB = np.sin(B), np.cos(B)
    om = np.array([m[4] for m in modes])
    Tc, Ts = np.cos(np.outer(t, om)), np.sin(np.outer(t, om))
    amp = np.array([m[2] for m in modes])
    cu = amp * np.array([2 * math.pi * m[1] / Ly for m in modes])
    cv = -amp * np.array([2 * math.pi * m[0] / Lx for m in modes])
    shape = (nt,) + X.shape
    U = ((Tc * cu) @ SB - (Ts * cu) @ CB).reshape(shape)
    V = ((Tc * cv) @ SB - (Ts * cv) @ CB).reshape(shape)
    scale = rms_kmh / math.sqrt(np.mean(U ** 2 + V ** 2))
    return t, lon, lat, U * scale, V * scale


def field_sha256(U, V):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(U, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(V, dtype=np.float64).tobytes())
    return h.hexdigest()


def time_series_requests(nreq=48, seed=FIELD_SEED + 1, t0_first=72.0,
                         windows=(-24.0, -48.0, -72.0), nx_range=(201, 1351),
                         ny_range=(101, 501)):
    """FTLE time-series requests over the gridded field extended to 480
    hourly frames: hourly t0 from 72 h, window lengths cycling through
    -24/-48/-72 h, particle grids of seeded uneven sizes between 201x101 and
    1351x501 over the velocity box."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nreq):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        out.append(dict(t0=t0_first + k, T=windows[k % len(windows)],
                        x=np.linspace(-100.0, 35.0, nx), y=np.linspace(-5.0, 45.0, ny)))
    return out
