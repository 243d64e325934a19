"""Closed-form trajectories of get_flow_2D's spherical conversion
(flows.py:154-239 of the reference).

On a lon/lat degree grid with velocities in km/h the reference integrates
  dlon/dt = u * 180 / (pi r cos(lat * pi/180)),   dlat/dt = v * 180 / (pi r)
with lon wrapped only where the velocity is looked up: into [-180, 180) for
spherical=1, [0, 360) for spherical=2; the state itself is not wrapped.

The field used here is u = U0 (1 + a sin(lon)), v = V0 (km/h, steady).  Then
lat is linear in time, lat1 = lat0 + V0 T / r (radians), and the lon equation
separates:
  integral_{lon0}^{lon1} dlon / (1 + a sin lon) = U0 T / (r cos lat0)          (V0 = 0)
                                               = (U0/V0) (gd^-1(lat1) - gd^-1(lat0))
with gd^-1(phi) = atanh(sin phi), the rhumb line of a uniform field (a = 0,
where the left side is lon1 - lon0 itself).  integration direction p0 cancels:
the result depends on T only.

The velocity grid spans the whole circle ([-180, 180] or [0, 360], both ends
included) so wrapped lookups stay inside the box; sin(lon) has zero second
derivative at both ends, which the natural spline's end condition matches.
Particles start close to the wrap (+-180 for spherical=1, 0/360 for
spherical=2) and cross it.
"""
import math

import numpy as np
from scipy import integrate

R_EARTH = 6371.0
T0 = 72.0
WINDOWS = (24.0, -48.0)  # both inside the field's time axis 0..96 h


def field(spherical, a, U0=40.0, V0=20.0):
    t = np.linspace(0.0, 96.0, 5)
    lon = np.linspace(-180.0, 180.0, 361) if spherical == 1 else np.linspace(0.0, 360.0, 361)
    lat = np.linspace(-80.0, 80.0, 161)
    _, LON, _ = np.meshgrid(t, lon, lat, indexing="ij")
    U = U0 * (1.0 + a * np.sin(np.deg2rad(LON)))
    V = np.full_like(U, V0)
    return t, lon, lat, U, V


def particles(spherical):
    """Start longitudes next to the wrap, latitudes in both hemispheres."""
    lon0 = np.array([172.5, -173.0]) if spherical == 1 else np.array([354.0, 4.5])
    lat0 = np.array([-35.0, 0.0, 10.0, 47.5])
    return lon0, lat0


def lat_final(lat0, T, V0, r=R_EARTH):
    return np.rad2deg(np.deg2rad(lat0) + V0 * T / r)


def lon_integral(lon0, lon1, a):
    """integral of 1/(1 + a sin lon) dlon over radians, lon in degrees."""
    if a == 0.0:
        return np.deg2rad(lon1 - lon0)
    val, _ = integrate.quad(lambda th: 1.0 / (1.0 + a * math.sin(th)), math.radians(lon0),
                            math.radians(lon1), epsabs=1e-15, epsrel=1e-13, limit=200)
    return val


def lon_rhs(lat0, T, U0, V0, r=R_EARTH):
    """Right-hand side of the separated lon equation."""
    phi0 = math.radians(lat0)
    if V0 == 0.0:
        return U0 * T / (r * math.cos(phi0))
    phi1 = phi0 + V0 * T / r
    return (U0 / V0) * (math.atanh(math.sin(phi1)) - math.atanh(math.sin(phi0)))


def check(fm, spherical, a, T, U0, V0, lon_tol, lat_tol):
    """fm (nlon0, nlat0, 2) from the particles() grid: returns the largest
    lon residual (relative to the closed-form integral) and lat error (deg).
    For a = 0 the lon residual is in degrees (the closed form is explicit)."""
    lon0, lat0 = particles(spherical)
    worst_lon = worst_lat = 0.0
    for i, l0 in enumerate(lon0):
        for j, p0 in enumerate(lat0):
            lon1, lat1 = fm[i, j]
            worst_lat = max(worst_lat, abs(lat1 - lat_final(p0, T, V0)))
            want = lon_rhs(p0, T, U0, V0)
            if a == 0.0:
                res = abs(lon1 - (l0 + math.degrees(want)))
            else:
                res = abs(lon_integral(l0, lon1, a) - want) / abs(want)
            worst_lon = max(worst_lon, res)
    assert worst_lat <= lat_tol, worst_lat
    assert worst_lon <= lon_tol, worst_lon
    # the trajectories really cross the wrap
    crossed = np.abs(fm[..., 0] // 360 - np.asarray(lon0)[:, None] // 360) > 0
    if spherical == 1:
        crossed = np.abs(np.floor((fm[..., 0] + 180) / 360)
                         - np.floor((np.asarray(lon0)[:, None] + 180) / 360)) > 0
    assert crossed.any()
    return worst_lon, worst_lat
