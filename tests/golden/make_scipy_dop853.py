"""Write FP64 golden flow maps with scipy's DOP853 (Hairer's Fortran code,
run through scipy.integrate.ode, never read) -> scipy_dop853.npz.

scipy's ``dop853`` with ``dfactor=0.333`` (Hairer's default fac1, the value
the reference's integrator uses) reproduces the reference fixture fm.npy
bit-exactly after the float32 cast.  For T < 0 the sign-flipped problem
f(s, y) = p0 * v(p0 * s, y) is integrated from p0*t0 to p0*(t0+T), exactly
as integration.py:164 does.  The velocity formulas are the reference's
closed forms (flows.py:1152-1158 and 1189-1213), written out here.
"""
import math
import os

import numpy as np
from scipy.integrate import ode

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scipy_dop853.npz")


def dg_velocity(p):
    def f(s, y):
        tt = p[0] * s
        a = p[2] * math.sin(p[4] * tt + p[5])
        b = 1 - 2 * a
        fx = a * y[0] ** 2 + b * y[0]
        df = 2 * a * y[0] + b
        return [p[0] * (-math.pi * p[1] * math.sin(math.pi * fx) * math.cos(math.pi * y[1]) - p[3] * y[0]),
                p[0] * (math.pi * p[1] * math.cos(math.pi * fx) * math.sin(math.pi * y[1]) * df - p[3] * y[1])]
    return f


def bickley_velocity(p):
    def f(s, y):
        tt = p[0] * s
        Y = y[1] / p[2]
        sech2 = 1 / (math.cosh(Y) ** 2)
        u = p[0] * (p[1] * sech2 + 2 * p[1] * math.tanh(Y) * sech2 * (
            p[3] * math.cos(p[6] * (y[0] - p[9] * tt)) + p[4] * math.cos(p[7] * (y[0] - p[10] * tt))
            + p[5] * math.cos(p[8] * (y[0] - p[11] * tt))))
        v = -p[0] * (p[1] * p[2] * sech2 * (
            p[3] * p[6] * math.sin(p[6] * (y[0] - p[9] * tt)) + p[4] * p[7] * math.sin(p[7] * (y[0] - p[10] * tt))
            + p[5] * p[8] * math.sin(p[8] * (y[0] - p[11] * tt))))
        return [u, v]
    return f


def dg_params(direction):
    return np.array([direction, 0.1, 0.25, 0.0, 0.2 * math.pi, 0.0])


def bickley_params(direction):
    r_e = 6371.0e-3
    U0 = 86400 * 62.66e-6
    c2 = 0.205 * U0
    c3 = 0.461 * U0
    c1 = c3 + (math.sqrt(5) - 1) * (c2 - c3)
    return np.array([direction, U0, 1770.0e-3, 0.0075, 0.15, 0.3, 2.0 / r_e, 4.0 / r_e, 6.0 / r_e,
                     c1, c2, c3])


def solve_grid(fun, p, t0, T, x, y, rtol=1e-6, atol=1e-8):
    out = np.zeros((len(x), len(y), 2))
    s0, s1 = p[0] * t0, p[0] * (t0 + T)
    for i, xi in enumerate(x):
        for j, yj in enumerate(y):
            r = ode(fun).set_integrator("dop853", rtol=rtol, atol=atol, nsteps=100000, dfactor=0.333)
            r.set_initial_value([xi, yj], s0)
            out[i, j] = r.integrate(s1)
            assert r.successful()
    return out


def main():
    cases = {}
    # reference fixture world: 21x11 double gyre, T=+8 (fm.npy)
    x, y = np.linspace(0, 2, 21), np.linspace(0, 1, 11)
    cases["dg_fwd8"] = (x, y, 0.0, 8.0, dg_params(1.0), "dg")
    # double gyre 401x201, T=-10, every 16th particle
    x, y = np.linspace(0, 2, 401)[::16], np.linspace(0, 1, 201)[::16]
    cases["dg_c1_sub"] = (x, y, 0.0, -10.0, dg_params(-1.0), "dg")
    # double gyre 2001x1001, T=-20, every 125th particle
    x, y = np.linspace(0, 2, 2001)[::125], np.linspace(0, 1, 1001)[::125]
    cases["dg_t20_sub"] = (x, y, 0.0, -20.0, dg_params(-1.0), "dg")
    # Bickley jet 2001x801, T=-40 days, every 125th/100th particle
    r_e = 6371.0e-3
    x, y = np.linspace(0, math.pi * r_e, 2001)[::125], np.linspace(-3, 3, 801)[::100]
    cases["bickley_t40_sub"] = (x, y, 0.0, -40.0, bickley_params(-1.0), "bickley")
    d = {}
    for name, (x, y, t0, T, p, kind) in cases.items():
        fun = dg_velocity(p) if kind == "dg" else bickley_velocity(p)
        d[name + "_x"] = x
        d[name + "_y"] = y
        d[name + "_t0T"] = np.array([t0, T])
        d[name + "_params"] = p
        d[name + "_fm"] = solve_grid(fun, p, t0, T, x, y)
        print(name, d[name + "_fm"].shape)
    np.savez(OUT, **d)


if __name__ == "__main__":
    main()
