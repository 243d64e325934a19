// Device velocity fields ("flows").  Each flow is a small trivially-copyable
// functor passed by value in the kernel arguments and inlined into the
// integrator; it evaluates dy = p0 * v(p0 * t, y), the sign-flipped
// right-hand side convention of the reference (params[0] = integration
// direction, flows.py:1146-1158, docs userguide.rst:216-227).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace nbx {

constexpr double kPi = 3.141592653589793;

// Double gyre, flows.py:1146-1158.  p = [dir, A, eps, alpha, omega, psi].
struct DoubleGyre {
    double p0, A, eps, alpha, omega, psi;

    __device__ __forceinline__ void operator()(double t, double y0, double y1, double& f0,
                                               double& f1) const {
        const double tt = p0 * t;
        const double a = eps * sin(omega * tt + psi);
        const double b = 1 - 2 * a;
        const double f = a * (y0 * y0) + b * y0;
        const double df = 2 * a * y0 + b;
        double sf, cf, sy, cy;
        sincos(kPi * f, &sf, &cf);
        sincos(kPi * y1, &sy, &cy);
        f0 = p0 * (-kPi * A * sf * cy - alpha * y0);
        f1 = p0 * (kPi * A * cf * sy * df - alpha * y1);
    }
};

// Bickley jet with three travelling modes, flows.py:1182-1213.
// p = [dir, U0, L, A1, A2, A3, k1, k2, k3, c1, c2, c3].
struct BickleyJet {
    double p[12];

    __device__ __forceinline__ void operator()(double t, double y0, double y1, double& f0,
                                               double& f1) const {
        const double tt = p[0] * t;
        const double Y = y1 / p[2];
        const double ch = cosh(Y);
        const double sech2 = 1 / (ch * ch);
        double s1, c1, s2, c2, s3, c3;
        sincos(p[6] * (y0 - p[9] * tt), &s1, &c1);
        sincos(p[7] * (y0 - p[10] * tt), &s2, &c2);
        sincos(p[8] * (y0 - p[11] * tt), &s3, &c3);
        f0 = p[0] * (p[1] * sech2 + 2 * p[1] * tanh(Y) * sech2 * (p[3] * c1 + p[4] * c2 + p[5] * c3));
        f1 = -p[0] * (p[1] * p[2] * sech2 *
                      (p[3] * p[6] * s1 + p[4] * p[7] * s2 + p[5] * p[8] * s3));
    }
};

// Uniform cubic B-spline basis at local coordinate s in [0, 1].
template <class R>
__device__ __forceinline__ void bspline_weights(R s, R w[4]) {
    const R s2 = s * s, s3 = s2 * s, u = R(1) - s;
    w[0] = u * u * u / R(6);
    w[1] = (R(3) * s3 - R(6) * s2 + R(4)) / R(6);
    w[2] = (R(-3) * s3 + R(3) * s2 + R(3) * s + R(1)) / R(6);
    w[3] = s3 / R(6);
}

template <class R>
__device__ __forceinline__ void bspline_dweights(R s, R dw[4]) {
    const R s2 = s * s, u = R(1) - s;
    dw[0] = -u * u / R(2);
    dw[1] = (R(3) * s2 - R(4) * s) / R(2);
    dw[2] = (R(-3) * s2 + R(2) * s + R(1)) / R(2);
    dw[3] = s2 / R(2);
}

// Python's float modulo (result takes the sign of the divisor).
__device__ __forceinline__ double py_mod(double a, double m) {
    double r = fmod(a, m);
    if (r != 0.0) {
        if ((m < 0) != (r < 0)) r += m;
    } else {
        r = copysign(0.0, m);
    }
    return r;
}

// Coefficient storage element: (u, v) interleaved so one gather serves both
// velocity components.
template <class R> struct CoefPair;
template <> struct CoefPair<double> { using type = double2; };
template <> struct CoefPair<float> { using type = float2; };

// Tricubic natural B-spline velocity on a uniform (t, x, y) grid (the
// reference's get_flow_2D over interpolation.eval_spline, flows.py:121-258).
// C holds (nt+2, nx+2, ny+2) (u, v) pairs, y fastest.  R is the storage and
// evaluation precision of the coefficients (double for the plain path).
//
// Out-of-box queries (parity unpinned: the reference's dependency is not
// available): extrap 0/2 clamp the query into the box per axis (value of the
// nearest box point), extrap 1 adds the gradient at that point times the
// outward offset.
template <class R>
struct Spline3 {
    using P = typename CoefPair<R>::type;
    const P* __restrict__ C;
    double lo[3], hi[3], dinv[3];
    int nm2[3];   // n - 2 per axis: the largest cell index
    int sx, st;   // strides (in pairs) of the x and t axes
    int spherical, extrap;
    double r;     // sphere radius (km) for spherical = 1, 2
    double p0;    // integration direction

    __device__ __forceinline__ void cell(int d, double q, int& i, R& s) const {
        const double u = (q - lo[d]) * dinv[d];
        int k = static_cast<int>(floor(u));
        k = k > nm2[d] ? nm2[d] : k;
        k = k < 0 ? 0 : k;
        i = k;
        s = static_cast<R>(u - k);
    }

    // (u, v) at an in-box point; with grad != nullptr also d/dt, d/dx, d/dy.
    __device__ __forceinline__ void eval_in(double qt, double qx, double qy, R& su, R& sv,
                                            R (*grad)[2]) const {
        int it, ix, iy;
        R s0, s1, s2;
        cell(0, qt, it, s0);
        cell(1, qx, ix, s1);
        cell(2, qy, iy, s2);
        R wt[4], wx[4], wy[4];
        bspline_weights(s0, wt);
        bspline_weights(s1, wx);
        bspline_weights(s2, wy);
        const P* base = C + (static_cast<long>(it) * st + static_cast<long>(ix) * sx + iy);
        R au = 0, av = 0;
        if (grad == nullptr) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const P* row = base + (a * st + b * sx);
                    const P c0 = row[0], c1 = row[1], c2 = row[2], c3 = row[3];
                    R ru = wy[0] * c0.x;
                    R rv = wy[0] * c0.y;
                    ru += wy[1] * c1.x;
                    rv += wy[1] * c1.y;
                    ru += wy[2] * c2.x;
                    rv += wy[2] * c2.y;
                    ru += wy[3] * c3.x;
                    rv += wy[3] * c3.y;
                    const R wab = wt[a] * wx[b];
                    au += wab * ru;
                    av += wab * rv;
                }
            }
        } else {
            R dt[4], dx[4], dy[4];
            bspline_dweights(s0, dt);
            bspline_dweights(s1, dx);
            bspline_dweights(s2, dy);
            R g[3][2] = {{0, 0}, {0, 0}, {0, 0}};
            for (int a = 0; a < 4; ++a) {
                for (int b = 0; b < 4; ++b) {
                    const P* row = base + (a * st + b * sx);
                    R ru = 0, rv = 0, du = 0, dv = 0;
                    for (int c = 0; c < 4; ++c) {
                        const P cc = row[c];
                        ru += wy[c] * cc.x;
                        rv += wy[c] * cc.y;
                        du += dy[c] * cc.x;
                        dv += dy[c] * cc.y;
                    }
                    au += wt[a] * wx[b] * ru;
                    av += wt[a] * wx[b] * rv;
                    g[0][0] += dt[a] * wx[b] * ru;
                    g[0][1] += dt[a] * wx[b] * rv;
                    g[1][0] += wt[a] * dx[b] * ru;
                    g[1][1] += wt[a] * dx[b] * rv;
                    g[2][0] += wt[a] * wx[b] * du;
                    g[2][1] += wt[a] * wx[b] * dv;
                }
            }
            for (int d = 0; d < 3; ++d) {
                grad[d][0] = g[d][0] * static_cast<R>(dinv[d]);
                grad[d][1] = g[d][1] * static_cast<R>(dinv[d]);
            }
        }
        su = au;
        sv = av;
    }

    __device__ __forceinline__ void eval(double qt, double qx, double qy, R& su, R& sv) const {
        const double ct = fmin(fmax(qt, lo[0]), hi[0]);
        const double cx = fmin(fmax(qx, lo[1]), hi[1]);
        const double cy = fmin(fmax(qy, lo[2]), hi[2]);
        const bool inside = (ct == qt) & (cx == qx) & (cy == qy);
        if (inside || extrap != 1) {
            eval_in(ct, cx, cy, su, sv, nullptr);
        } else {
            R g[3][2];
            eval_in(ct, cx, cy, su, sv, g);
            su += g[0][0] * static_cast<R>(qt - ct) + g[1][0] * static_cast<R>(qx - cx) +
                  g[2][0] * static_cast<R>(qy - cy);
            sv += g[0][1] * static_cast<R>(qt - ct) + g[1][1] * static_cast<R>(qx - cx) +
                  g[2][1] * static_cast<R>(qy - cy);
        }
    }

    __device__ __forceinline__ void operator()(double t, double y0, double y1, double& f0,
                                               double& f1) const {
        const double tt = p0 * t;
        R su, sv;
        if (spherical == 0) {
            eval(tt, y0, y1, su, sv);
            f0 = p0 * static_cast<double>(su);
            f1 = p0 * static_cast<double>(sv);
        } else {
            const double xx = spherical == 1 ? py_mod(y0 - 180, 360) - 180 : py_mod(y0, 360);
            eval(tt, xx, y1, su, sv);
            f0 = ((p0 * static_cast<double>(su)) * 180) / (kPi * r * cos(y1 * kPi / 180));
            f1 = ((p0 * static_cast<double>(sv)) * 180) / (kPi * r);
        }
    }
};

}  // namespace nbx
