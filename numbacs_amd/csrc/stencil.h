// Point stencils shared by the FTLE launches: central-difference flow-map
// gradient (utils.py:41-44 of the reference), the Cauchy-Green tensor and its
// symmetric 2x2 eigen-decomposition.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace nbx {

struct Grad {
    double dxdx, dxdy, dydx, dydy;
};

__device__ __forceinline__ Grad grad_at(const double2* __restrict__ F, long k, int ny, double dx,
                                        double dy) {
    const double2 e = F[k + ny], w = F[k - ny], n = F[k + 1], s = F[k - 1];
    Grad g;
    g.dxdx = (e.x - w.x) / (2 * dx);
    g.dxdy = (n.x - s.x) / (2 * dy);
    g.dydx = (e.y - w.y) / (2 * dx);
    g.dydy = (n.y - s.y) / (2 * dy);
    return g;
}

// Symmetric 2x2 [[a, b], [b, d]]: ascending eigenvalues and eigenvectors as
// columns.  lmax = (a+d)/2 + sqrt(((a-d)/2)^2 + b^2); lmin from the
// determinant when it is small relative to lmax; the eigenvector of lmax is
// taken orthogonal to the larger row of (C - lmax I).
__device__ __forceinline__ void eig2_sym(double a, double b, double d, double ev[2],
                                         double V[2][2]) {
    const double h = 0.5 * (a - d);
    const double rad = sqrt(h * h + b * b);
    const double m = 0.5 * (a + d);
    const double lmax = m + rad;
    double lmin = m - rad;
    if (lmax != 0.0) {
        const double alt = (a * d - b * b) / lmax;
        if (fabs(lmin) < 0.5 * fabs(lmax)) lmin = alt;
    }
    ev[0] = lmin;
    ev[1] = lmax;
    double ex, ey;
    if (b == 0.0) {
        if (a > d) { ex = 1.0; ey = 0.0; }
        else { ex = 0.0; ey = 1.0; }
    } else {
        const double r1x = b, r1y = lmax - a;
        const double r2x = lmax - d, r2y = b;
        const double n1 = r1x * r1x + r1y * r1y, n2 = r2x * r2x + r2y * r2y;
        if (n1 >= n2) {
            const double s = 1.0 / sqrt(n1);
            ex = r1x * s;
            ey = r1y * s;
        } else {
            const double s = 1.0 / sqrt(n2);
            ex = r2x * s;
            ey = r2y * s;
        }
    }
    V[0][1] = ex;
    V[1][1] = ey;
    V[0][0] = -ey;
    V[1][0] = ex;
}

__device__ __forceinline__ bool interior(int i, int j, int nx, int ny) {
    return i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1;
}

// FTLE of one interior point from its flow-map gradient (diagnostics.py:
// 56-63: C = gradF^T gradF, its largest eigenvalue, log / (2|T|) if > 1).
__device__ __forceinline__ double ftle_of(const Grad& g, double scaling) {
    const double off = g.dxdx * g.dxdy + g.dydx * g.dydy;
    const double a = g.dxdx * g.dxdx + g.dydx * g.dydx;
    const double d = g.dxdy * g.dxdy + g.dydy * g.dydy;
    const double trace = a + d;
    const double disc = sqrt((a - d) * (a - d) + 4 * (off * off));
    const double lmax = 0.5 * (trace + disc);
    return lmax > 1 ? scaling * log(lmax) : 0.0;
}

// The same point's eigen-decomposition (C_eig_2D) and ftle_from_eig.
__device__ __forceinline__ void ftle_eig_of(const Grad& g, double two_abs_T, double& ftle,
                                            double2& evec) {
    double ev[2], V[2][2];
    const double off = g.dxdx * g.dxdy + g.dydx * g.dydy;
    const double a = g.dxdx * g.dxdx + g.dydx * g.dydx;
    const double d = g.dxdy * g.dxdy + g.dydy * g.dydy;
    eig2_sym(a, off, d, ev, V);
    ftle = ev[1] > 1 ? log(ev[1]) / two_abs_T : 0.0;
    evec = make_double2(V[0][1], V[1][1]);
}

// FTLE at one interior point (diagnostics.py:56-63): C = gradF^T gradF,
// lmax by eigvalsh_max_2D's closed form (utils.py:185-189).
__device__ __forceinline__ double ftle_point(const double2* __restrict__ F, long k, int ny,
                                             double scaling, double dx, double dy) {
    return ftle_of(grad_at(F, k, ny, dx, dy), scaling);
}

// C_eig_2D at one interior point: ascending eigenvalues and eigenvectors.
__device__ __forceinline__ void ceig_point(const double2* __restrict__ F, long k, int ny,
                                           double dx, double dy, double ev[2], double V[2][2]) {
    const Grad g = grad_at(F, k, ny, dx, dy);
    const double off = g.dxdx * g.dxdy + g.dydx * g.dydy;
    const double a = g.dxdx * g.dxdx + g.dydx * g.dydx;
    const double d = g.dxdy * g.dxdy + g.dydy * g.dydy;
    eig2_sym(a, off, d, ev, V);
}

}  // namespace nbx
