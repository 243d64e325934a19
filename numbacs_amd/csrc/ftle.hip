// FTLE and Cauchy-Green eigen stencils over a flow map.
//
// Reference: ftle_grid_2D (diagnostics.py:21-65) with gradF_stencil_2D
// (utils.py:9-46) and eigvalsh_max_2D (utils.py:168-189); C_eig_2D
// (diagnostics.py:200-244) and ftle_from_eig (diagnostics.py:247-269).
//
// Layout: the flow map is (nx, ny) double2 in HBM, y fastest.  A 64x4 thread
// block covers 64 consecutive y-points of 4 consecutive rows; the +-1 row and
// column neighbours it reads are shared with the adjacent lanes/blocks and
// come from L1/L2, so HBM sees each flow-map element about once: 16 B read +
// 8 B written per point for FTLE (24 B), 40 B per point for FTLE + eigvec_max.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "stencil.h"

namespace nbx {

__global__ __launch_bounds__(256) void k_ftle(const double2* __restrict__ F, int nx, int ny,
                                              double scaling, double dx, double dy,
                                              const uint8_t* __restrict__ mask,
                                              double* __restrict__ ftle) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long k = static_cast<long>(i) * ny + j;
    double out = 0.0;
    if (interior(i, j, nx, ny) && !(mask != nullptr && mask[k])) {
        const Grad g = grad_at(F, k, ny, dx, dy);
        const double off = g.dxdx * g.dxdy + g.dydx * g.dydy;
        const double a = g.dxdx * g.dxdx + g.dydx * g.dydx;
        const double d = g.dxdy * g.dxdy + g.dydy * g.dydy;
        const double trace = a + d;
        const double disc = sqrt((a - d) * (a - d) + 4 * (off * off));
        const double lmax = 0.5 * (trace + disc);
        if (lmax > 1) out = scaling * log(lmax);
    }
    ftle[k] = out;
}

__global__ __launch_bounds__(256) void k_C_eig(const double2* __restrict__ F, int nx, int ny,
                                               double dx, double dy,
                                               const uint8_t* __restrict__ mask,
                                               double2* __restrict__ evals,
                                               double4* __restrict__ evecs) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long k = static_cast<long>(i) * ny + j;
    double ev[2] = {0.0, 0.0}, V[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (interior(i, j, nx, ny) && !(mask != nullptr && mask[k])) {
        const Grad g = grad_at(F, k, ny, dx, dy);
        const double off = g.dxdx * g.dxdy + g.dydx * g.dydy;
        const double a = g.dxdx * g.dxdx + g.dydx * g.dydx;
        const double d = g.dxdy * g.dxdy + g.dydy * g.dydy;
        eig2_sym(a, off, d, ev, V);
    }
    evals[k] = make_double2(ev[0], ev[1]);
    evecs[k] = make_double4(V[0][0], V[0][1], V[1][0], V[1][1]);
}

// C_eig_2D + ftle_from_eig fused: the ridge path's FTLE comes from the
// eigen-solver's largest eigenvalue (diagnostics.py:266-269).
__global__ __launch_bounds__(256) void k_ftle_eig(const double2* __restrict__ F, int nx, int ny,
                                                  double two_abs_T, double dx, double dy,
                                                  const uint8_t* __restrict__ mask,
                                                  double* __restrict__ ftle,
                                                  double2* __restrict__ evec_max) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long k = static_cast<long>(i) * ny + j;
    double ev[2] = {0.0, 0.0}, V[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (interior(i, j, nx, ny) && !(mask != nullptr && mask[k])) {
        const Grad g = grad_at(F, k, ny, dx, dy);
        const double off = g.dxdx * g.dxdy + g.dydx * g.dydy;
        const double a = g.dxdx * g.dxdx + g.dydx * g.dydx;
        const double d = g.dxdy * g.dxdy + g.dydy * g.dydy;
        eig2_sym(a, off, d, ev, V);
    }
    ftle[k] = ev[1] > 1 ? log(ev[1]) / two_abs_T : 0.0;
    evec_max[k] = make_double2(V[0][1], V[1][1]);
}

}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_ftle_grid_2d(const double* fm, int nx, int ny, double T, double dx, double dy,
                     const uint8_t* mask, double* ftle, nbx_stream stream) {
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    if (static_cast<int64_t>(nx) * ny == 0) return 0;
    NBX_REQUIRE(fm != nullptr && ftle != nullptr, "NULL array");
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_ftle<<<grid, block, 0, as_stream(stream)>>>(reinterpret_cast<const double2*>(fm), nx, ny,
                                                  1 / (2 * fabs(T)), dx, dy, mask, ftle);
    NBX_LAUNCHED("ftle");
    return 0;
}

int nbx_C_eig_2d(const double* fm, int nx, int ny, double dx, double dy, const uint8_t* mask,
                 double* eigvals, double* eigvecs, nbx_stream stream) {
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    if (static_cast<int64_t>(nx) * ny == 0) return 0;
    NBX_REQUIRE(fm != nullptr && eigvals != nullptr && eigvecs != nullptr, "NULL array");
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_C_eig<<<grid, block, 0, as_stream(stream)>>>(reinterpret_cast<const double2*>(fm), nx, ny,
                                                   dx, dy, mask,
                                                   reinterpret_cast<double2*>(eigvals),
                                                   reinterpret_cast<double4*>(eigvecs));
    NBX_LAUNCHED("C_eig");
    return 0;
}

int nbx_ftle_eig_2d(const double* fm, int nx, int ny, double T, double dx, double dy,
                    const uint8_t* mask, double* ftle, double* eigvec_max, nbx_stream stream) {
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    if (static_cast<int64_t>(nx) * ny == 0) return 0;
    NBX_REQUIRE(fm != nullptr && ftle != nullptr && eigvec_max != nullptr, "NULL array");
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_ftle_eig<<<grid, block, 0, as_stream(stream)>>>(
        reinterpret_cast<const double2*>(fm), nx, ny, 2 * fabs(T), dx, dy, mask, ftle,
        reinterpret_cast<double2*>(eigvec_max));
    NBX_LAUNCHED("ftle_eig");
    return 0;
}

}  // extern "C"
