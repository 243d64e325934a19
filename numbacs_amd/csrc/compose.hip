// Flow-map composition (integration.py:609-737 of the reference): the flow
// map over [t0, t0 + nT h] as the chain of nT short flow maps over
// [t0 + k h, t0 + (k+1) h], each sampled at the previous map's image with
// bilinear interpolation on the particle grid (eval_linear, CONSTANT
// extrapolation: 0 outside the closed grid box -- five double-gyre
// trajectories of the reference's composition fixture end 1.8e-15 beyond
// x = 2 and the fixture holds 0 there).
//
// Layout: nslots slots of (nx, ny) double2 flow maps; map k of the chain
// lives in slot (first + k) % nslots, so a sliding window over t0 replaces
// one slot per step instead of shifting the others.  One thread per
// particle walks the chain: nT-1 bilinear lookups of 4 double2 (64 B)
// each, served by L2 for grids up to a few MB per map.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"

namespace nbx {

struct Box2 {
    double x0, x1, y0, y1, dx, dy;  // dx = (x1 - x0) / (nx - 1), as eval_linear's grid step
    int nx, ny;
};

// Both components of one flow map at (qx, qy).  The arithmetic follows the
// oracle's source order with contraction off, so the two agree to the bit.
__device__ __forceinline__ double2 bilinear0(const double2* __restrict__ F, const Box2& b,
                                             double qx, double qy) {
#pragma clang fp contract(off)
    if (qx < b.x0 || qx > b.x1 || qy < b.y0 || qy > b.y1) return make_double2(0.0, 0.0);
    const double ux = (qx - b.x0) / b.dx;
    const double uy = (qy - b.y0) / b.dy;
    int i = static_cast<int>(floor(ux)), j = static_cast<int>(floor(uy));
    i = i > b.nx - 2 ? b.nx - 2 : i;
    j = j > b.ny - 2 ? b.ny - 2 : j;
    i = i < 0 ? 0 : i;
    j = j < 0 ? 0 : j;
    const double lx = ux - i, ly = uy - j;
    const long k = static_cast<long>(i) * b.ny + j;
    const double2 v00 = F[k], v01 = F[k + 1], v10 = F[k + b.ny], v11 = F[k + b.ny + 1];
    double2 r;
    r.x = (1 - lx) * ((1 - ly) * v00.x + ly * v01.x) + lx * ((1 - ly) * v10.x + ly * v11.x);
    r.y = (1 - lx) * ((1 - ly) * v00.y + ly * v01.y) + lx * ((1 - ly) * v10.y + ly * v11.y);
    return r;
}

__global__ __launch_bounds__(256) void k_compose(const double2* __restrict__ maps, int nslots,
                                                 int first, int nT, Box2 b,
                                                 double2* __restrict__ out) {
    const long n = static_cast<long>(b.nx) * b.ny;
    const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (p >= n) return;
    auto slot = [&](int k) { return maps + n * ((first + k) % nslots); };
    double2 q = slot(0)[p];
    for (int k = 1; k < nT - 1; ++k) q = bilinear0(slot(k), b, q.x, q.y);
    out[p] = bilinear0(slot(nT - 1), b, q.x, q.y);
}

}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_flowmap_compose(const double* flowmaps, int nslots, int first, int nT, double x0,
                        double x1, int nx, double y0, double y1, int ny, double* out,
                        nbx_stream stream) {
    NBX_REQUIRE(nx >= 2 && ny >= 2, "the composition grid needs at least 2 points per axis");
    NBX_REQUIRE(x1 > x0 && y1 > y0, "grid axes must be ascending");
    NBX_REQUIRE(nT >= 1 && nslots >= nT, "need 1 <= nT <= nslots");
    NBX_REQUIRE(first >= 0 && first < nslots, "first slot out of range");
    NBX_REQUIRE(flowmaps != nullptr && out != nullptr, "NULL array");
    const Box2 b{x0, x1, y0, y1, (x1 - x0) / (nx - 1), (y1 - y0) / (ny - 1), nx, ny};
    const long n = static_cast<long>(nx) * ny;
    k_compose<<<static_cast<unsigned>((n + 255) / 256), 256, 0, as_stream(stream)>>>(
        reinterpret_cast<const double2*>(flowmaps), nslots, first, nT, b,
        reinterpret_cast<double2*>(out));
    NBX_LAUNCHED("compose");
    return 0;
}

}  // extern "C"
