// The native flow handle behind the C ABI and the host helpers that turn it
// into a device functor; shared by the single-grid and batched launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "flows.h"

struct nbx_flow {
    int kind;
    // gridded flows
    double2* coef = nullptr;  // spline: (nt+2, nx+2, ny+2) interleaved (u, v) coefficients;
                              // trilinear: (nt, nx, ny) interleaved samples; device
    float4* coef32 = nullptr; // FP32 copy, time-paired, for the FP32-displacement path
    int64_t ncoef = 0;
    double grid9[9] = {0};
    int spherical = 0;
    int extrap = 0;
    double r = 6371.0;
};

namespace nbx {

// Launch shape per (flow kind, precision), set by the autotuner through
// nbx_set_launch_config:
//   lanes       lanes per particle for spline flows (1, 2 or 4);
//   block       threads per block (64, 128 or 256);
//   persistent  0: one group per particle, one launch covering all of them;
//               n > 0: a persistent grid of n blocks per CU pulling
//               particles from a device work queue.
struct LaunchCfg {
    int lanes;
    int block;
    int persistent;
};

LaunchCfg launch_cfg(int kind, int precision);

// L: compile the linear-extrapolation (gradient) path into the functor; the
// dispatchers set it exactly when the flow asks for extrap 1
template <class R, bool L>
Spline3<R, 1, L> make_spline(const nbx_flow* fl, double p0) {
    Spline3<R, 1, L> s;
    if constexpr (sizeof(R) == 8) s.C = fl->coef;
    else s.C = fl->coef32;
    // 16-byte elements either way; 0 (never used) past 4 GiB
    const int64_t bytes = fl->ncoef * 16;
    s.cbytes = bytes < (int64_t(1) << 32) ? static_cast<unsigned>(bytes) : 0u;
    for (int d = 0; d < 3; ++d) {
        s.lo[d] = fl->grid9[3 * d];
        s.hi[d] = fl->grid9[3 * d + 1];
        const int n = static_cast<int>(fl->grid9[3 * d + 2]);
        s.dinv[d] = (n - 1) / (s.hi[d] - s.lo[d]);
        s.nm2[d] = n - 2;
    }
    const int nx = static_cast<int>(fl->grid9[5]), ny = static_cast<int>(fl->grid9[8]);
    s.sx = ny + 2;
    s.st = (nx + 2) * (ny + 2);
    s.spherical = fl->spherical;
    s.extrap = fl->extrap;
    s.r = fl->r;
    s.p0 = p0;
    s.kdeg = p0 * (180 / (kPi * fl->r));
    return s;
}

template <class R>
Linear3<R> make_linear(const nbx_flow* fl, double p0) {
    Linear3<R> s;
    s.V = fl->coef;
    for (int d = 0; d < 3; ++d) {
        s.lo[d] = fl->grid9[3 * d];
        s.hi[d] = fl->grid9[3 * d + 1];
        s.n[d] = static_cast<int>(fl->grid9[3 * d + 2]);
        s.step[d] = (s.hi[d] - s.lo[d]) / (s.n[d] - 1);
    }
    s.sx = s.n[2];
    s.st = s.n[1] * s.n[2];
    s.spherical = fl->spherical;
    s.extrap = fl->extrap;
    s.r = fl->r;
    s.p0 = p0;
    return s;
}

template <int G, class R, bool S = false>
Spline3<R, G, false, S> regroup(const Spline3<R, 1, false>& a) {
    Spline3<R, G, false, S> s;
    s.C = a.C;
    for (int d = 0; d < 3; ++d) {
        s.lo[d] = a.lo[d];
        s.hi[d] = a.hi[d];
        s.dinv[d] = a.dinv[d];
        s.nm2[d] = a.nm2[d];
    }
    s.sx = a.sx;
    s.st = a.st;
    s.cbytes = a.cbytes;
    s.spherical = a.spherical;
    s.extrap = a.extrap;
    s.r = a.r;
    s.p0 = a.p0;
    s.kdeg = a.kdeg;
    return s;
}

}  // namespace nbx
