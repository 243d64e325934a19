// FTLE and Cauchy-Green eigen stencils over a flow map.
//
// Reference: ftle_grid_2D (diagnostics.py:21-65) with gradF_stencil_2D
// (utils.py:9-46) and eigvalsh_max_2D (utils.py:168-189); C_eig_2D
// (diagnostics.py:200-244) and ftle_from_eig (diagnostics.py:247-269);
// C_tensor_2D and C_eig_aux_2D (diagnostics.py:67-197) on the aux grid.
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

// Row strips: one wavefront walks down R rows of a 62-column strip (lane l
// holds column j0 - 1 + l; lanes 0 and 63 are the strip's halo), keeping
// the rows above and below its point in registers and taking the column
// neighbours from the adjacent lanes, so each flow-map element is read from
// memory once per strip (plus one halo row at each end) -- the 16 B per point
// of the algorithmic count -- where 64 x 4 blocks re-read the rows shared
// with the blocks above and below.  Gradient as gradF_stencil_2D
// (utils.py:41-44, the e/w/n/s of grad_at): results equal the point form's.
template <bool EIG>
__global__ __launch_bounds__(64) void k_ftle_strip(const double2* __restrict__ F, int nx, int ny,
                                                   int R, double tscale, double dx, double dy,
                                                   const uint8_t* __restrict__ mask,
                                                   double* __restrict__ ftle,
                                                   double2* __restrict__ evec) {
    const int lane = static_cast<int>(threadIdx.x);
    const int j = static_cast<int>(blockIdx.x) * 62 - 1 + lane;
    const int i0 = static_cast<int>(blockIdx.y) * R;
    const int i1 = i0 + R < nx ? i0 + R : nx;
    const bool jin = j >= 0 && j < ny;
    const bool writer = lane >= 1 && lane <= 62 && jin;
    auto row = [&](int i) {
        return (jin && i >= 0 && i < nx) ? F[static_cast<long>(i) * ny + j]
                                         : make_double2(0.0, 0.0);
    };
    // the point of row i from its rows above and below and its own row
    // (column neighbours from the adjacent lanes; every lane shuffles)
    auto emit = [&](int i, const double2& up, const double2& cur, const double2& down) {
        double2 n, s;
        n.x = __shfl_down(cur.x, 1);
        n.y = __shfl_down(cur.y, 1);
        s.x = __shfl_up(cur.x, 1);
        s.y = __shfl_up(cur.y, 1);
        if (writer) {
            const long k = static_cast<long>(i) * ny + j;
            const bool in = interior(i, j, nx, ny) && !(mask != nullptr && mask[k]);
            Grad g;
            g.dxdx = (down.x - up.x) / (2 * dx);
            g.dxdy = (n.x - s.x) / (2 * dy);
            g.dydx = (down.y - up.y) / (2 * dx);
            g.dydy = (n.y - s.y) / (2 * dy);
            if constexpr (EIG) {
                double f = 0.0;
                double2 e = make_double2(0.0, 0.0);
                if (in) ftle_eig_of(g, tscale, f, e);
                ftle[k] = f;
                evec[k] = e;
            } else {
                ftle[k] = in ? ftle_of(g, tscale) : 0.0;
            }
        }
    };
    double2 up = row(i0 - 1), cur = row(i0);
    int i = i0;
    // four rows a pass: their loads are in flight together
    for (; i + 4 <= i1; i += 4) {
        const double2 d0 = row(i + 1), d1 = row(i + 2), d2 = row(i + 3), d3 = row(i + 4);
        emit(i, up, cur, d0);
        emit(i + 1, cur, d0, d1);
        emit(i + 2, d0, d1, d2);
        emit(i + 3, d1, d2, d3);
        up = d2;
        cur = d3;
    }
    for (; i < i1; ++i) {
        const double2 down = row(i + 1);
        emit(i, up, cur, down);
        up = cur;
        cur = down;
    }
}

// Rows per strip: about 32768 wavefronts (four per SIMD slot: more rows in
// flight across the machine; 16-row strips on an 8001 x 4001 map ran 3%
// faster than 64-row ones), 4 to 64 rows (one halo row re-read per strip).
inline int strip_rows(int nx, int ny) {
    const long cols = (ny + 61) / 62;
    long r = (static_cast<long>(nx) * cols + 32767) / 32768;
    r = r < 4 ? 4 : (r > 64 ? 64 : r);
    return static_cast<int>(r);
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

// Auxiliary-grid stencils: F is (nx, ny, naux) double2, the particles of
// each grid point started at offsets (h,0), (-h,0), (0,h), (0,-h) and, for
// naux = 5, (0,0).  gradF_aux_stencil_2D (utils.py:50-84).
__device__ __forceinline__ Grad grad_aux(const double2* __restrict__ F, long c, int naux,
                                         double h) {
    const double2* p = F + c * naux;
    const double2 a = p[0], b = p[1], u = p[2], w = p[3];
    Grad g;
    g.dxdx = (a.x - b.x) / (2 * h);
    g.dxdy = (u.x - w.x) / (2 * h);
    g.dydx = (a.y - b.y) / (2 * h);
    g.dydy = (u.y - w.y) / (2 * h);
    return g;
}

// C_tensor_2D (diagnostics.py:67-110): (C11, C12, C22) on i, j in [2, n-3]
__global__ __launch_bounds__(256) void k_C_tensor_aux(const double2* __restrict__ F, int nx,
                                                      int ny, int naux, double h,
                                                      const uint8_t* __restrict__ mask,
                                                      double* __restrict__ C) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long c = static_cast<long>(i) * ny + j;
    double c11 = 0, c12 = 0, c22 = 0;
    if (i >= 2 && i < nx - 2 && j >= 2 && j < ny - 2 && !(mask != nullptr && mask[c])) {
        const Grad g = grad_aux(F, c, naux, h);
        c11 = g.dxdx * g.dxdx + g.dydx * g.dydx;
        c12 = g.dxdx * g.dxdy + g.dydx * g.dydy;
        c22 = g.dxdy * g.dxdy + g.dydy * g.dydy;
    }
    C[3 * c] = c11;
    C[3 * c + 1] = c12;
    C[3 * c + 2] = c22;
}

// C_eig_aux_2D (diagnostics.py:113-197).  eig_main: on i, j in [2, n-3],
// eigenvectors of the aux-grid tensor and eigenvalues of the main-grid
// tensor (gradF_main_stencil_2D, utils.py:88-124: central differences of
// the neighbours' centre particles); otherwise both from the aux-grid
// tensor on i, j in [1, n-2].
__global__ __launch_bounds__(256) void k_C_eig_aux(const double2* __restrict__ F, int nx, int ny,
                                                   int naux, double dx, double dy, double h,
                                                   int eig_main,
                                                   const uint8_t* __restrict__ mask,
                                                   double2* __restrict__ evals,
                                                   double4* __restrict__ evecs) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long c = static_cast<long>(i) * ny + j;
    const int lo = eig_main ? 2 : 1;
    double ev[2] = {0.0, 0.0}, V[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (i >= lo && i < nx - lo && j >= lo && j < ny - lo && !(mask != nullptr && mask[c])) {
        const Grad g = grad_aux(F, c, naux, h);
        eig2_sym(g.dxdx * g.dxdx + g.dydx * g.dydx, g.dxdx * g.dxdy + g.dydx * g.dydy,
                 g.dxdy * g.dxdy + g.dydy * g.dydy, ev, V);
        if (eig_main) {
            const int m = naux - 1;
            const double2 e = F[(c + ny) * naux + m], w = F[(c - ny) * naux + m];
            const double2 n = F[(c + 1) * naux + m], s = F[(c - 1) * naux + m];
            const double dxdx = (e.x - w.x) / (2 * dx), dxdy = (n.x - s.x) / (2 * dy);
            const double dydx = (e.y - w.y) / (2 * dx), dydy = (n.y - s.y) / (2 * dy);
            double ev2[2], V2[2][2];
            eig2_sym(dxdx * dxdx + dydx * dydx, dxdx * dxdy + dydx * dydy,
                     dxdy * dxdy + dydy * dydy, ev2, V2);
            ev[0] = ev2[0];
            ev[1] = ev2[1];
        }
    }
    evals[c] = make_double2(ev[0], ev[1]);
    evecs[c] = make_double4(V[0][0], V[0][1], V[1][0], V[1][1]);
}

}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_ftle_grid_2d(const double* fm, int nx, int ny, double T, double dx, double dy,
                     const uint8_t* mask, double* ftle, nbx_stream stream) {
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    if (static_cast<int64_t>(nx) * ny == 0) return 0;
    NBX_REQUIRE(fm != nullptr && ftle != nullptr, "NULL array");
    const int R = strip_rows(nx, ny);
    const dim3 grid((ny + 61) / 62, (nx + R - 1) / R);
    k_ftle_strip<false><<<grid, 64, 0, as_stream(stream)>>>(
        reinterpret_cast<const double2*>(fm), nx, ny, R, 1 / (2 * fabs(T)), dx, dy, mask, ftle,
        nullptr);
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
    const int R = strip_rows(nx, ny);
    const dim3 grid((ny + 61) / 62, (nx + R - 1) / R);
    k_ftle_strip<true><<<grid, 64, 0, as_stream(stream)>>>(
        reinterpret_cast<const double2*>(fm), nx, ny, R, 2 * fabs(T), dx, dy, mask, ftle,
        reinterpret_cast<double2*>(eigvec_max));
    NBX_LAUNCHED("ftle_eig");
    return 0;
}

int nbx_C_tensor_aux_2d(const double* fm_aux, int nx, int ny, int naux, double h,
                        const uint8_t* mask, double* C, nbx_stream stream) {
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    NBX_REQUIRE(naux == 4 || naux == 5, "naux must be 4 or 5");
    if (static_cast<int64_t>(nx) * ny == 0) return 0;
    NBX_REQUIRE(fm_aux != nullptr && C != nullptr, "NULL array");
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_C_tensor_aux<<<grid, block, 0, as_stream(stream)>>>(
        reinterpret_cast<const double2*>(fm_aux), nx, ny, naux, h, mask, C);
    NBX_LAUNCHED("C_tensor_aux");
    return 0;
}

int nbx_C_eig_aux_2d(const double* fm_aux, int nx, int ny, int naux, double dx, double dy,
                     double h, int eig_main, const uint8_t* mask, double* eigvals,
                     double* eigvecs, nbx_stream stream) {
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    NBX_REQUIRE(naux == 4 || naux == 5, "naux must be 4 or 5");
    NBX_REQUIRE(!eig_main || naux == 5, "eig_main needs the centre particle (naux = 5)");
    if (static_cast<int64_t>(nx) * ny == 0) return 0;
    NBX_REQUIRE(fm_aux != nullptr && eigvals != nullptr && eigvecs != nullptr, "NULL array");
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_C_eig_aux<<<grid, block, 0, as_stream(stream)>>>(
        reinterpret_cast<const double2*>(fm_aux), nx, ny, naux, dx, dy, h, eig_main ? 1 : 0, mask,
        reinterpret_cast<double2*>(eigvals), reinterpret_cast<double4*>(eigvecs));
    NBX_LAUNCHED("C_eig_aux");
    return 0;
}

}  // extern "C"
