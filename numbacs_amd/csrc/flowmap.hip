// Flow handles, spline prefilter and the DOP853 flow-map launches.
//
// Reference path: flowmap_grid_2D (integration.py:123-182) runs one
// numbalsoda.dop853 solve per grid point over a numba cfunc right-hand side.
// Here one GPU thread owns one particle: the right-hand side is a device
// functor inlined into the integrator, the state stays in registers for the
// whole solve, and the final position is written once, coalesced.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <string>
#include <vector>

#include "common.h"
#include "dop853.h"
#include "flow_handle.h"
#include "flows.h"
#include "queue.h"

namespace nbx {

namespace {
thread_local std::string g_last_error;
struct WsEntry {
    std::mutex m;  // held by the Workspace that uses the buffer
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;  // guards the map, not the buffers
std::map<std::tuple<int, hipStream_t, int>, std::unique_ptr<WsEntry>> g_ws;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

Workspace::Workspace(hipStream_t stream, int slot, size_t bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        err_ = -1;
        return;
    }
    WsEntry* e = nullptr;
    {
        std::lock_guard<std::mutex> g(g_ws_mutex);
        auto& slot_entry = g_ws[std::make_tuple(dev, stream, slot)];
        if (!slot_entry) slot_entry = std::make_unique<WsEntry>();
        e = slot_entry.get();
    }
    lock_ = std::unique_lock<std::mutex>(e->m);
    if (e->bytes < bytes) {
        if (e->ptr) {
            // hipFree waits for the device work still using the old buffer
            (void)hipFree(e->ptr);
            e->ptr = nullptr;
            e->bytes = 0;
        }
        const size_t want = bytes < 4096 ? 4096 : bytes;
        if (hipMalloc(&e->ptr, want) != hipSuccess) {
            set_error("workspace hipMalloc failed");
            err_ = -1;
            return;
        }
        e->bytes = want;
    }
    ptr_ = static_cast<char*>(e->ptr);
}

// ---------------------------------------------------------------------------
// Spline prefilter: natural cubic B-spline, one thread per line.
// Per axis n samples -> n+2 coefficients c[-1..n]: (c[i-1] + 4c[i] + c[i+1])/6
// = f[i], zero second derivative at both ends, i.e. c[0] = f[0],
// c[n-1] = f[n-1], interior by the Thomas algorithm, then
// c[-1] = 2c[0]-c[1], c[n] = 2c[n-1]-c[n-2].

__device__ void prefilter_line(double* p, long stride, int n) {
    auto C = [&](int k) -> double& { return p[(static_cast<long>(k) + 1) * stride]; };
    if (n == 1) {
        C(-1) = C(0);
        C(1) = C(0);
        return;
    }
    if (n > 2) {
        // forward sweep; the modified super-diagonal cp[k] = 1/(4 - cp[k-1])
        // is recomputed on the way back instead of stored
        const int m = n - 2;
        double cp = 0.0, prev_d = 0.0;
        for (int k = 0; k < m; ++k) {
            const int i = k + 1;
            double rhs = 6.0 * C(i);
            if (i == 1) rhs -= C(0);
            if (i == n - 2) rhs -= C(n - 1);
            const double denom = 4.0 - (k > 0 ? cp : 0.0);
            cp = 1.0 / denom;
            const double d = (rhs - (k > 0 ? prev_d : 0.0)) / denom;
            C(i) = d;
            prev_d = d;
        }
        // back substitution needs cp[k] for k = m-2 .. 0.  The recurrence
        // cp[k] = 1/(4 - cp[k-1]) reaches its double-precision fixed point
        // (2 - sqrt(3)) by k = 13 and stays there, so 32 stored terms give
        // every cp[k] exactly.
        constexpr int kLocal = 32;
        double cps[kLocal];
        double c = 0.0;
        for (int k = 0; k < kLocal; ++k) {
            c = 1.0 / (4.0 - (k > 0 ? c : 0.0));
            cps[k] = c;
        }
        for (int k = m - 2; k >= 0; --k)
            C(k + 1) = C(k + 1) - cps[k < kLocal ? k : kLocal - 1] * C(k + 2);
    }
    C(-1) = 2 * C(0) - C(1);
    C(n) = 2 * C(n - 1) - C(n - 2);
}

__global__ void k_prefilter_scatter(const double* __restrict__ U, int nt, int nx, int ny,
                                    double* __restrict__ C) {
    const long n2 = ny + 2, n1 = static_cast<long>(nx + 2) * n2;
    const long total = static_cast<long>(nt + 2) * n1;
    for (long k = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; k < total;
         k += static_cast<long>(gridDim.x) * blockDim.x) {
        const int a = static_cast<int>(k / n1);
        const int b = static_cast<int>((k / n2) % (nx + 2));
        const int c = static_cast<int>(k % n2);
        double v = 0.0;
        if (a >= 1 && a <= nt && b >= 1 && b <= nx && c >= 1 && c <= ny)
            v = U[(static_cast<long>(a - 1) * nx + (b - 1)) * ny + (c - 1)];
        C[k] = v;
    }
}

// axis: 2 = y lines (a in 1..nt, b in 1..nx), 1 = x lines (a in 1..nt,
// c in 0..ny+1), 0 = t lines (b in 0..nx+1, c in 0..ny+1)
__global__ void k_prefilter_axis(double* __restrict__ C, int nt, int nx, int ny, int axis) {
    const long n2 = ny + 2, n1 = static_cast<long>(nx + 2) * n2;
    long lines, stride;
    int n;
    if (axis == 2) { lines = static_cast<long>(nt) * nx; stride = 1; n = ny; }
    else if (axis == 1) { lines = static_cast<long>(nt) * (ny + 2); stride = n2; n = nx; }
    else { lines = static_cast<long>(nx + 2) * (ny + 2); stride = n1; n = nt; }
    const long L = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (L >= lines) return;
    double* p;
    if (axis == 2) {
        const long a = 1 + L / nx, b = 1 + L % nx;
        p = C + a * n1 + b * n2;
    } else if (axis == 1) {
        const long a = 1 + L / (ny + 2), c = L % (ny + 2);
        p = C + a * n1 + c;
    } else {
        const long b = L / (ny + 2), c = L % (ny + 2);
        p = C + b * n2 + c;
    }
    prefilter_line(p, stride, n);
}

// (u, v) pairs in FP64, and FP32 elements pairing time index t with t + 1
// (slab = one time slice of coefficients; the last slice repeats itself).
__global__ void k_interleave(const double* __restrict__ u, const double* __restrict__ v, int64_t n,
                             int64_t slab, double2* __restrict__ out,
                             float4* __restrict__ out32) {
    for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double a = u[k], b = v[k];
        out[k] = make_double2(a, b);
        if (out32 == nullptr) continue;
        const int64_t k1 = k + slab < n ? k + slab : k;
        out32[k] = make_float4(static_cast<float>(a), static_cast<float>(b),
                               static_cast<float>(u[k1]), static_cast<float>(v[k1]));
    }
}

// 'xy' velocity samples (nt, ny, nx) to the 'ij' layout (nt, nx, ny), scaled
// (the moveaxis + unit conversion of plot_merra_ftle.py:37-41: m/s -> km/h
// is scale 3.6).  Writes are coalesced along y; the reads stride by nx.
// One 32x32 tile of one time slice per block, staged in LDS (a padding
// column keeps the column reads free of bank conflicts): the block reads 32
// rows of the 'xy' slice along x and writes 32 rows of the 'ij' slice along
// y, both coalesced.  blockIdx.y runs over (time, y-tile).
constexpr int kIngestTile = 32;
constexpr int kIngestRows = 8;  // threads per block: 32 x 8

__global__ __launch_bounds__(kIngestTile * kIngestRows) void k_ingest_xy(
    const double* __restrict__ in, int nt, int ny, int nx, double scale,
    double* __restrict__ out) {
    __shared__ double tile[kIngestTile][kIngestTile + 1];
    const int nty = (ny + kIngestTile - 1) / kIngestTile;
    const int64_t t = blockIdx.y / nty;
    const int by = static_cast<int>(blockIdx.y % nty) * kIngestTile;
    const int bx = static_cast<int>(blockIdx.x) * kIngestTile;
    const double* src = in + t * ny * static_cast<int64_t>(nx);
    double* dst = out + t * nx * static_cast<int64_t>(ny);
    const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
    for (int r = 0; r < kIngestTile; r += kIngestRows) {
        const int j = by + ty + r, i = bx + tx;
        if (j < ny && i < nx) tile[ty + r][tx] = src[static_cast<int64_t>(j) * nx + i];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kIngestTile; r += kIngestRows) {
        const int i = bx + ty + r, j = by + tx;
        if (i < nx && j < ny) dst[static_cast<int64_t>(i) * ny + j] = tile[tx][ty + r] * scale;
    }
}

// ---------------------------------------------------------------------------
// Flow-map kernels.

// Particle sources: a tensor grid x[i], y[j] in 'ij' order, or a point list.
// pos() returns false for a particle the source itself leaves out (its
// outputs are written as for a masked one); cell() is its mask index.
struct GridSource {
    const double* __restrict__ xs;
    const double* __restrict__ ys;
    int ny;
    int nx;
    // optional launch order of the 4 x 4 tiles (16 slots each): slot s runs
    // in tile order[s / 16] (see tile_order_edge_first)
    const int* __restrict__ order = nullptr;
    // slots of the grid launch: the tile order of tile_slots/tile_particle
    __host__ __device__ int64_t slots() const { return tile_slots(nx, ny); }
    __device__ __forceinline__ int64_t particle(int64_t slot) const {
        if (order != nullptr) slot = (static_cast<int64_t>(order[slot >> 4]) << 4) | (slot & 15);
        return tile_particle(slot, nx, ny);
    }
    __device__ __forceinline__ bool pos(int64_t p, double& px, double& py) const {
        const int64_t i = p / ny;
        px = xs[i];
        py = ys[p - i * ny];
        return true;
    }
    __device__ __forceinline__ int64_t cell(int64_t p) const { return p; }
};

struct PointSource {
    const double2* __restrict__ pts;
    __device__ __forceinline__ bool pos(int64_t p, double& px, double& py) const {
        const double2 q = pts[p];
        px = q.x;
        py = q.y;
        return true;
    }
    __device__ __forceinline__ int64_t cell(int64_t p) const { return p; }
};

// The auxiliary grid of flowmap_aux_grid_2D (integration.py:249-464):
// particle p = (i*ny + j)*naux + k starts at (x[i], y[j]) + offset k, the
// offsets being (h,0), (-h,0), (0,h), (0,-h) and, with eig_main, (0,0).
// On edge points only the centre particle runs (eig_main) or all four (not
// eig_main); without compute_edge edge points are left out.
struct AuxSource {
    const double* __restrict__ xs;
    const double* __restrict__ ys;
    int nx, ny, naux;
    double h;
    bool eig_main, compute_edge;
    __device__ __forceinline__ bool pos(int64_t p, double& px, double& py) const {
        const int64_t c = p / naux;
        const int k = static_cast<int>(p - c * naux);
        const int64_t i = c / ny;
        const int64_t j = c - i * ny;
        const bool edge = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
        if (edge && (!compute_edge || (eig_main && k != 4))) return false;
        const double ox = k == 0 ? h : (k == 1 ? -h : 0.0);
        const double oy = k == 2 ? h : (k == 3 ? -h : 0.0);
        px = xs[i] + ox;
        py = ys[j] + oy;
        return true;
    }
    __device__ __forceinline__ int64_t cell(int64_t p) const { return p / naux; }
};

// One solve per group of G lanes (G = 1 for analytic flows); the group's
// lanes run identical solver arithmetic and split the spline gathers.
template <class Flow, class R, int G, class Src>
__global__ __launch_bounds__(256, NBX_FLOWMAP_MIN_WAVES) void k_flowmap(Flow f, Src src, int64_t P, double s0, double s1,
                                                 double rtol, double atol,
                                                 const uint8_t* __restrict__ mask, long long nmax,
                                                 double2* __restrict__ out,
                                                 int32_t* __restrict__ status,
                                                 int32_t* __restrict__ nsteps) {
    f.prepare();
    const unsigned bid = xcd_block<Flow>();
    const int64_t tid = bid * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    int64_t p = tid / G;
    if (p >= P) return;  // whole groups leave together (P counts slots for grids)
    if constexpr (std::is_same<Src, GridSource>::value) {
        p = src.particle(p);
        if (p < 0) return;
    }
    const bool lead = (tid & (G - 1)) == 0;
    double px, py;
    const bool run = src.pos(p, px, py);
    if (!run || (mask != nullptr && mask[src.cell(p)])) {
        if (lead) {
            out[p] = make_double2(0.0, 0.0);
            if (status) status[p] = NBX_ST_MASKED;
            if (nsteps) nsteps[p] = 0;
        }
        return;
    }
    Dop853State<R> s;
    dop853_init(f, s, s0, s1, px, py, rtol, atol);
    while (dop853_step(f, s, rtol, atol, nmax)) {
    }
    if (lead) {
        out[p] = make_double2(s.px(), s.py());
        if (status) status[p] = s.idid;
        if (nsteps) nsteps[p] = s.nstep;
    }
}

// Spline coefficients staged through LDS, 4 lanes per particle, one
// wavefront per 64-thread block (16 particles, a 4 x 4 tile of the grid).
// Before every attempted step the wavefront bounds where its stencils will
// fall -- times p0*[x, x+h] of its particles, positions from the current
// point along h times the current derivative with half that displacement
// as margin on either side, in the spline's lookup coordinates -- copies
// that box of coefficient elements (bt x bx x by, at most kStageElems of 16
// bytes) from HBM/L2 into LDS with row-contiguous loads, and every stencil
// inside it is then read from LDS (Spline3<..., S = true>); stencils outside
// (a stage that strays, a box too large) are gathered from global memory as
// in k_flowmap.  Per step a wavefront's stencils span a few hours and a few
// grid cells: median box 8 KiB, 23 KiB at the largest step of a typical
// tile on the gridded workload, against 192 KiB of gathered coefficients.
// The integrator's arithmetic is that of k_flowmap, so results are equal.
constexpr int kStageElems = 1280;  // 20 KiB: 8 one-wave blocks per CU

// One global_load_lds_dwordx4: lane l's 16 bytes at src go to LDS at
// dst + 16 l (dst wavefront-uniform).  The builtin exists only in the device
// compilation pass.
__device__ __forceinline__ void load_lds16(const void* src, void* dst) {
#if __HIP_DEVICE_COMPILE__
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0,
                                     0);
#endif
}

__device__ __forceinline__ double wave_min_groups(double v) {
    // the four lanes of a group hold equal values: reduce over groups
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_max_groups(double v) {
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

template <class R>
__global__ __launch_bounds__(64) void k_flowmap_staged(
    Spline3<R, 4, false, true> f, GridSource src, int64_t P, double s0, double s1, double rtol,
    double atol, const uint8_t* __restrict__ mask, long long nmax, double2* __restrict__ out,
    int32_t* __restrict__ status, int32_t* __restrict__ nsteps) {
    using Pt = typename CoefPair<R>::type;
    __shared__ Pt stage[kStageElems];
    f.prepare();
    const int lane = static_cast<int>(threadIdx.x);
    const int64_t slot = static_cast<int64_t>(blockIdx.x) * 16 + (lane >> 2);
    const int64_t p = slot < P ? src.particle(slot) : -1;
    const bool lead = (lane & 3) == 0;
    bool running = p >= 0;
    if (running && mask != nullptr && mask[p]) {
        if (lead) {
            out[p] = make_double2(0.0, 0.0);
            if (status) status[p] = NBX_ST_MASKED;
            if (nsteps) nsteps[p] = 0;
        }
        running = false;
    }
    f.box.lds = stage;
    f.box.on = false;
    Dop853State<R> s;
    if (running) {
        double px, py;
        src.pos(p, px, py);
        dop853_init(f, s, s0, s1, px, py, rtol, atol);
    }
    auto finish = [&]() {
        if (lead) {
            out[p] = make_double2(s.px(), s.py());
            if (status) status[p] = s.idid;
            if (nsteps) nsteps[p] = s.nstep;
        }
    };
    const double inf = __builtin_inf();
    for (;;) {
        bool stepping = false;
        if (running) {
            stepping = dop853_step_begin(s, nmax);
            if (!stepping) {
                finish();
                running = false;
            }
        }
        if (!__any(stepping)) break;
        // ---- the step's box (wavefront-uniform)
        double tlo = inf, thi = -inf, xlo = inf, xhi = -inf, ylo = inf, yhi = -inf;
        if (stepping) {
            const double t0 = f.p0 * s.x, t1 = f.p0 * (s.x + s.h);
            tlo = fmin(t0, t1);
            thi = fmax(t0, t1);
            const double ex = s.h * static_cast<double>(s.k1x);
            const double ey = s.h * static_cast<double>(s.k1y);
            const double m = 0.5 * (fabs(ex) + fabs(ey));
            const double px = s.px(), py = s.py();
            const double a = f.lookup_x(px + fmin(0.0, ex) - m);
            const double b = f.lookup_x(px + fmax(0.0, ex) + m);
            if (a <= b) {
                xlo = a;
                xhi = b;
            } else {  // the box straddles the longitude wrap: no staging
                xlo = -inf;
                xhi = inf;
            }
            ylo = py + fmin(0.0, ey) - m;
            yhi = py + fmax(0.0, ey) + m;
        }
        tlo = wave_min_groups(tlo);
        thi = wave_max_groups(thi);
        xlo = wave_min_groups(xlo);
        xhi = wave_max_groups(xhi);
        ylo = wave_min_groups(ylo);
        yhi = wave_max_groups(yhi);
        int c0[3], c1[3];
        const double q0[3] = {tlo, xlo, ylo}, q1[3] = {thi, xhi, yhi};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            R fr;
            f.cell(d, fmin(fmax(q0[d], f.lo[d]), f.hi[d]), c0[d], fr);
            f.cell(d, fmin(fmax(q1[d], f.lo[d]), f.hi[d]), c1[d], fr);
        }
        const int bt = c1[0] - c0[0] + 4, bx = c1[1] - c0[1] + 4, by = c1[2] - c0[2] + 4;
        const int size = bt * bx * by;
        const bool fits = xlo > -inf && by <= 64 && ((size + 63) & ~63) <= kStageElems;
        f.box.it0 = c0[0];
        f.box.ix0 = c0[1];
        f.box.iy0 = c0[2];
        f.box.bt = bt;
        f.box.bx = bx;
        f.box.by = by;
        f.box.on = fits;
        if (fits) {
            // element e = (a*bx + b)*by + c of the box, 64 a pass: each
            // global_load_lds_dwordx4 writes the wavefront's 64 elements to
            // consecutive LDS slots, the loads all in flight at once.  Lanes
            // past the end (last pass) reload the last element into slots
            // below the rounded-up size, which no stencil reads.
            const float rby = __builtin_amdgcn_rcpf(static_cast<float>(by));
            const float rbx = __builtin_amdgcn_rcpf(static_cast<float>(bx));
            const int passes = (size + 63) >> 6;
            for (int k = 0; k < passes; ++k) {
                const int e = min(k * 64 + lane, size - 1);
                const int r = static_cast<int>((static_cast<float>(e) + 0.5f) * rby);
                const int c = e - r * by;
                const int a = static_cast<int>((static_cast<float>(r) + 0.5f) * rbx);
                const int b = r - a * bx;
                const Pt* src_e = f.C + (static_cast<long>(c0[0] + a) * f.st +
                                         static_cast<long>(c0[1] + b) * f.sx + (c0[2] + c));
                load_lds16(src_e, stage + k * 64);
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
        if (stepping) {
            running = dop853_step_body(f, s, rtol, atol);
            if (!running) finish();
        }
        __syncthreads();  // every read of the box is done before the next copy
    }
}

template <class Flow, class R>
__global__ void k_flow_eval(Flow f, const double* __restrict__ t, const double2* __restrict__ pts,
                            int64_t n, double2* __restrict__ out) {
    f.prepare();
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n) return;
    R a, b;
    f(t[k], pts[k].x, pts[k].y, a, b);
    out[k] = make_double2(static_cast<double>(a), static_cast<double>(b));
}

// defaults: the best shapes measured on the gridded-velocity and
// double-gyre workloads (spline flows: 4 lanes per particle, 128-thread
// blocks); the autotuner overrides them per workload
LaunchCfg g_cfg[4][2] = {{{1, 256, 0}, {1, 256, 0}},
                         {{1, 256, 0}, {1, 256, 0}},
                         {{4, 128, 0}, {4, 128, 0}},
                         {{1, 128, 0}, {1, 128, 0}}};
// edge-first tile order of the spline grid launch, per precision
int g_edge_first[2] = {1, 1};
std::mutex g_cfg_mutex;

LaunchCfg launch_cfg(int kind, int precision) {
    std::lock_guard<std::mutex> lock(g_cfg_mutex);
    return g_cfg[kind][precision];
}

struct SolveArgs {
    int64_t P;
    double s0, s1, rtol, atol;
    const uint8_t* mask;
    long long nmax;
    double2* out;
    int32_t* status;
    int32_t* nsteps;
};

int device_cu_count() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    if (static_cast<int>(cache.size()) <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        cache[dev] = n;
    }
    return cache[dev];
}

// The single-grid request as a queue entry (grid sources only).
inline ReqDev single_request(const GridSource& src, const SolveArgs& a) {
    return ReqDev{a.s0, a.s1, src.xs, src.ys, static_cast<int>(a.P / src.ny), src.ny, a.mask,
                  a.out, a.status, a.nsteps};
}

template <class Flow, class R, int G, class Src>
int launch_solve(const Flow& f, const Src& src, const SolveArgs& a, const LaunchCfg& cfg,
                 hipStream_t s) {
    const int block = cfg.block;
    if constexpr (std::is_same<Src, GridSource>::value) {
        if (cfg.persistent > 0) {  // persistent grid over the queue of particle tiles
            const int cus = device_cu_count();
            NBX_REQUIRE(cus > 0, "cannot query the compute-unit count");
            const int64_t slots = src.slots();
            if (slots == 0) return 0;
            const int64_t ntiles = slots / (64 / G);
            NBX_REQUIRE(ntiles < (int64_t(1) << 31), "too many particle tiles for the queue");
            const unsigned long long ring = chunk_ring(ntiles);
            const size_t state_off = ring * sizeof(unsigned long long) + 256;
            Workspace ws(s, 3, state_off + sizeof(Dop853State<R>) * static_cast<size_t>(slots));
            if (ws.error()) return ws.error();
            char* w = ws.get();
            const ChunkQueue q{reinterpret_cast<unsigned long long*>(w),
                               reinterpret_cast<unsigned long long*>(w + state_off - 256),
                               reinterpret_cast<unsigned long long*>(w + state_off - 128),
                               reinterpret_cast<unsigned long long*>(w + state_off - 64),
                               ring - 1, static_cast<unsigned long long>(ntiles)};
            k_chunk_init<<<blocks_for(static_cast<int64_t>(ring), 256), 256, 0, s>>>(q, ntiles);
            NBX_LAUNCHED("chunk_init");
            const int64_t want = (ntiles * 64 + block - 1) / block;
            const int64_t grid = std::min<int64_t>(want, static_cast<int64_t>(cus) * cfg.persistent);
            k_flowmap_chunks<Flow, R, G><<<static_cast<unsigned>(grid), block, 0, s>>>(
                f, single_request(src, a), nullptr, nullptr, 1, q,
                reinterpret_cast<Dop853State<R>*>(w + state_off), a.rtol, a.atol, a.nmax);
            NBX_LAUNCHED("flowmap_chunks");
            return 0;
        }
    }
    int64_t slots = a.P;
    if constexpr (std::is_same<Src, GridSource>::value) slots = src.slots();
    k_flowmap<Flow, R, G, Src><<<blocks_for(slots * G, block), block, 0, s>>>(
        f, src, slots, a.s0, a.s1, a.rtol, a.atol, a.mask, a.nmax, a.out, a.status, a.nsteps);
    NBX_LAUNCHED("flowmap");
    return 0;
}

// Edge-first launch order of a grid's 4 x 4 particle tiles.  A particle
// that leaves the velocity box crosses the step from the spline's values to
// zero (extrapolation 0), where DOP853 cuts its steps and rejects some, so
// tiles near the box's edge take the most steps (gridded workload: 29.5
// steps per particle within 2 degrees of the edge, 21 beyond 10); launched
// in plain tile order some of them start last and form the launch's tail.
// The tiles are counted into 64 buckets by the distance of their centre
// particle to the box edge in velocity-grid cells (the last bucket also
// takes 63 cells and more, and the tiles past the particle grid's edge) and
// scattered bucket by bucket: order[k] is the tile launched k-th.  Within a
// bucket the order is whatever the atomics give; the launch order never
// changes a particle's result.
// Edge-first order: a counting sort of the tiles by bucket (distance of the
// centre of the tile's block of kOrderBlock x kOrderBlock particles to the
// nearest edge of the velocity box, in grid cells, clipped to 0..63).  Each block histograms its tiles in LDS and stores its
// 64 counts; the scatter derives every block's offsets from those counts
// (no global atomics: nearly every interior tile falls in the last bucket,
// and a shared global counter serialised them).
constexpr int kOrderThreads = 256;
#ifndef NBX_ORDER_BLOCK
#define NBX_ORDER_BLOCK 32
#endif
constexpr int64_t kOrderBlock = NBX_ORDER_BLOCK;  // particles per side, a multiple of 8

template <class Flow>
__global__ __launch_bounds__(kOrderThreads) void k_tile_order_keys(
    Flow f, GridSource src, int64_t tiles, int* __restrict__ bucket,
    unsigned* __restrict__ block_count) {
    __shared__ unsigned hist[64];
    if (threadIdx.x < 64) hist[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t u = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (u < tiles) {
        const int64_t T = u >> 2, q = u & 3, ntj = (src.ny + 7) / 8;
        const int64_t i0 = (T / ntj) * 8 + (q >> 1) * 4, j0 = (T % ntj) * 8 + (q & 1) * 4;
        int b = 63;
        if (i0 < src.nx && j0 < src.ny) {
            // the centre of the tile's block of kOrderBlock x kOrderBlock
            // particles: whole blocks share a bucket, so each bucket is swept
            // band by band in tile order (compact working sets in L2)
            const int64_t ib = (i0 / kOrderBlock) * kOrderBlock + kOrderBlock / 2;
            const int64_t jb = (j0 / kOrderBlock) * kOrderBlock + kOrderBlock / 2;
            const int64_t i = ib < src.nx ? ib : src.nx - 1;
            const int64_t j = jb < src.ny ? jb : src.ny - 1;
            const double x = f.lookup_x(src.xs[i]), y = src.ys[j];
            const double d = fmin(fmin((x - f.lo[1]) * f.dinv[1], (f.hi[1] - x) * f.dinv[1]),
                                  fmin((y - f.lo[2]) * f.dinv[2], (f.hi[2] - y) * f.dinv[2]));
            b = d > 0 ? (d < 63 ? static_cast<int>(d) : 63) : 0;
        }
        bucket[u] = b;
        atomicAdd(&hist[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) block_count[blockIdx.x * 64 + threadIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kOrderThreads) void k_tile_order_scatter(
    const int* __restrict__ bucket, const unsigned* __restrict__ block_count, int nblocks,
    int64_t tiles, int* __restrict__ order) {
    // per bucket: the tiles of all blocks (total) and of the blocks before
    // this one (before), four partial sums per bucket over the block counts
    __shared__ unsigned total[4][64], before[4][64];
    __shared__ unsigned base[64], cursor[64];
    const int bk = threadIdx.x & 63, part = threadIdx.x >> 6;
    unsigned tot = 0u, bef = 0u;
    for (int k = part; k < nblocks; k += 4) {
        const unsigned c = block_count[k * 64 + bk];
        tot += c;
        bef += k < static_cast<int>(blockIdx.x) ? c : 0u;
    }
    total[part][bk] = tot;
    before[part][bk] = bef;
    __syncthreads();
    if (threadIdx.x < 64) {
        base[bk] = total[0][bk] + total[1][bk] + total[2][bk] + total[3][bk];
        cursor[bk] = before[0][bk] + before[1][bk] + before[2][bk] + before[3][bk];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned acc = 0u;
        for (int b = 0; b < 64; ++b) {
            const unsigned n = base[b];
            base[b] = acc;
            acc += n;
        }
    }
    __syncthreads();
    const int64_t u = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (u >= tiles) return;
    const int b = bucket[u];
    order[base[b] + atomicAdd(&cursor[b], 1u)] = static_cast<int>(u);
}

inline int64_t tile_order_blocks(int64_t tiles) {
    return (tiles + kOrderThreads - 1) / kOrderThreads;
}

// Enqueue the edge-first order of src's tiles into the stream's workspace
// and return src with it (src unchanged for an empty grid).
template <class Flow>
int tile_order_edge_first(const Flow& f, const GridSource& src, hipStream_t s, Workspace& ws,
                          GridSource& out) {
    out = src;
    const int64_t tiles = src.slots() / 16;
    if (tiles == 0) return 0;
    NBX_REQUIRE(tiles < (int64_t(1) << 31), "too many particle tiles");
    if (ws.error()) return ws.error();
    const int64_t nb = tile_order_blocks(tiles);
    char* w = ws.get();
    unsigned* block_count = reinterpret_cast<unsigned*>(w);  // 64 per block
    int* bucket = reinterpret_cast<int*>(w + 256 * nb);
    int* order = bucket + tiles;
    k_tile_order_keys<<<static_cast<unsigned>(nb), kOrderThreads, 0, s>>>(f, src, tiles, bucket,
                                                                          block_count);
    NBX_LAUNCHED("tile_order_keys");
    k_tile_order_scatter<<<static_cast<unsigned>(nb), kOrderThreads, 0, s>>>(
        bucket, block_count, static_cast<int>(nb), tiles, order);
    NBX_LAUNCHED("tile_order_scatter");
    out.order = order;
    return 0;
}

inline size_t tile_order_bytes(const GridSource& src) {
    const int64_t tiles = src.slots() / 16;
    return 256 * static_cast<size_t>(tile_order_blocks(tiles)) +
           2 * sizeof(int) * static_cast<size_t>(tiles);
}

template <class R, class Src>
int dispatch_solve(const nbx_flow* fl, const double* params, int np, const Src& src,
                   const SolveArgs& a, int precision, hipStream_t s) {
    const LaunchCfg cfg = launch_cfg(fl->kind, precision);
    if (fl->kind == NBX_FLOW_SPLINE3) {
        // linear extrapolation: the kernel with the gradient path, one lane
        // per particle (the lane-group setting is not used for such flows)
        if (fl->extrap == 1)
            return launch_solve<Spline3<R, 1, true>, R, 1>(make_spline<R, true>(fl, params[0]),
                                                           src, a, cfg, s);
        const Spline3<R, 1> sp = make_spline<R, false>(fl, params[0]);
        if constexpr (std::is_same<Src, GridSource>::value) {
            if (cfg.persistent < 0) {  // LDS-staged coefficients, 4 lanes, one-wave blocks
                const int64_t slots = src.slots();
                if (slots == 0) return 0;
                k_flowmap_staged<R><<<blocks_for(slots * 4, 64), 64, 0, s>>>(
                    regroup<4, R, true>(sp), src, slots, a.s0, a.s1, a.rtol, a.atol, a.mask,
                    a.nmax, a.out, a.status, a.nsteps);
                NBX_LAUNCHED("flowmap_staged");
                return 0;
            }
        }
        Src run_src = src;
        std::unique_ptr<Workspace> order_ws;  // held until the launch is enqueued
        if constexpr (std::is_same<Src, GridSource>::value) {
            bool edge_first;
            {
                std::lock_guard<std::mutex> lock(g_cfg_mutex);
                edge_first = g_edge_first[precision == NBX_FP32_DISP ? 1 : 0] != 0;
            }
            if (edge_first && cfg.persistent == 0) {
                order_ws = std::make_unique<Workspace>(s, 4, tile_order_bytes(src));
                if (int rc = tile_order_edge_first(sp, src, s, *order_ws, run_src)) return rc;
            }
        }
        if (cfg.lanes == 4) {
            NBX_REQUIRE(sp.cbytes != 0, "the 4-lane spline kernel needs a coefficient array "
                                        "under 4 GiB (set 1 or 2 lanes per particle)");
            return launch_solve<Spline3<R, 4>, R, 4>(regroup<4>(sp), run_src, a, cfg, s);
        }
        if (cfg.lanes == 2)
            return launch_solve<Spline3<R, 2>, R, 2>(regroup<2>(sp), run_src, a, cfg, s);
        return launch_solve<Spline3<R, 1>, R, 1>(sp, run_src, a, cfg, s);
    }
    if (fl->kind == NBX_FLOW_LINEAR3)
        return launch_solve<Linear3<R>, R, 1>(make_linear<R>(fl, params[0]), src, a, cfg, s);
    if (fl->kind == NBX_FLOW_DOUBLE_GYRE) {
        NBX_REQUIRE(np >= 6, "double_gyre needs 6 parameters [dir, A, eps, alpha, omega, psi]");
        const DoubleGyre<R> dg{params[0], params[1], params[2], params[3], params[4], params[5]};
        return launch_solve<DoubleGyre<R>, R, 1>(dg, src, a, cfg, s);
    }
    if (fl->kind == NBX_FLOW_BICKLEY_JET) {
        NBX_REQUIRE(np >= 12, "bickley_jet needs 12 parameters");
        BickleyJet<R> bj;
        for (int k = 0; k < 12; ++k) bj.p[k] = params[k];
        return launch_solve<BickleyJet<R>, R, 1>(bj, src, a, cfg, s);
    }
    set_error("unknown flow kind");
    return -2;
}

template <class Src>
int solve(const nbx_flow* fl, const double* params, int np, const Src& src, const SolveArgs& a,
          int precision, hipStream_t s) {
    if (precision == NBX_FP32_DISP) return dispatch_solve<float>(fl, params, np, src, a, precision, s);
    return dispatch_solve<double>(fl, params, np, src, a, precision, s);
}

}  // namespace nbx

using namespace nbx;

extern "C" {

const char* nbx_last_error(void) { return nbx::g_last_error.c_str(); }

int nbx_version(void) { return 1; }

int nbx_flow_create_analytic(int kind, nbx_flow** out) {
    NBX_REQUIRE(out != nullptr, "out is NULL");
    NBX_REQUIRE(kind == NBX_FLOW_DOUBLE_GYRE || kind == NBX_FLOW_BICKLEY_JET,
                "unknown analytic flow kind");
    nbx_flow* fl = new nbx_flow();
    fl->kind = kind;
    *out = fl;
    return 0;
}

int nbx_flow_create_spline3(const double* grid9_host, const double* Cu, const double* Cv,
                            int spherical, int extrap, double r, nbx_flow** out,
                            nbx_stream stream) {
    NBX_REQUIRE(out != nullptr && grid9_host != nullptr && Cu != nullptr && Cv != nullptr,
                "NULL argument");
    NBX_REQUIRE(spherical >= 0 && spherical <= 2, "spherical must be 0, 1 or 2");
    NBX_REQUIRE(extrap >= 0 && extrap <= 2, "unknown extrapolation mode");
    int n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = static_cast<int>(grid9_host[3 * d + 2]);
        NBX_REQUIRE(n[d] >= 2 && n[d] == grid9_host[3 * d + 2],
                    "spline grids need at least 2 points per axis");
        NBX_REQUIRE(grid9_host[3 * d + 1] > grid9_host[3 * d], "grid axes must be ascending");
    }
    const int64_t ncoef = static_cast<int64_t>(n[0] + 2) * (n[1] + 2) * (n[2] + 2);
    nbx_flow* fl = new nbx_flow();
    fl->kind = NBX_FLOW_SPLINE3;
    std::memcpy(fl->grid9, grid9_host, sizeof(fl->grid9));
    fl->spherical = spherical;
    fl->extrap = extrap;
    fl->r = r;
    fl->ncoef = ncoef;
    if (hipMalloc(&fl->coef, sizeof(double2) * ncoef) != hipSuccess ||
        hipMalloc(&fl->coef32, sizeof(float4) * ncoef) != hipSuccess) {
        nbx_flow_destroy(fl);
        set_error("hipMalloc of spline coefficients failed");
        return -1;
    }
    hipStream_t s = as_stream(stream);
    k_interleave<<<blocks_for(ncoef, 256) < 4096 ? blocks_for(ncoef, 256) : 4096, 256, 0, s>>>(
        Cu, Cv, ncoef, static_cast<int64_t>(n[1] + 2) * (n[2] + 2), fl->coef, fl->coef32);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        nbx_flow_destroy(fl);
        set_error(std::string("interleave launch: ") + hipGetErrorString(e));
        return -1;
    }
    *out = fl;
    return 0;
}

int nbx_flow_create_linear3(const double* grid9_host, const double* U, const double* V,
                            int spherical, int extrap, double r, nbx_flow** out,
                            nbx_stream stream) {
    NBX_REQUIRE(out != nullptr && grid9_host != nullptr && U != nullptr && V != nullptr,
                "NULL argument");
    NBX_REQUIRE(spherical >= 0 && spherical <= 2, "spherical must be 0, 1 or 2");
    NBX_REQUIRE(extrap >= 0 && extrap <= 2, "unknown extrapolation mode");
    int n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = static_cast<int>(grid9_host[3 * d + 2]);
        NBX_REQUIRE(n[d] >= 2 && n[d] == grid9_host[3 * d + 2],
                    "trilinear grids need at least 2 points per axis");
        NBX_REQUIRE(grid9_host[3 * d + 1] > grid9_host[3 * d], "grid axes must be ascending");
    }
    const int64_t nval = static_cast<int64_t>(n[0]) * n[1] * n[2];
    nbx_flow* fl = new nbx_flow();
    fl->kind = NBX_FLOW_LINEAR3;
    std::memcpy(fl->grid9, grid9_host, sizeof(fl->grid9));
    fl->spherical = spherical;
    fl->extrap = extrap;
    fl->r = r;
    fl->ncoef = nval;
    if (hipMalloc(&fl->coef, sizeof(double2) * nval) != hipSuccess) {
        nbx_flow_destroy(fl);
        set_error("hipMalloc of velocity samples failed");
        return -1;
    }
    hipStream_t s = as_stream(stream);
    k_interleave<<<blocks_for(nval, 256) < 4096 ? blocks_for(nval, 256) : 4096, 256, 0, s>>>(
        U, V, nval, 0, fl->coef, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        nbx_flow_destroy(fl);
        set_error(std::string("interleave launch: ") + hipGetErrorString(e));
        return -1;
    }
    *out = fl;
    return 0;
}

int nbx_ingest_xy(const double* U_xy, int nt, int ny, int nx, double scale, double* U_ij,
                  nbx_stream stream) {
    NBX_REQUIRE(nt >= 0 && ny >= 0 && nx >= 0, "negative size");
    const int64_t n = static_cast<int64_t>(nt) * ny * nx;
    if (n == 0) return 0;
    NBX_REQUIRE(U_xy != nullptr && U_ij != nullptr && U_xy != U_ij, "NULL or aliased array");
    const unsigned tiles_x = static_cast<unsigned>((nx + kIngestTile - 1) / kIngestTile);
    const int64_t tiles_ty = static_cast<int64_t>(nt) * ((ny + kIngestTile - 1) / kIngestTile);
    // the dispatch's y extent in work-items (tiles x 8 rows) must fit 32 bits
    NBX_REQUIRE(tiles_ty * kIngestRows <= 4294967295LL, "too many time slices for one launch");
    k_ingest_xy<<<dim3(tiles_x, static_cast<unsigned>(tiles_ty)), dim3(kIngestTile, kIngestRows),
                  0, as_stream(stream)>>>(U_xy, nt, ny, nx, scale, U_ij);
    NBX_LAUNCHED("ingest_xy");
    return 0;
}

void nbx_flow_destroy(nbx_flow* fl) {
    if (!fl) return;
    if (fl->coef) (void)hipFree(fl->coef);
    if (fl->coef32) (void)hipFree(fl->coef32);
    delete fl;
}

int nbx_prefilter3(const double* U, int nt, int nx, int ny, double* C, nbx_stream stream) {
    NBX_REQUIRE(U != nullptr && C != nullptr, "NULL argument");
    NBX_REQUIRE(nt >= 1 && nx >= 1 && ny >= 1, "empty grid");
    hipStream_t s = as_stream(stream);
    const long total = static_cast<long>(nt + 2) * (nx + 2) * (ny + 2);
    const unsigned nb = blocks_for(total, 256) < 8192 ? blocks_for(total, 256) : 8192;
    k_prefilter_scatter<<<nb, 256, 0, s>>>(U, nt, nx, ny, C);
    NBX_LAUNCHED("prefilter_scatter");
    const long lines2 = static_cast<long>(nt) * nx;
    k_prefilter_axis<<<blocks_for(lines2, 128), 128, 0, s>>>(C, nt, nx, ny, 2);
    NBX_LAUNCHED("prefilter_axis_y");
    const long lines1 = static_cast<long>(nt) * (ny + 2);
    k_prefilter_axis<<<blocks_for(lines1, 128), 128, 0, s>>>(C, nt, nx, ny, 1);
    NBX_LAUNCHED("prefilter_axis_x");
    const long lines0 = static_cast<long>(nx + 2) * (ny + 2);
    k_prefilter_axis<<<blocks_for(lines0, 128), 128, 0, s>>>(C, nt, nx, ny, 0);
    NBX_LAUNCHED("prefilter_axis_t");
    return 0;
}

int nbx_flow_eval(const nbx_flow* fl, const double* params, int np, const double* t,
                  const double* pts, int64_t n, double* out, nbx_stream stream) {
    NBX_REQUIRE(fl != nullptr && params != nullptr && np >= 1, "flow and params are required");
    if (n == 0) return 0;
    NBX_REQUIRE(t != nullptr && pts != nullptr && out != nullptr, "NULL array");
    hipStream_t s = as_stream(stream);
    const auto* P = reinterpret_cast<const double2*>(pts);
    auto* O = reinterpret_cast<double2*>(out);
    const unsigned nb = blocks_for(n, 256);
    if (fl->kind == NBX_FLOW_SPLINE3) {
        k_flow_eval<Spline3<double, 1, true>, double><<<nb, 256, 0, s>>>(
            make_spline<double, true>(fl, params[0]), t, P, n, O);
    } else if (fl->kind == NBX_FLOW_LINEAR3) {
        k_flow_eval<Linear3<double>, double><<<nb, 256, 0, s>>>(
            make_linear<double>(fl, params[0]), t, P, n, O);
    } else if (fl->kind == NBX_FLOW_DOUBLE_GYRE) {
        NBX_REQUIRE(np >= 6, "double_gyre needs 6 parameters");
        const DoubleGyre<double> dg{params[0], params[1], params[2], params[3], params[4], params[5]};
        k_flow_eval<DoubleGyre<double>, double><<<nb, 256, 0, s>>>(dg, t, P, n, O);
    } else {
        NBX_REQUIRE(np >= 12, "bickley_jet needs 12 parameters");
        BickleyJet<double> bj;
        for (int k = 0; k < 12; ++k) bj.p[k] = params[k];
        k_flow_eval<BickleyJet<double>, double><<<nb, 256, 0, s>>>(bj, t, P, n, O);
    }
    NBX_LAUNCHED("flow_eval");
    return 0;
}

static int check_solve_args(const nbx_flow* fl, const double* params, int np, double rtol,
                            double atol, int64_t max_steps, int precision) {
    NBX_REQUIRE(fl != nullptr, "flow handle is NULL");
    NBX_REQUIRE(params != nullptr && np >= 1, "params must hold at least the direction");
    NBX_REQUIRE(rtol > 0 && atol >= 0, "tolerances must be positive");
    NBX_REQUIRE(max_steps >= 1, "max_steps must be >= 1");
    NBX_REQUIRE(precision == NBX_FP64 || precision == NBX_FP32_DISP, "unknown precision");
    return 0;
}

int nbx_flowmap_grid_2d(const nbx_flow* fl, double t0, double T, const double* x, int nx,
                        const double* y, int ny, const double* params, int np, double rtol,
                        double atol, const uint8_t* mask, int64_t max_steps, int precision,
                        double* out, int32_t* status, int32_t* nsteps, nbx_stream stream) {
    if (int rc = check_solve_args(fl, params, np, rtol, atol, max_steps, precision)) return rc;
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    const int64_t P = static_cast<int64_t>(nx) * ny;
    if (P == 0) return 0;
    NBX_REQUIRE(x != nullptr && y != nullptr && out != nullptr, "NULL array");
    SolveArgs a{P, params[0] * t0, params[0] * (t0 + T), rtol, atol, mask, max_steps,
                reinterpret_cast<double2*>(out), status, nsteps};
    return solve(fl, params, np, GridSource{x, y, ny, nx}, a, precision, as_stream(stream));
}

int nbx_flowmap_aux_grid_2d(const nbx_flow* fl, double t0, double T, const double* x, int nx,
                            const double* y, int ny, double h, int eig_main, int compute_edge,
                            const double* params, int np, double rtol, double atol,
                            const uint8_t* mask, int64_t max_steps, int precision, double* out,
                            int32_t* status, int32_t* nsteps, nbx_stream stream) {
    if (int rc = check_solve_args(fl, params, np, rtol, atol, max_steps, precision)) return rc;
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    const int naux = eig_main ? 5 : 4;
    const int64_t P = static_cast<int64_t>(nx) * ny * naux;
    if (P == 0) return 0;
    NBX_REQUIRE(x != nullptr && y != nullptr && out != nullptr, "NULL array");
    SolveArgs a{P, params[0] * t0, params[0] * (t0 + T), rtol, atol, mask, max_steps,
                reinterpret_cast<double2*>(out), status, nsteps};
    const AuxSource src{x, y, nx, ny, naux, h, eig_main != 0, compute_edge != 0};
    return solve(fl, params, np, src, a, precision, as_stream(stream));
}

int nbx_flowmap_pts(const nbx_flow* fl, double t0, double T, const double* pts, int64_t npts,
                    const double* params, int np, double rtol, double atol, const uint8_t* mask,
                    int64_t max_steps, int precision, double* out, int32_t* status,
                    int32_t* nsteps, nbx_stream stream) {
    if (int rc = check_solve_args(fl, params, np, rtol, atol, max_steps, precision)) return rc;
    NBX_REQUIRE(npts >= 0, "negative point count");
    if (npts == 0) return 0;
    NBX_REQUIRE(pts != nullptr && out != nullptr, "NULL array");
    SolveArgs a{npts, params[0] * t0, params[0] * (t0 + T), rtol, atol, mask, max_steps,
                reinterpret_cast<double2*>(out), status, nsteps};
    return solve(fl, params, np, PointSource{reinterpret_cast<const double2*>(pts)}, a, precision,
                 as_stream(stream));
}

int nbx_set_launch_config(int kind, int precision, int lanes_per_particle, int block_threads,
                          int persistent_blocks_per_cu) {
    NBX_REQUIRE(kind >= 0 && kind <= 3, "unknown flow kind");
    NBX_REQUIRE(precision == NBX_FP64 || precision == NBX_FP32_DISP, "unknown precision");
    NBX_REQUIRE(lanes_per_particle == 1 || lanes_per_particle == 2 || lanes_per_particle == 4,
                "lanes_per_particle must be 1, 2 or 4");
    NBX_REQUIRE(block_threads == 64 || block_threads == 128 || block_threads == 256,
                "block_threads must be 64, 128 or 256");
    NBX_REQUIRE(kind == NBX_FLOW_SPLINE3 || lanes_per_particle == 1,
                "analytic flows run one lane per particle");
    NBX_REQUIRE(persistent_blocks_per_cu >= -1 && persistent_blocks_per_cu <= 64,
                "persistent_blocks_per_cu must be in [-1, 64]");
    NBX_REQUIRE(persistent_blocks_per_cu != -1 ||
                    (kind == NBX_FLOW_SPLINE3 && lanes_per_particle == 4 && block_threads == 64),
                "the LDS-staged launch (-1) is for spline flows at 4 lanes, 64-thread blocks");
    std::lock_guard<std::mutex> lock(g_cfg_mutex);
    g_cfg[kind][precision] = LaunchCfg{lanes_per_particle, block_threads, persistent_blocks_per_cu};
    return 0;
}

int nbx_set_tile_order(int precision, int edge_first) {
    NBX_REQUIRE(precision == NBX_FP64 || precision == NBX_FP32_DISP, "unknown precision");
    NBX_REQUIRE(edge_first == 0 || edge_first == 1, "edge_first must be 0 or 1");
    std::lock_guard<std::mutex> lock(g_cfg_mutex);
    g_edge_first[precision == NBX_FP32_DISP ? 1 : 0] = edge_first;
    return 0;
}

int nbx_get_tile_order(int precision, int* edge_first) {
    NBX_REQUIRE(precision == NBX_FP64 || precision == NBX_FP32_DISP, "unknown precision");
    NBX_REQUIRE(edge_first != nullptr, "NULL output");
    std::lock_guard<std::mutex> lock(g_cfg_mutex);
    *edge_first = g_edge_first[precision == NBX_FP32_DISP ? 1 : 0];
    return 0;
}

int nbx_get_launch_config(int kind, int precision, int* lanes_per_particle, int* block_threads,
                          int* persistent_blocks_per_cu) {
    NBX_REQUIRE(kind >= 0 && kind <= 3, "unknown flow kind");
    NBX_REQUIRE(precision == NBX_FP64 || precision == NBX_FP32_DISP, "unknown precision");
    NBX_REQUIRE(lanes_per_particle != nullptr && block_threads != nullptr &&
                    persistent_blocks_per_cu != nullptr,
                "NULL output");
    const LaunchCfg c = launch_cfg(kind, precision);
    *lanes_per_particle = c.lanes;
    *block_threads = c.block;
    *persistent_blocks_per_cu = c.persistent;
    return 0;
}

}  // extern "C"
