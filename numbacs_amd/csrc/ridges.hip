// FTLE ridge extraction: percentile threshold, ridge predicate, stream
// compaction, 8-connected labelling and grouping into ridges.
//
// Reference: ftle_ridge_pts (extraction/ridges.py:9-76), _ftle_ridges
// (ridges.py:79-154) and ftle_ridges (ridges.py:186-229), which use
// np.percentile (linear method) and scipy.ndimage.label with a full 3x3
// structure.
//
// Device pipeline (with a percentile threshold no host round trip before the
// end, the grouping stage being sized by a bound on the ridge-point count;
// without one the count is read back to size it):
//   percentile  : two radix selects (8 passes of 8 bits over order-preserving
//                 64-bit keys, each pass's blocks first picking the previous
//                 pass's digit) for the order statistics around the virtual
//                 index, then numpy's two-sided lerp -- f_min stays in HBM;
//   predicate   : one thread per grid point, the reference's central
//                 differences and Steger sub-pixel test;
//   compaction  : exclusive scan over the ridge flags (tile scans, then the
//                 tile offsets, which consumers of the grouping path add
//                 themselves), so points come out in raveled (i*ny+j) order;
//   labelling   : lock-free union-find (atomicMin linking of the larger root
//                 under the smaller) over the 4 raster-earlier neighbours,
//                 then path flattening; every root is its component's
//                 smallest raveled index, so ordering components by root is
//                 scipy.ndimage.label's raster order of first pixels;
//   grouping    : bitonic sort of (root << 32 | index) keys, segment
//                 boundaries, min_ridge_pts filter and a scatter.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "common.h"
#include "ordered_ridges.h"

// Scratch of nbx_ridges_noalloc: a fixed 64 MiB device array in place of the
// allocated workspace, for callers that cannot allocate.  The host never
// uses the array's name as an address: each device's copy lives at its own
// address, which hipGetSymbolAddress returns (once per device, with that
// device current, in scratch_address below) and which is then passed to the
// kernels.
static __device__ __align__(256) char scratch[static_cast<size_t>(64) << 20];

namespace nbx {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;  // 1024 flags per block

constexpr int kSortThreads = 512;
constexpr int kSortChunk = 2 * kSortThreads;  // keys per LDS chunk

// Radix-select state.  The digit picked from pass p's histograms is
// computed by every block of pass p+1 (and by the finishing kernel after the
// last pass) instead of by a launch of its own, so the histograms rotate
// over three buffers (pass p adds into hist[p % 3], reads hist[(p-1) % 3]
// and clears hist[(p+1) % 3], which no block of p reads or writes) and the
// digit state over two (pass p's blocks read the state of pass p-1's
// search, block 0 stores the updated one for pass p+1).
struct SelState {
    unsigned long long prefix[2][2];  // [pass & 1][search]
    unsigned long long mask[2][2];
    long long k[2][2];
    unsigned int hist[3][2][256];     // [pass % 3][search][bin]
    unsigned int nan_count;
    double result;
};

__device__ __forceinline__ unsigned long long order_key(double v) {
    const unsigned long long b = __double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double(static_cast<long long>(b));
}

__global__ void k_sel_init(SelState* st, long long k_lo, long long k_hi) {
    const int t = threadIdx.x;
    if (t < 256) {
        st->hist[0][0][t] = 0;
        st->hist[0][1][t] = 0;
    }
    if (t == 0) {
        st->prefix[0][0] = st->prefix[0][1] = 0;
        st->mask[0][0] = st->mask[0][1] = 0;
        st->k[0][0] = k_lo;
        st->k[0][1] = k_hi;
        st->nan_count = 0;
        st->result = 0.0;
    }
}

// Digit state of both searches after pass `prev` (digit at `shift`), into
// LDS: an inclusive scan of the pass's 256-bin histograms, then the bin
// whose cumulative count first exceeds k.  Called by all 256 threads; ends
// with a barrier.
__device__ void sel_advance(const SelState* st, int prev, int shift, long long (*cum)[256],
                            unsigned long long* sp, unsigned long long* sm, long long* sk) {
    const int b = threadIdx.x;
    const int sb = prev & 1, hb = prev % 3;
    cum[0][b] = st->hist[hb][0][b];
    cum[1][b] = st->hist[hb][1][b];
    if (b < 2) {
        sp[b] = st->prefix[sb][b];
        sm[b] = st->mask[sb][b];
        sk[b] = st->k[sb][b];
    }
    __syncthreads();
    // read both ranks before any thread updates them (the scan's barriers
    // order these reads before the single writer below)
    const long long kk[2] = {sk[0], sk[1]};
    for (int off = 1; off < 256; off <<= 1) {
        const long long a0 = b >= off ? cum[0][b - off] : 0;
        const long long a1 = b >= off ? cum[1][b - off] : 0;
        __syncthreads();
        cum[0][b] += a0;
        cum[1][b] += a1;
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const long long k = kk[s];
        const long long before = b > 0 ? cum[s][b - 1] : 0;
        // exactly one bin has before <= k < cum (k < total always holds)
        if (before <= k && k < cum[s][b]) {
            sk[s] = k - before;
            sp[s] |= static_cast<unsigned long long>(b) << shift;
            sm[s] |= 0xffull << shift;
        }
    }
    __syncthreads();
}

// Pass `pass` (digit at bits 56 - 8*pass) of both searches.
__global__ __launch_bounds__(256) void k_sel_hist(const double* __restrict__ f, int64_t n,
                                                  SelState* st, int pass) {
    __shared__ unsigned int h[2][256];
    __shared__ unsigned int nans;
    __shared__ long long cum[2][256];
    __shared__ unsigned long long sp[2], sm[2];
    __shared__ long long sk[2];
    const int t = threadIdx.x;
    const int shift = 56 - 8 * pass;
    h[0][t] = 0;
    h[1][t] = 0;
    if (t == 0) nans = 0;
    if (pass == 0) {
        if (t < 2) {
            sp[t] = st->prefix[0][t];
            sm[t] = st->mask[0][t];
            sk[t] = st->k[0][t];
        }
        __syncthreads();
    } else {
        sel_advance(st, pass - 1, shift + 8, cum, sp, sm, sk);
    }
    if (blockIdx.x == 0) {
        if (pass > 0 && t < 2) {
            st->prefix[pass & 1][t] = sp[t];
            st->mask[pass & 1][t] = sm[t];
            st->k[pass & 1][t] = sk[t];
        }
        st->hist[(pass + 1) % 3][0][t] = 0;
        st->hist[(pass + 1) % 3][1][t] = 0;
    }
    const unsigned long long p0 = sp[0], m0 = sm[0];
    const unsigned long long p1 = sp[1], m1 = sm[1];
    for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + t; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double v = f[k];
        if (pass == 0 && v != v) atomicAdd(&nans, 1u);
        const unsigned long long key = order_key(v);
        const unsigned int bin = static_cast<unsigned int>((key >> shift) & 0xffu);
        if ((key & m0) == p0) atomicAdd(&h[0][bin], 1u);
        if ((key & m1) == p1) atomicAdd(&h[1][bin], 1u);
    }
    __syncthreads();
    unsigned int* gh = st->hist[pass % 3][0];
    if (h[0][t]) atomicAdd(&gh[t], h[0][t]);
    if (h[1][t]) atomicAdd(&gh[256 + t], h[1][t]);
    if (t == 0 && nans) atomicAdd(&st->nan_count, nans);
}

// The last digit (pass 7, bits 0-7), then numpy's _lerp: a + (b-a)*g, or
// b - (b-a)*(1-g) when g >= 0.5.
__global__ __launch_bounds__(256) void k_sel_finish(SelState* st, double gamma, double* out) {
    __shared__ long long cum[2][256];
    __shared__ unsigned long long sp[2], sm[2];
    __shared__ long long sk[2];
    sel_advance(st, 7, 0, cum, sp, sm, sk);
    if (threadIdx.x != 0) return;
    const double a = key_value(sp[0]);
    const double b = key_value(sp[1]);
    const double diff = b - a;
    double r = gamma >= 0.5 ? b - diff * (1 - gamma) : a + diff * gamma;
    if (st->nan_count) r = __longlong_as_double(0x7ff8000000000000ll);
    st->result = r;
    if (out) *out = r;
}

int percentile_device(const double* f, int64_t n, double q, SelState* st, double* out,
                      hipStream_t s) {
    NBX_REQUIRE(n >= 1, "percentile of an empty array");
    NBX_REQUIRE(q >= 0 && q <= 100, "percentile must be in [0, 100]");
    // numpy 'linear': virtual index (n-1) * (q/100); neighbours clamped to the
    // last element at or above n-1
    const double vi = static_cast<double>(n - 1) * (q / 100.0);
    const double lo = floor(vi);
    long long klo = static_cast<long long>(lo);
    long long khi = klo + 1;
    if (vi >= static_cast<double>(n - 1)) klo = khi = n - 1;
    if (khi > n - 1) khi = n - 1;
    const double gamma = vi - lo;
    k_sel_init<<<1, 256, 0, s>>>(st, klo, khi);
    NBX_LAUNCHED("sel_init");
    // about 8 keys per thread: every block adds its nonzero bins to the same
    // 512 global counters, and on the first passes nearly all keys share a
    // few bins, so fewer blocks serialise fewer atomics on each counter
    unsigned nb = blocks_for(n, 256 * 8);
    if (nb > 1024) nb = 1024;
    for (int pass = 0; pass < 8; ++pass) {
        k_sel_hist<<<nb, 256, 0, s>>>(f, n, st, pass);
        NBX_LAUNCHED("sel_hist");
    }
    k_sel_finish<<<1, 256, 0, s>>>(st, gamma, out);
    NBX_LAUNCHED("sel_finish");
    return 0;
}

// ---------------------------------------------------------------------------
// Ridge predicate (ridges.py:52-74): flag[k] = 1 and pts[k] = sub-pixel ridge
// point where the test passes, flag[k] = 0 elsewhere.
__global__ __launch_bounds__(256) void k_ridge_predicate(
    const double* __restrict__ f, const double2* __restrict__ evec, const double* __restrict__ xs,
    const double* __restrict__ ys, int nx, int ny, double dx, double dy, double sdd_thresh,
    const double* __restrict__ fmin_dev, int32_t* __restrict__ flag, double2* __restrict__ pts,
    double* __restrict__ sdd_out, int32_t* __restrict__ uf_par) {
    // the reference's expressions in its source order, uncontracted: ridge
    // points and second derivatives equal the oracle's bit for bit, so the
    // ordered-ridge linking (which compares them) takes the same decisions
#pragma clang fp contract(off)
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const long k = static_cast<long>(i) * ny + j;
    int32_t hit = 0;
    double sdd = 0.0;
    if (i >= 2 && i < nx - 2 && j >= 2 && j < ny - 2) {
        const double f_min = fmin_dev ? *fmin_dev : 0.0;
        const double f0 = f[k];
        if (f0 > f_min) {
            const double fe = f[k + ny], fw = f[k - ny], fn = f[k + 1], fs = f[k - 1];
            const double fx = (fe - fw) / (2 * dx);
            const double fy = (fn - fs) / (2 * dy);
            const double fxx = (fe - 2 * f0 + fw) / (dx * dx);
            const double fyy = (fn - 2 * f0 + fs) / (dy * dy);
            const double fxy = (f[k + ny + 1] - f[k + ny - 1] - f[k - ny + 1] + f[k - ny - 1]) /
                               (4 * dx * dy);
            const double2 e = evec[k];
            const double ex = e.x, ey = e.y;
            const double c2 = ex * (fxx * ex + fxy * ey) + ey * (fxy * ex + fyy * ey);
            if (c2 < -sdd_thresh) {
                const double t = -(fx * ex + fy * ey) / c2;
                if (fabs(t * ex) <= dx / 2 && fabs(t * ey) <= dy / 2) {
                    pts[k] = make_double2(xs[i] + t * ex, ys[j] + t * ey);
                    hit = 1;
                    sdd = c2;
                }
            }
        }
    }
    flag[k] = hit;
    if (sdd_out != nullptr) sdd_out[k] = sdd;
    // union-find start (each ridge point its own root), when labelling follows
    if (uf_par != nullptr) uf_par[k] = hit ? static_cast<int32_t>(k) : -1;
}

// ---------------------------------------------------------------------------
// Exclusive scan of int32 values, or of int32 pairs scanned together (sums
// must fit in int32).

struct I2 {
    int32_t a, b;
};
__device__ __forceinline__ I2 operator+(I2 x, I2 y) { return I2{x.a + y.a, x.b + y.b}; }
__device__ __forceinline__ I2& operator+=(I2& x, I2 y) {
    x.a += y.a;
    x.b += y.b;
    return x;
}

template <class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* sh, T* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const T add = t >= off ? sh[t - off] : T{};
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const T ex = t > 0 ? sh[t - 1] : T{};
    *total = sh[kScanThreads - 1];
    __syncthreads();
    return ex;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const T* __restrict__ in, int64_t n,
                                                             T* __restrict__ out,
                                                             T* __restrict__ tile_sums) {
    __shared__ T sh[kScanThreads];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    T v[kScanItems];
    T sum{};
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        v[q] = base + q < n ? in[base + q] : T{};
        sum += v[q];
    }
    T total;
    T run = block_exclusive_scan(sum, sh, &total);
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// single block: exclusive scan of the tile sums in place; grand total to *total
template <class T>
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(T* __restrict__ sums, int64_t m,
                                                            T* __restrict__ total) {
    __shared__ T sh[kScanThreads];
    T carry{};
    for (int64_t base = 0; base < m; base += kScanThreads) {
        const int64_t k = base + threadIdx.x;
        const T v = k < m ? sums[k] : T{};
        T chunk;
        const T ex = block_exclusive_scan(v, sh, &chunk);
        if (k < m) sums[k] = carry + ex;
        carry += chunk;
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class T>
__global__ __launch_bounds__(256) void k_scan_add(T* __restrict__ out, int64_t n,
                                                  const T* __restrict__ tile_sums) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) out[k] += tile_sums[k / kScanTile];
}

template <class T>
__global__ void k_set_zero(T* p) {
    *p = T{};
}

// out = exclusive scan of in; *total (device) = sum.  tmp holds
// ceil(n / 1024) values of T.
template <class T>
int exclusive_scan(const T* in, int64_t n, T* out, T* tmp, T* total, hipStream_t s) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (n == 0) {
        k_set_zero<<<1, 1, 0, s>>>(total);
        NBX_LAUNCHED("set_zero");
        return 0;
    }
    k_scan_tiles<<<static_cast<unsigned>(tiles), kScanThreads, 0, s>>>(in, n, out, tmp);
    NBX_LAUNCHED("scan_tiles");
    k_scan_sums<<<1, kScanThreads, 0, s>>>(tmp, tiles, total);
    NBX_LAUNCHED("scan_sums");
    k_scan_add<<<blocks_for(n, 256), 256, 0, s>>>(out, n, tmp);
    NBX_LAUNCHED("scan_add");
    return 0;
}

// The first two phases only: out holds each value's exclusive scan within
// its 1024-value tile and tile_off the tiles' offsets, so a consumer reads
// out[k] + tile_off[k / kScanTile] and no third pass is launched.
template <class T>
int exclusive_scan_tiled(const T* in, int64_t n, T* out, T* tile_off, T* total, hipStream_t s) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (n == 0) {
        k_set_zero<<<1, 1, 0, s>>>(total);
        NBX_LAUNCHED("set_zero");
        return 0;
    }
    k_scan_tiles<<<static_cast<unsigned>(tiles), kScanThreads, 0, s>>>(in, n, out, tile_off);
    NBX_LAUNCHED("scan_tiles");
    k_scan_sums<<<1, kScanThreads, 0, s>>>(tile_off, tiles, total);
    NBX_LAUNCHED("scan_sums");
    return 0;
}

// ---------------------------------------------------------------------------
// Compaction of flagged points in raveled order.
// pos is a complete exclusive scan, or with tile_off the tiled form.
__global__ void k_compact(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                          const double2* __restrict__ pts_dense, int64_t n,
                          double2* __restrict__ pts_out, int32_t* __restrict__ idx_out,
                          const int32_t* __restrict__ tile_off) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n || !flag[k]) return;
    const int32_t p = pos[k] + (tile_off ? tile_off[k / kScanTile] : 0);
    if (pts_out) pts_out[p] = pts_dense[k];
    if (idx_out) idx_out[p] = static_cast<int32_t>(k);
}

// ---------------------------------------------------------------------------
// Union-find labelling.
__device__ __forceinline__ int32_t uf_load(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t uf_root(const int32_t* par, int32_t a) {
    int32_t p = uf_load(par + a);
    while (p != a) {
        a = p;
        p = uf_load(par + a);
    }
    return a;
}

__device__ void uf_unite(int32_t* par, int32_t a, int32_t b) {
    for (;;) {
        a = uf_root(par, a);
        b = uf_root(par, b);
        if (a == b) return;
        if (a < b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        // a > b: hang root a under b; if a stopped being a root meanwhile,
        // continue with what it now points to
        const int32_t old = atomicMin(par + a, b);
        if (old == a) return;
        a = old;
    }
}

__global__ void k_uf_init(const int32_t* __restrict__ flag, int64_t n, int32_t* __restrict__ par) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) par[k] = flag[k] ? static_cast<int32_t>(k) : -1;
}

__global__ __launch_bounds__(256) void k_uf_union(const int32_t* __restrict__ flag, int nx, int ny,
                                                  int32_t* par) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const int32_t k = i * ny + j;
    if (!flag[k]) return;
    if (j > 0 && flag[k - 1]) uf_unite(par, k, k - 1);
    if (i > 0) {
        const int32_t up = k - ny;
        if (j > 0 && flag[up - 1]) uf_unite(par, k, up - 1);
        if (flag[up]) uf_unite(par, k, up);
        if (j + 1 < ny && flag[up + 1]) uf_unite(par, k, up + 1);
    }
}

__global__ void k_uf_flatten(const int32_t* __restrict__ flag, int64_t n, int32_t* par) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n && flag[k]) par[k] = uf_root(par, static_cast<int32_t>(k));
}

__global__ void k_bytes_to_flags(const uint8_t* __restrict__ b, int64_t n,
                                 int32_t* __restrict__ flag) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) flag[k] = b[k] ? 1 : 0;
}

__global__ void k_is_root(const int32_t* __restrict__ par, int64_t n, int32_t* __restrict__ out) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) out[k] = par[k] == static_cast<int32_t>(k) ? 1 : 0;
}

__global__ void k_label_out(const int32_t* __restrict__ par, const int32_t* __restrict__ root_rank,
                            int64_t n, int32_t* __restrict__ labels) {
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k < n) labels[k] = par[k] >= 0 ? root_rank[par[k]] + 1 : 0;
}

// ---------------------------------------------------------------------------
// Bitonic sort of 64-bit keys; N is a power of two and a multiple of
// kSortChunk.

__device__ __forceinline__ void cmp_swap(unsigned long long& a, unsigned long long& b, bool up) {
    if ((a > b) == up) {
        const unsigned long long t = a;
        a = b;
        b = t;
    }
}

// all stages k <= kSortChunk inside one chunk
__global__ __launch_bounds__(kSortThreads) void k_bitonic_chunk_full(unsigned long long* keys) {
    __shared__ unsigned long long sh[kSortChunk];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortChunk;
    const int t = threadIdx.x;
    sh[t] = keys[base + t];
    sh[t + kSortThreads] = keys[base + t + kSortThreads];
    __syncthreads();
    for (int k = 2; k <= kSortChunk; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const int lo = 2 * j * (t / j) + (t % j);
            const bool up = (((base + lo) & k) == 0);
            cmp_swap(sh[lo], sh[lo + j], up);
            __syncthreads();
        }
    }
    keys[base + t] = sh[t];
    keys[base + t + kSortThreads] = sh[t + kSortThreads];
}

// stage k (> kSortChunk), steps j < kSortChunk inside one chunk
__global__ __launch_bounds__(kSortThreads) void k_bitonic_chunk_tail(unsigned long long* keys,
                                                                     int64_t k) {
    __shared__ unsigned long long sh[kSortChunk];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortChunk;
    const int t = threadIdx.x;
    sh[t] = keys[base + t];
    sh[t + kSortThreads] = keys[base + t + kSortThreads];
    __syncthreads();
    const bool up = ((base & k) == 0);
    for (int j = kSortChunk >> 1; j > 0; j >>= 1) {
        const int lo = 2 * j * (t / j) + (t % j);
        cmp_swap(sh[lo], sh[lo + j], up);
        __syncthreads();
    }
    keys[base + t] = sh[t];
    keys[base + t + kSortThreads] = sh[t + kSortThreads];
}

__global__ void k_bitonic_global(unsigned long long* keys, int64_t N, int64_t k, int64_t j) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= N / 2) return;
    const int64_t lo = 2 * j * (t / j) + (t % j);
    const bool up = ((lo & k) == 0);
    unsigned long long a = keys[lo], b = keys[lo + j];
    if ((a > b) == up) {
        keys[lo] = b;
        keys[lo + j] = a;
    }
}

// Two global steps of stage k in one pass, j and h = j/2: thread t owns the
// four keys lo, lo+h, lo+j, lo+j+h (lo with bits j and h clear), a set that
// both steps map onto itself, and all four share the direction of bit k.
__global__ void k_bitonic_global2(unsigned long long* keys, int64_t N, int64_t k, int64_t h) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= N / 4) return;
    const int64_t lo = 4 * h * (t / h) + (t % h);
    const bool up = ((lo & k) == 0);
    unsigned long long a0 = keys[lo], a1 = keys[lo + h], a2 = keys[lo + 2 * h],
                       a3 = keys[lo + 3 * h];
    cmp_swap(a0, a2, up);
    cmp_swap(a1, a3, up);
    cmp_swap(a0, a1, up);
    cmp_swap(a2, a3, up);
    keys[lo] = a0;
    keys[lo + h] = a1;
    keys[lo + 2 * h] = a2;
    keys[lo + 3 * h] = a3;
}

int bitonic_sort(unsigned long long* keys, int64_t N, hipStream_t s) {
    const unsigned chunks = static_cast<unsigned>(N / kSortChunk);
    k_bitonic_chunk_full<<<chunks, kSortThreads, 0, s>>>(keys);
    NBX_LAUNCHED("bitonic_chunk_full");
    for (int64_t k = 2 * kSortChunk; k <= N; k <<= 1) {
        int64_t j = k >> 1;
        for (; j >= 2 * kSortChunk; j >>= 2) {
            k_bitonic_global2<<<blocks_for(N / 4, 256), 256, 0, s>>>(keys, N, k, j >> 1);
            NBX_LAUNCHED("bitonic_global2");
        }
        if (j == kSortChunk) {
            k_bitonic_global<<<blocks_for(N / 2, 256), 256, 0, s>>>(keys, N, k, j);
            NBX_LAUNCHED("bitonic_global");
        }
        k_bitonic_chunk_tail<<<chunks, kSortThreads, 0, s>>>(keys, k);
        NBX_LAUNCHED("bitonic_chunk_tail");
    }
    return 0;
}

__global__ void k_offsets_zero(int64_t* __restrict__ offsets) { offsets[0] = 0; }

__global__ void k_make_keys(const int32_t* __restrict__ idx, const int32_t* __restrict__ par,
                            const int32_t* __restrict__ count_dev, int64_t N,
                            unsigned long long* __restrict__ keys) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= N) return;
    const int32_t cnt = *count_dev;
    if (t < cnt) {
        const int32_t k = idx[t];
        keys[t] = (static_cast<unsigned long long>(static_cast<uint32_t>(par[k])) << 32) |
                  static_cast<uint32_t>(k);
    } else {
        keys[t] = ~0ull;
    }
}

// seg_start[t] = 1 where a new component begins in the sorted keys
__global__ void k_seg_start(const unsigned long long* __restrict__ keys,
                            const int32_t* __restrict__ count_dev, int64_t cap,
                            int32_t* __restrict__ seg_start) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= cap) return;
    const int32_t cnt = *count_dev;
    int32_t v = 0;
    if (t < cnt) v = (t == 0 || (keys[t] >> 32) != (keys[t - 1] >> 32)) ? 1 : 0;
    seg_start[t] = v;
}

// seg_first[s] = position of the first key of segment s; seg_first[nseg] = cnt
// seg_ex/seg_tile: the tiled exclusive scan of seg_start
__global__ void k_seg_first(const int32_t* __restrict__ seg_start,
                            const int32_t* __restrict__ seg_ex,
                            const int32_t* __restrict__ seg_tile,
                            const int32_t* __restrict__ count_dev,
                            const int32_t* __restrict__ nseg_dev, int64_t cap,
                            int32_t* __restrict__ seg_first) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t < cap && t < *count_dev && seg_start[t])
        seg_first[seg_ex[t] + seg_tile[t / kScanTile]] = static_cast<int32_t>(t);
    if (t == 0) seg_first[*nseg_dev] = *count_dev;
}

// kept ridge sizes: len or 0
__global__ void k_seg_keep(const int32_t* __restrict__ seg_first, const int32_t* __restrict__ nseg_dev,
                           int64_t cap, int min_pts, I2* __restrict__ kept) {
    const int64_t s = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (s >= cap) return;
    int32_t len = 0;
    if (s < *nseg_dev) len = seg_first[s + 1] - seg_first[s];
    const bool keep = len >= min_pts && len > 0;
    kept[s] = I2{keep ? len : 0, keep ? 1 : 0};
}

__global__ void k_group_scatter(const unsigned long long* __restrict__ keys,
                                const int32_t* __restrict__ seg_start,
                                const int32_t* __restrict__ seg_ex,
                                const int32_t* __restrict__ seg_tile,
                                const int32_t* __restrict__ seg_first,
                                const I2* __restrict__ kept,
                                const I2* __restrict__ out_off_idx,
                                const I2* __restrict__ out_tile,
                                const I2* __restrict__ totals,
                                const double2* __restrict__ pts_dense,
                                const int32_t* __restrict__ count_dev, int64_t cap,
                                double2* __restrict__ pts_out, int64_t* __restrict__ offsets) {
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t == 0) {
        const I2 tot = *totals;  // (points kept, ridges kept)
        offsets[tot.b] = tot.a;
    }
    if (t >= cap || t >= *count_dev) return;
    // segment of t: inclusive count of starts up to t, minus one
    const int32_t s = seg_ex[t] + seg_tile[t / kScanTile] + seg_start[t] - 1;
    if (!kept[s].b) return;
    const I2 o = out_off_idx[s] + out_tile[s / kScanTile];  // output offset, ridge number
    const int32_t k = static_cast<int32_t>(keys[t] & 0xffffffffull);
    const int32_t r = o.a + (static_cast<int32_t>(t) - seg_first[s]);
    pts_out[r] = pts_dense[k];
    if (t == seg_first[s]) offsets[o.b] = o.a;
}

// Bump allocator over one workspace block.
struct Arena {
    char* base;
    size_t used;
    template <class T>
    T* take(int64_t count) {
        const size_t bytes = (sizeof(T) * static_cast<size_t>(count) + 255) & ~static_cast<size_t>(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        return p;
    }
};

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Read `count` (<= 16) int32 results back to the host: one DMA into a pinned
// staging buffer per host thread (allocated on the thread's first call and
// kept, 64 bytes), then a wait on the stream.  A pageable destination would
// make the runtime stage each copy itself, tens of microseconds apiece.
int read_counts(const int32_t* dev, int count, int32_t* host, hipStream_t s) {
    thread_local int32_t* pinned = nullptr;
    if (pinned == nullptr) NBX_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned), 64));
    NBX_HIP(hipMemcpyAsync(pinned, dev, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    memcpy(host, pinned, sizeof(int32_t) * count);
    return 0;
}

// Ridge predicate + optional percentile; leaves flag/pts_dense in the arena.
int run_predicate(const double* f, const double* evec, const double* x, const double* y, int nx,
                  int ny, double dx, double dy, double sdd_thresh, double percentile,
                  SelState* st, int32_t* flag, double2* pts_dense, hipStream_t s,
                  double* sdd = nullptr, int32_t* uf_par = nullptr) {
    const int64_t n = static_cast<int64_t>(nx) * ny;
    const double* fmin_dev = nullptr;
    if (percentile != 0) {
        int rc = percentile_device(f, n, percentile, st, nullptr, s);
        if (rc) return rc;
        fmin_dev = &st->result;
    }
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_ridge_predicate<<<grid, block, 0, s>>>(f, reinterpret_cast<const double2*>(evec), x, y, nx,
                                             ny, dx, dy, sdd_thresh, fmin_dev, flag, pts_dense,
                                             sdd, uf_par);
    NBX_LAUNCHED("ridge_predicate");
    return 0;
}

// Host linking and joining shared by both ordered-ridge entry points.
int link_and_join(const ordered::RidgeField& F, double dist_tol, double ep_tan_ang,
                  int min_ridge_pts, double sdd_thresh, double c, double* pts_host,
                  int64_t pts_cap, int64_t* offsets_host, int64_t offsets_cap,
                  int64_t* nridges_host, int64_t* npts_host) {
    const ordered::Linked L = ordered::link(F, sdd_thresh, c);
    std::vector<double> out;
    std::vector<int64_t> offsets;
    ordered::join(L, dist_tol, ep_tan_ang, min_ridge_pts, out, offsets);
    const int64_t nr = static_cast<int64_t>(offsets.size()) - 1;
    const int64_t np_ = static_cast<int64_t>(out.size() / 2);
    *nridges_host = nr;
    *npts_host = np_;
    NBX_REQUIRE(np_ <= pts_cap && nr + 1 <= offsets_cap,
                "output buffers too small (the counts hold the sizes needed)");
    std::copy(out.begin(), out.end(), pts_host);
    std::copy(offsets.begin(), offsets.end(), offsets_host);
    return 0;
}

int check_grid(int nx, int ny) {
    NBX_REQUIRE(nx >= 0 && ny >= 0, "negative grid size");
    NBX_REQUIRE(static_cast<int64_t>(nx) * ny < (static_cast<int64_t>(1) << 31) - 1,
                "grid too large for 32-bit point indices");
    return 0;
}

constexpr size_t kNoAllocScratchBytes = sizeof(scratch);  // 64 MiB
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_addr[kMaxDevices] = {};  // per-device address of `scratch`

int scratch_address(char** out) {
    int dev = 0;
    NBX_HIP(hipGetDevice(&dev));
    NBX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_addr[dev] == nullptr) {
        void* p = nullptr;
        NBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NBX_REQUIRE(p != nullptr, "hipGetSymbolAddress returned NULL for the scratch array");
        g_scratch_addr[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_addr[dev];
    return 0;
}

// Scratch layout of the grouped-ridge pipeline for n grid points.
struct RidgeLayout {
    SelState* st;
    // total[0..1] = the pair scan's (points, ridges) total, total[2] = count,
    // the number of ridge points: one read-back fetches all three
    int32_t *flag, *pos, *tmp, *count, *par, *idx, *seg_start, *seg_ex, *seg_first, *nseg,
        *total;
    I2 *kept, *out_off_idx, *tmp2;
    double2* dense;
    unsigned long long* keys;
    void carve(Arena& A, int64_t n) {
        const int64_t tiles = (n + kScanTile - 1) / kScanTile;
        int64_t N = kSortChunk;
        while (N < n) N <<= 1;
        st = A.take<SelState>(1);
        flag = A.take<int32_t>(n);
        pos = A.take<int32_t>(n);
        dense = A.take<double2>(n);
        tmp = A.take<int32_t>(tiles + 1);   // tile offsets: compaction, then segments
        tmp2 = A.take<I2>(tiles + 1);       // tile offsets of the pair scan
        par = A.take<int32_t>(n);
        idx = A.take<int32_t>(n);
        keys = A.take<unsigned long long>(N);
        seg_start = A.take<int32_t>(n);
        seg_ex = A.take<int32_t>(n);
        seg_first = A.take<int32_t>(n + 1);
        kept = A.take<I2>(n);         // (kept length, kept flag) per segment
        out_off_idx = A.take<I2>(n);  // their exclusive scans
        nseg = A.take<int32_t>(1);
        total = A.take<int32_t>(3);
        count = total + 2;
    }
    static size_t bytes(int64_t n) {
        RidgeLayout L;
        Arena measure{nullptr, 0};
        L.carve(measure, n);
        return measure.used;
    }
};

// Count read-back without the pinned staging buffer (for callers that must
// not allocate): a copy into pageable host memory, then a wait.
int read_counts_pageable(const int32_t* dev, int count, int32_t* host, hipStream_t s) {
    NBX_HIP(hipMemcpyAsync(host, dev, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    return 0;
}

// Counts of the asynchronous form: (ridges kept, points kept, ridge points
// before the length filter); -1 ridges if the points exceeded the grouping
// capacity m (not reachable: m bounds them, see ridges_run).
// `grouped` = 0 when the grouping stage did not run (capacity 0): its totals
// are then not this call's.
__global__ void k_counts_out(const int32_t* __restrict__ total, int64_t m, int grouped,
                             int64_t* __restrict__ counts) {
    if (threadIdx.x == 0) {
        const int32_t all = total[2];
        const int32_t kept = grouped ? total[0] : 0, ridges = grouped ? total[1] : 0;
        counts[0] = all > m ? -1 : ridges;
        counts[1] = kept;
        counts[2] = all;
    }
}

// (a kernel, not a memset: the call may be captured in a graph)
__global__ void k_counts_clear(int64_t* __restrict__ counts) {
    if (threadIdx.x < 3) counts[threadIdx.x] = 0;
}

// ftle_ridges on a scratch arena of at least RidgeLayout::bytes(nx*ny)
// bytes at `arena`; n = nx*ny > 0 and the arguments already checked.  With
// counts_dev set the counts go to device memory and nothing waits on the
// stream (then the grouping capacity without a percentile is the grid size:
// no count is read back).
int ridges_run(const double* f, const double* evec, const double* x, int nx, const double* y,
               int ny, double dx, double dy, double sdd_thresh, double percentile,
               int min_ridge_pts, double* pts, int64_t* offsets, int64_t* nridges_host,
               int64_t* npts_host, char* arena, bool pinned_counts, hipStream_t s,
               int64_t* counts_dev = nullptr) {
    const int64_t n = static_cast<int64_t>(nx) * ny;
    auto read = [&](const int32_t* dev, int count, int32_t* host) {
        return pinned_counts ? read_counts(dev, count, host, s)
                             : read_counts_pageable(dev, count, host, s);
    };
    RidgeLayout L;
    Arena A{arena, 0};
    L.carve(A, n);
    SelState* st = L.st;
    int32_t *flag = L.flag, *pos = L.pos, *tmp = L.tmp, *count = L.count, *par = L.par,
            *idx = L.idx, *seg_start = L.seg_start, *seg_ex = L.seg_ex,
            *seg_first = L.seg_first, *nseg = L.nseg, *total = L.total;
    double2* dense = L.dense;
    unsigned long long* keys = L.keys;

    if (int rc = run_predicate(f, evec, x, y, nx, ny, dx, dy, sdd_thresh, percentile, st, flag,
                               dense, s, nullptr, par))
        return rc;
    // label components (the predicate started the union-find)
    const unsigned nb = blocks_for(n, 256);
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_uf_union<<<grid, block, 0, s>>>(flag, nx, ny, par);
    NBX_LAUNCHED("uf_union");
    k_uf_flatten<<<nb, 256, 0, s>>>(flag, n, par);
    NBX_LAUNCHED("uf_flatten");
    // compact ridge points (raveled order)
    if (int rc = exclusive_scan_tiled(flag, n, pos, tmp, count, s)) return rc;
    k_compact<<<nb, 256, 0, s>>>(flag, pos, dense, n, nullptr, idx, tmp);
    NBX_LAUNCHED("compact");
    // The grouping stage works on the ridge points only (typically a few
    // percent of the grid), sized by a bound on their count; its kernels read
    // the exact count on the device.  With a percentile threshold q the bound
    // needs no read-back: a ridge point has f > f_min >= the order statistic
    // of rank floor((n-1) q/100), which at most n-1-that rank values exceed,
    // so the whole pipeline stays queued behind the work before it.  Without
    // one the count itself is read back.
    int64_t m = 0;
    if (percentile != 0) {
        const double vi = static_cast<double>(n - 1) * (percentile / 100.0);
        int64_t klo = static_cast<int64_t>(floor(vi));
        if (vi >= static_cast<double>(n - 1)) klo = n - 1;
        m = n - 1 - klo;
    } else if (counts_dev) {
        m = n;
    } else {
        int32_t cnt = 0;
        if (int rc = read(count, 1, &cnt)) return rc;
        m = cnt;
    }
    if (m == 0) {
        k_offsets_zero<<<1, 1, 0, s>>>(offsets);
        NBX_LAUNCHED("offsets_zero");
        if (counts_dev) {
            // no ridge point passes (the grouping totals were not written)
            k_counts_out<<<1, 64, 0, s>>>(total, 0, 0, counts_dev);
            NBX_LAUNCHED("counts_out");
            return 0;
        }
        NBX_HIP(hipStreamSynchronize(s));
        return 0;
    }
    int64_t Ns = kSortChunk;
    while (Ns < m) Ns <<= 1;
    const unsigned nbm = blocks_for(m, 256);
    // order by (component root, raveled index)
    k_make_keys<<<blocks_for(Ns, 256), 256, 0, s>>>(idx, par, count, Ns, keys);
    NBX_LAUNCHED("make_keys");
    if (int rc = bitonic_sort(keys, Ns, s)) return rc;
    // segments = components
    k_seg_start<<<nbm, 256, 0, s>>>(keys, count, m, seg_start);
    NBX_LAUNCHED("seg_start");
    if (int rc = exclusive_scan_tiled(seg_start, m, seg_ex, tmp, nseg, s)) return rc;
    k_seg_first<<<nbm, 256, 0, s>>>(seg_start, seg_ex, tmp, count, nseg, m, seg_first);
    NBX_LAUNCHED("seg_first");
    k_seg_keep<<<nbm, 256, 0, s>>>(seg_first, nseg, m, min_ridge_pts, L.kept);
    NBX_LAUNCHED("seg_keep");
    // total and nridges are adjacent: the pair scan's total lands in both
    I2* totals = reinterpret_cast<I2*>(total);
    if (int rc = exclusive_scan_tiled(L.kept, m, L.out_off_idx, L.tmp2, totals, s)) return rc;
    // the segment scan's tile offsets (tmp) are still intact: the pair scan
    // used tmp2; the scatter also writes offsets[nridges] = total
    k_group_scatter<<<nbm, 256, 0, s>>>(keys, seg_start, seg_ex, tmp, seg_first, L.kept,
                                        L.out_off_idx, L.tmp2, totals, dense, count, m,
                                        reinterpret_cast<double2*>(pts), offsets);
    NBX_LAUNCHED("group_scatter");
    if (counts_dev) {
        k_counts_out<<<1, 64, 0, s>>>(total, m, 1, counts_dev);
        NBX_LAUNCHED("counts_out");
        return 0;
    }
    int32_t h[3] = {0, 0, 0};  // points kept, ridges kept, ridge points (adjacent)
    if (int rc = read(total, 3, h)) return rc;
    // the grouping kernels clip to the capacity m; the percentile bound
    // guarantees count <= m, and a violation is reported, not truncated
    NBX_REQUIRE(h[2] <= m, "ridge-point count exceeds the grouping capacity");
    *nridges_host = h[1];
    *npts_host = h[0];
    return 0;
}


}  // namespace
}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_percentile(const double* f, int64_t n, double q, double* result, nbx_stream stream) {
    NBX_REQUIRE(f != nullptr && result != nullptr, "NULL array");
    hipStream_t s = as_stream(stream);
    Workspace ws(s, 2, sizeof(SelState));
    if (ws.error()) return ws.error();
    return percentile_device(f, n, q, reinterpret_cast<SelState*>(ws.get()), result, s);
}

int nbx_ridge_pts(const double* f, const double* evec, const double* x, int nx, const double* y,
                  int ny, double dx, double dy, double sdd_thresh, double percentile, double* pts,
                  int64_t* npts_host, nbx_stream stream) {
    if (int rc = check_grid(nx, ny)) return rc;
    NBX_REQUIRE(npts_host != nullptr, "npts_host is NULL");
    const int64_t n = static_cast<int64_t>(nx) * ny;
    *npts_host = 0;
    if (n == 0) return 0;
    NBX_REQUIRE(f && evec && x && y && pts, "NULL array");
    hipStream_t s = as_stream(stream);
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    const size_t bytes = align256(sizeof(SelState)) + align256(4 * n) * 2 + align256(16 * n) +
                         align256(4 * (tiles + 1)) + align256(4);
    Workspace ws(s, 1, bytes);
    if (ws.error()) return ws.error();
    Arena A{ws.get(), 0};
    auto* st = A.take<SelState>(1);
    auto* flag = A.take<int32_t>(n);
    auto* pos = A.take<int32_t>(n);
    auto* dense = A.take<double2>(n);
    auto* tmp = A.take<int32_t>(tiles + 1);
    auto* total = A.take<int32_t>(1);
    if (int rc = run_predicate(f, evec, x, y, nx, ny, dx, dy, sdd_thresh, percentile, st, flag,
                               dense, s))
        return rc;
    if (int rc = exclusive_scan(flag, n, pos, tmp, total, s)) return rc;
    k_compact<<<blocks_for(n, 256), 256, 0, s>>>(flag, pos, dense, n,
                                                 reinterpret_cast<double2*>(pts), nullptr, nullptr);
    NBX_LAUNCHED("compact");
    int32_t cnt = 0;
    if (int rc = read_counts(total, 1, &cnt, s)) return rc;
    *npts_host = cnt;
    return 0;
}

int nbx_ridges(const double* f, const double* evec, const double* x, int nx, const double* y,
               int ny, double dx, double dy, double sdd_thresh, double percentile,
               int min_ridge_pts, double* pts, int64_t* offsets, int64_t* nridges_host,
               int64_t* npts_host, nbx_stream stream) {
    if (int rc = check_grid(nx, ny)) return rc;
    NBX_REQUIRE(nridges_host != nullptr && npts_host != nullptr, "NULL count pointer");
    const int64_t n = static_cast<int64_t>(nx) * ny;
    *nridges_host = 0;
    *npts_host = 0;
    if (n == 0) return 0;
    NBX_REQUIRE(f && evec && x && y && pts && offsets, "NULL array");
    hipStream_t s = as_stream(stream);
    Workspace ws(s, 1, RidgeLayout::bytes(n));
    if (ws.error()) return ws.error();
    return ridges_run(f, evec, x, nx, y, ny, dx, dy, sdd_thresh, percentile, min_ridge_pts, pts,
                      offsets, nridges_host, npts_host, ws.get(), true, s);
}

int nbx_ridges_async(const double* f, const double* evec, const double* x, int nx,
                     const double* y, int ny, double dx, double dy, double sdd_thresh,
                     double percentile, int min_ridge_pts, double* pts, int64_t* offsets,
                     int64_t* counts, nbx_stream stream) {
    if (int rc = check_grid(nx, ny)) return rc;
    NBX_REQUIRE(counts != nullptr, "NULL counts");
    const int64_t n = static_cast<int64_t>(nx) * ny;
    hipStream_t s = as_stream(stream);
    if (n == 0) {
        k_counts_clear<<<1, 64, 0, s>>>(counts);
        NBX_LAUNCHED("counts_clear");
        return 0;
    }
    NBX_REQUIRE(f && evec && x && y && pts && offsets, "NULL array");
    Workspace ws(s, 1, RidgeLayout::bytes(n));
    if (ws.error()) return ws.error();
    int64_t unused_r = 0, unused_p = 0;
    return ridges_run(f, evec, x, nx, y, ny, dx, dy, sdd_thresh, percentile, min_ridge_pts, pts,
                      offsets, &unused_r, &unused_p, ws.get(), true, s, counts);
}

int nbx_ridges_noalloc(const double* f, const double* evec, const double* x, int nx,
                       const double* y, int ny, double dx, double dy, double sdd_thresh,
                       double percentile, int min_ridge_pts, double* pts, int64_t* offsets,
                       int64_t* nridges_host, int64_t* npts_host, nbx_stream stream) {
    if (int rc = check_grid(nx, ny)) return rc;
    NBX_REQUIRE(nridges_host != nullptr && npts_host != nullptr, "NULL count pointer");
    const int64_t n = static_cast<int64_t>(nx) * ny;
    *nridges_host = 0;
    *npts_host = 0;
    if (n == 0) return 0;
    NBX_REQUIRE(f && evec && x && y && pts && offsets, "NULL array");
    const size_t need = RidgeLayout::bytes(n);
    if (need > kNoAllocScratchBytes) {
        set_error("nbx_ridges_noalloc: a grid of " + std::to_string(n) + " points needs " +
                  std::to_string(need) + " bytes of scratch, above its fixed " +
                  std::to_string(kNoAllocScratchBytes) + " (use nbx_ridges)");
        return -2;
    }
    char* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    // every buffer the pipeline reads is written earlier in the same call
    // (counters and histograms are cleared by its own kernels), so the
    // previous call's contents need no clearing
    return ridges_run(f, evec, x, nx, y, ny, dx, dy, sdd_thresh, percentile, min_ridge_pts, pts,
                      offsets, nridges_host, npts_host, base, false, as_stream(stream));
}

int nbx_label8(const uint8_t* b, int nx, int ny, int32_t* labels, int32_t* nlabels_host,
               nbx_stream stream) {
    if (int rc = check_grid(nx, ny)) return rc;
    NBX_REQUIRE(nlabels_host != nullptr, "nlabels_host is NULL");
    const int64_t n = static_cast<int64_t>(nx) * ny;
    *nlabels_host = 0;
    if (n == 0) return 0;
    NBX_REQUIRE(b != nullptr && labels != nullptr, "NULL array");
    hipStream_t s = as_stream(stream);
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    const size_t bytes = align256(4 * n) * 4 + align256(4 * (tiles + 1)) + align256(4);
    Workspace ws(s, 1, bytes);
    if (ws.error()) return ws.error();
    Arena A{ws.get(), 0};
    auto* flag = A.take<int32_t>(n);
    auto* par = A.take<int32_t>(n);
    auto* isroot = A.take<int32_t>(n);
    auto* rank = A.take<int32_t>(n);
    auto* tmp = A.take<int32_t>(tiles + 1);
    auto* total = A.take<int32_t>(1);
    const unsigned nb = blocks_for(n, 256);
    k_bytes_to_flags<<<nb, 256, 0, s>>>(b, n, flag);
    NBX_LAUNCHED("bytes_to_flags");
    k_uf_init<<<nb, 256, 0, s>>>(flag, n, par);
    NBX_LAUNCHED("uf_init");
    const dim3 block(64, 4), grid((ny + 63) / 64, (nx + 3) / 4);
    k_uf_union<<<grid, block, 0, s>>>(flag, nx, ny, par);
    NBX_LAUNCHED("uf_union");
    k_uf_flatten<<<nb, 256, 0, s>>>(flag, n, par);
    NBX_LAUNCHED("uf_flatten");
    k_is_root<<<nb, 256, 0, s>>>(par, n, isroot);
    NBX_LAUNCHED("is_root");
    if (int rc = exclusive_scan(isroot, n, rank, tmp, total, s)) return rc;
    k_label_out<<<nb, 256, 0, s>>>(par, rank, n, labels);
    NBX_LAUNCHED("label_out");
    int32_t cnt = 0;
    if (int rc = read_counts(total, 1, &cnt, s)) return rc;
    *nlabels_host = cnt;
    return 0;
}

int nbx_ordered_ridges(const double* f, const double* evec, const double* x, int nx,
                       const double* y, int ny, double dist_tol, double ep_tan_ang,
                       int min_ridge_pts, double sdd_thresh, double percentile, double c,
                       double* pts_host, int64_t pts_cap, int64_t* offsets_host,
                       int64_t offsets_cap, int64_t* nridges_host, int64_t* npts_host,
                       nbx_stream stream) {
    if (int rc = check_grid(nx, ny)) return rc;
    NBX_REQUIRE(nridges_host != nullptr && npts_host != nullptr, "NULL count pointer");
    *nridges_host = 0;
    *npts_host = 0;
    const int64_t n = static_cast<int64_t>(nx) * ny;
    if (n == 0) return 0;
    NBX_REQUIRE(nx >= 2 && ny >= 2, "the ridge stencil needs at least 2 points per axis");
    NBX_REQUIRE(f && evec && x && y && pts_host && offsets_host, "NULL array");
    hipStream_t s = as_stream(stream);
    const size_t bytes = align256(sizeof(SelState)) + align256(4 * n) + align256(16 * n) +
                         align256(8 * n);
    Workspace ws(s, 1, bytes);
    if (ws.error()) return ws.error();
    Arena A{ws.get(), 0};
    auto* st = A.take<SelState>(1);
    auto* flag = A.take<int32_t>(n);
    auto* dense = A.take<double2>(n);
    auto* sdd = A.take<double>(n);
    double xh[2], yh[2];
    NBX_HIP(hipMemcpyAsync(xh, x, sizeof(xh), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipMemcpyAsync(yh, y, sizeof(yh), hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    const double dx = xh[1] - xh[0], dy = yh[1] - yh[0];
    if (int rc = run_predicate(f, evec, x, y, nx, ny, dx, dy, sdd_thresh, percentile, st, flag,
                               dense, s, sdd))
        return rc;
    // the serial linking runs on the host over the dense per-point buffers
    std::vector<int32_t> fl(n);
    std::vector<double2> pt(n), ev(n);
    ordered::RidgeField F;
    F.nx = nx;
    F.ny = ny;
    F.h = dx < dy ? dx : dy;
    F.sdd.resize(n);
    NBX_HIP(hipMemcpyAsync(fl.data(), flag, 4 * n, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipMemcpyAsync(pt.data(), dense, 16 * n, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipMemcpyAsync(ev.data(), evec, 16 * n, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipMemcpyAsync(F.sdd.data(), sdd, 8 * n, hipMemcpyDeviceToHost, s));
    NBX_HIP(hipStreamSynchronize(s));
    F.px.resize(n);
    F.py.resize(n);
    F.vx.resize(n);
    F.vy.resize(n);
    for (int64_t k = 0; k < n; ++k) {
        F.px[k] = fl[k] ? pt[k].x : -1.0;
        F.py[k] = fl[k] ? pt[k].y : -1.0;
        F.vx[k] = ev[k].x;
        F.vy[k] = ev[k].y;
    }
    return link_and_join(F, dist_tol, ep_tan_ang, min_ridge_pts, sdd_thresh, c, pts_host, pts_cap,
                         offsets_host, offsets_cap, nridges_host, npts_host);
}

int nbx_ordered_ridges_link(const double* pts_dense, const double* eigvec_max, const double* sdd,
                            int nx, int ny, double h, double dist_tol, double ep_tan_ang,
                            int min_ridge_pts, double sdd_thresh, double c, double* pts_host,
                            int64_t pts_cap, int64_t* offsets_host, int64_t offsets_cap,
                            int64_t* nridges_host, int64_t* npts_host) {
    if (int rc = check_grid(nx, ny)) return rc;
    NBX_REQUIRE(nridges_host != nullptr && npts_host != nullptr, "NULL count pointer");
    *nridges_host = 0;
    *npts_host = 0;
    const int64_t n = static_cast<int64_t>(nx) * ny;
    if (n == 0) return 0;
    NBX_REQUIRE(pts_dense && eigvec_max && sdd && pts_host && offsets_host, "NULL array");
    NBX_REQUIRE(h > 0, "grid spacing must be positive");
    ordered::RidgeField F;
    F.nx = nx;
    F.ny = ny;
    F.h = h;
    F.px.resize(n);
    F.py.resize(n);
    F.vx.resize(n);
    F.vy.resize(n);
    F.sdd.assign(sdd, sdd + n);
    for (int64_t k = 0; k < n; ++k) {
        F.px[k] = pts_dense[2 * k];
        F.py[k] = pts_dense[2 * k + 1];
        F.vx[k] = eigvec_max[2 * k];
        F.vy[k] = eigvec_max[2 * k + 1];
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            NBX_REQUIRE(F.sdd[static_cast<int64_t>(i) * ny + j] >= 0 ||
                            (i >= 2 && i < nx - 2 && j >= 2 && j < ny - 2),
                        "ridge points (sdd < 0) must lie two or more points inside the grid");
    return link_and_join(F, dist_tol, ep_tan_ang, min_ridge_pts, sdd_thresh, c, pts_host, pts_cap,
                         offsets_host, offsets_cap, nridges_host, npts_host);
}

}  // extern "C"
