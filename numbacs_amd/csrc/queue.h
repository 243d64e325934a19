// Persistent DOP853 launch over a device queue of particle tiles.
//
// A wavefront's particles finish together only when their step counts
// agree, and a grid launch dispatches wavefronts in order: the slowest tiles
// that start late form the launch's tail (on the gridded workload the
// in-order schedule is ~1.28x the balanced one, and launching tiles longest
// first measured 11% faster in FP64, 19% in FP32).  Step counts are not
// known before the solve, so the queue balances instead: a wavefront pops a
// tile (one wavefront's particles), runs at most kChunkSteps attempted steps
// per particle, and when some particle is still running stores the tile's
// solver states and pushes the tile to the back of the queue.  Long tiles
// are thus spread over whichever wavefronts are free, in chunks.
//
// The queue is a ring of 64-bit entries (reservation index << 32 | payload)
// with a claim counter (head), a push counter (tail) and a count of finished
// tiles (done).  A wavefront claims ring index h by fetch-and-add on head
// (one atomic per claim, whatever the number of wavefronts: a
// compare-and-swap claim measured 50-150x slower with two thousand
// wavefronts retrying) and takes the entry of slot h once it carries
// reservation index h; until then it sleeps and polls, and it leaves when
// every tile has finished.  A claimed index is always filled unless all
// tiles finish first: every unfinished tile is either in the ring or held
// by a running wavefront that will push it, and pushes take the reservation
// indices in order.  The ring holds at least four times as many entries as
// there are tiles (each tile is in the ring at most once), so a slot's next
// reservation comes long after its entry has been taken.  Each solve is
// bounded by max_steps, so the tiles finish and every wavefront reaches the
// exit; a poll limit of about ten seconds bounds the wait regardless.
//
// The queue spans a table of requests (flow-map grids with their own window
// t0, T and particle axes), which is how a whole FTLE time series runs in one
// launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dop853.h"
#include "flows.h"

// Minimum wavefronts per SIMD requested from the compiler for the flow-map
// kernels (caps their VGPR budget at 512 / n registers).
#ifndef NBX_FLOWMAP_MIN_WAVES
#define NBX_FLOWMAP_MIN_WAVES 1
#endif

namespace nbx {

// Particle order of the one-solve-per-group grid launch: 8 x 8 tiles of the
// particle grid, each walked as its four 4 x 4 quadrants, so any run of 16,
// 32 or 64 aligned slots -- one wavefront at 4, 2 or 1 lanes per particle --
// is a compact 4 x 4, 4 x 8 or 8 x 8 tile.  A wavefront issues until its
// slowest particle finishes and step counts vary smoothly in space, so
// compact tiles idle fewer lanes than runs along y (issued solver work x1.20
// of the mean instead of x1.26 on the gridded workload) while their particles
// still share stencil rows.  Slots past the grid's edge map to no particle.
// (The work queue keeps row order: its refills already keep lanes busy, and
// tile order measured no faster there.)
__host__ __device__ inline int64_t tile_slots(int nx, int ny) {
    return static_cast<int64_t>((nx + 7) / 8) * 8 * ((ny + 7) / 8) * 8;
}
__host__ __device__ inline int64_t tile_particle(int64_t slot, int nx, int ny) {
    const int64_t t = slot >> 6, r = slot & 63, quad = r >> 4, q = r & 15;
    const int64_t ntj = (ny + 7) / 8;
    const int64_t i = (t / ntj) * 8 + (quad >> 1) * 4 + (q >> 2);
    const int64_t j = (t % ntj) * 8 + (quad & 1) * 4 + (q & 3);
    return (i < nx && j < ny) ? i * ny + j : -1;
}

// One flow-map request as the kernel sees it.
struct ReqDev {
    double s0, s1;                // p0*t0 and p0*(t0+T)
    const double* xs;             // particle axes (device)
    const double* ys;
    int nx, ny;
    const uint8_t* mask;          // nx*ny or nullptr
    double2* out;                 // (nx, ny) positions
    int32_t* status;              // nullable
    int32_t* nsteps;              // nullable
};

// Blocks per chunk of the spline tile launches' block remap (0: no remap).
#ifndef NBX_XCD_CHUNK
#define NBX_XCD_CHUNK 64
#endif

template <class T>
struct IsSpline3 : std::false_type {};
template <class R, int G, bool L, bool S>
struct IsSpline3<Spline3<R, G, L, S>> : std::true_type {};

// The work block of this block in a tile launch over a spline flow.
// Consecutive blocks go to the eight XCDs in turn; XCD x is given the work
// of chunks x, x + 8, ... of NBX_XCD_CHUNK consecutive blocks, so the tiles
// an XCD runs together are neighbours (their stencils share its own L2)
// while every XCD still takes an even share of each stretch of the launch
// order.  Only when every XCD gets at least 10 chunks (fewer, larger chunks
// are regions whose long tiles pile up on one XCD), and only for spline
// flows, where it was measured faster (the trilinear flow map was 15%
// slower with it).  A bijection on the blocks of whole groups of 8 chunks;
// the remaining blocks keep their index.
template <class Flow>
__device__ __forceinline__ unsigned xcd_block() {
    unsigned bid = blockIdx.x;
#if NBX_XCD_CHUNK > 0
    if constexpr (IsSpline3<Flow>::value) {
        constexpr unsigned C = NBX_XCD_CHUNK;
        if (gridDim.x >= 80u * C) {
            const unsigned full = gridDim.x / (8u * C) * (8u * C);
            if (bid < full) {
                const unsigned x = bid % 8u, g = bid / 8u;
                bid = ((g / C) * 8u + x) * C + g % C;
            }
        }
    }
#endif
    return bid;
}

// Attempted steps per particle before a wavefront hands its tile back.
constexpr int kChunkSteps = 16;

struct ChunkQueue {
    unsigned long long* items;  // ring: (reservation index << 32) | payload
    unsigned long long* head;   // next reservation index to pop
    unsigned long long* tail;   // next reservation index to push
    unsigned long long* done;   // tiles whose particles have all finished
    unsigned long long mask;    // ring entries - 1 (a power of two above the tile count)
    unsigned long long ntiles;
};
// Polls of an unfilled claim before a wavefront gives up waiting (about ten
// seconds at ~1 us of sleep per poll): a safety bound, far beyond any solve
// here; a wavefront that gives up drops its claim.
constexpr int kMaxPolls = 10 << 20;
// payload bit 31: the tile has been started (its states are stored)
constexpr unsigned kResume = 0x80000000u;

// Ring entries for ntiles tiles: a power of two >= 4 ntiles (at least 1024).
inline unsigned long long chunk_ring(int64_t ntiles) {
    unsigned long long n = 1024;
    while (n < 4 * static_cast<unsigned long long>(ntiles)) n <<= 1;
    return n;
}

// Entries 0..ntiles-1 hold tiles 0..ntiles-1 (reservations 0..ntiles-1);
// the others carry an index no reservation of their slot can have (the
// complement of the slot number: odd minus even never matches).  A kernel
// rather than a memset, so the reset can be captured in a graph.
static __global__ void k_chunk_init(ChunkQueue q, int64_t ntiles) {
    const unsigned long long i = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x;
    if (i > q.mask) return;
    q.items[i] = i < static_cast<unsigned long long>(ntiles)
                     ? ((i & 0xffffffffull) << 32) | i
                     : ((~i & 0xffffffffull) << 32);
    if (i == 0) {
        *q.head = 0ull;
        *q.tail = static_cast<unsigned long long>(ntiles);
        *q.done = 0ull;
    }
}

// Memory ordering without cache-wide fences.  An agent-scope acquire or
// release writes back or invalidates the whole L2 of an XCD, which drops
// the spline coefficients every wavefront there is gathering (measured:
// 40-270x slower).  Instead the handed-on solver states are stored and
// loaded word by word as agent-scope relaxed atomics, which reach the
// point of coherence of all XCDs per access; a workgroup-scope release
// fence (a wait for the wavefront's stores to complete, and a compiler
// barrier) orders them before the ring entry that publishes the tile, and
// a reader's state addresses depend on the entry it popped.
template <class T>
__device__ __forceinline__ void coherent_store(T* dst, const T& v) {
    static_assert(sizeof(T) % 8 == 0 && alignof(T) >= 8, "8-byte words");
    unsigned long long w[sizeof(T) / 8];
    __builtin_memcpy(w, &v, sizeof(T));
    auto* d = reinterpret_cast<unsigned long long*>(dst);
#pragma unroll
    for (int i = 0; i < static_cast<int>(sizeof(T) / 8); ++i)
        __hip_atomic_store(d + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T coherent_load(const T* src) {
    unsigned long long w[sizeof(T) / 8];
    const auto* d = reinterpret_cast<const unsigned long long*>(src);
#pragma unroll
    for (int i = 0; i < static_cast<int>(sizeof(T) / 8); ++i)
        w[i] = __hip_atomic_load(const_cast<unsigned long long*>(d + i), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// Claim the next ring index and wait for its tile: the payload, or -1 once
// every tile has finished (or after kMaxPolls polls).
__device__ __forceinline__ long long chunk_pop(const ChunkQueue& q) {
    const unsigned long long h = atomicAdd(q.head, 1ull);
    for (int polls = 0;; ++polls) {
        const unsigned long long e =
            __hip_atomic_load(q.items + (h & q.mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((e >> 32) == (h & 0xffffffffull)) return static_cast<long long>(e & 0xffffffffull);
        if (__hip_atomic_load(q.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= q.ntiles ||
            polls >= kMaxPolls)
            return -1;
        __builtin_amdgcn_s_sleep(32);
    }
}

__device__ __forceinline__ void chunk_push(const ChunkQueue& q, unsigned payload) {
    const unsigned long long r = atomicAdd(q.tail, 1ull);
    __hip_atomic_store(q.items + (r & q.mask), ((r & 0xffffffffull) << 32) | payload,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Number of compute units of the current device (cached per device).
int device_cu_count();

// The queue's tiles are the grid launch's wavefronts: 64 / G consecutive
// slots in tile order (a 4 x 4 particle tile at G = 4).  Slots are numbered
// across the requests (request r owns [slot_off[r], slot_off[r+1]), a
// multiple of 64 each, so no tile straddles two requests); state holds one
// solver state per slot.  Masked particles are written at their tile's first
// chunk; a particle's outputs are written when its solve ends.
template <class Flow, class R, int G>
__global__ __launch_bounds__(256, sizeof(R) == 4 ? 3 : NBX_FLOWMAP_MIN_WAVES) void k_flowmap_chunks(
    Flow f, ReqDev single, const ReqDev* __restrict__ table, const int64_t* __restrict__ slot_off,
    int nreq, ChunkQueue q, Dop853State<R>* __restrict__ state, double rtol, double atol,
    long long nmax) {
    constexpr int kSlots = 64 / G;
    // (only the factors, in FP64: this kernel's solver state, tile
    // bookkeeping and handoff leave no vector registers for the box and
    // spacing, and the FP32 instance none at its 3-wave budget)
    if constexpr (sizeof(R) == sizeof(double)) f.prepare_factors();
    const int lane = static_cast<int>(__lane_id());
    const bool lead = (lane & (G - 1)) == 0;
    for (;;) {
        long long e = 0;
        if (lane == 0) e = chunk_pop(q);
        e = __shfl(e, 0);
        if (e < 0) break;
        const unsigned u = static_cast<unsigned>(e) & ~kResume;
        const bool resume = (static_cast<unsigned>(e) & kResume) != 0;
        const int64_t slot = static_cast<int64_t>(u) * kSlots + lane / G;
        int r = 0;
        if (table != nullptr) {
            int lo = 0, hi = nreq - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (slot_off[mid] <= slot) lo = mid;
                else hi = mid - 1;
            }
            r = lo;
        }
        const ReqDev& Rq = table == nullptr ? single : table[r];
        const int64_t p =
            tile_particle(slot - (table == nullptr ? 0 : slot_off[r]), Rq.nx, Rq.ny);
        Dop853State<R> s;
        bool active = false;
        if (p >= 0) {
            if (resume) {
                s = coherent_load(state + slot);
                active = s.idid == 0;
            } else if (Rq.mask != nullptr && Rq.mask[p]) {
                if (lead) {
                    Rq.out[p] = make_double2(0.0, 0.0);
                    if (Rq.status) Rq.status[p] = 0;
                    if (Rq.nsteps) Rq.nsteps[p] = 0;
                }
                s.idid = 1;  // nothing left to do if the tile comes back
            } else {
                const int64_t i = p / Rq.ny;
                dop853_init(f, s, Rq.s0, Rq.s1, Rq.xs[i], Rq.ys[p - i * Rq.ny], rtol, atol);
                active = true;
            }
        }
        for (int n = 0; active && n < kChunkSteps; ++n) {
            if (!dop853_step(f, s, rtol, atol, nmax)) {
                if (lead) {
                    Rq.out[p] = make_double2(s.px(), s.py());
                    if (Rq.status) Rq.status[p] = s.idid;
                    if (Rq.nsteps) Rq.nsteps[p] = s.nstep;
                }
                active = false;
            }
        }
        if (__ballot(active) != 0ull) {
            if (lead && p >= 0) coherent_store(state + slot, s);
            // the states complete before the entry that hands them on
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0) chunk_push(q, u | kResume);
        } else if (lane == 0) {
            atomicAdd(q.done, 1ull);
        }
    }
}

// One non-persistent launch over a table of requests in the grid launch's
// tile order: request r owns the launch's slots [slot_off[r], slot_off[r+1])
// (tile_slots(nx, ny) of them, a multiple of 64, so every wavefront lies in
// one request) and maps them to its particles with tile_particle.  The
// batch's alternative to the work queue when its particles' spatial
// locality is worth more than the queue's refills (the autotuner times
// both: persistent occupancy 0 selects this kernel).
template <class Flow, class R, int G>
__global__ __launch_bounds__(256, NBX_FLOWMAP_MIN_WAVES) void k_flowmap_tiles(
    Flow f, const ReqDev* __restrict__ table, const int64_t* __restrict__ slot_off, int nreq,
    int64_t total_slots, double rtol, double atol, long long nmax) {
    f.prepare();
    const int64_t tid = xcd_block<Flow>() * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    const int64_t q = tid / G;
    if (q >= total_slots) return;  // whole groups leave together
    int lo = 0, hi = nreq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slot_off[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    const ReqDev& Rq = table[lo];
    const int64_t p = tile_particle(q - slot_off[lo], Rq.nx, Rq.ny);
    if (p < 0) return;
    const bool lead = (tid & (G - 1)) == 0;
    if (Rq.mask != nullptr && Rq.mask[p]) {
        if (lead) {
            Rq.out[p] = make_double2(0.0, 0.0);
            if (Rq.status) Rq.status[p] = 0;
            if (Rq.nsteps) Rq.nsteps[p] = 0;
        }
        return;
    }
    const int64_t i = p / Rq.ny;
    Dop853State<R> s;
    dop853_init(f, s, Rq.s0, Rq.s1, Rq.xs[i], Rq.ys[p - i * Rq.ny], rtol, atol);
    while (dop853_step(f, s, rtol, atol, nmax)) {
    }
    if (lead) {
        Rq.out[p] = make_double2(s.px(), s.py());
        if (Rq.status) Rq.status[p] = s.idid;
        if (Rq.nsteps) Rq.nsteps[p] = s.nstep;
    }
}

}  // namespace nbx
