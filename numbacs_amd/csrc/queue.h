// Persistent DOP853 launch with a device work queue.
//
// A grid sized to the machine (not to the problem) pulls particles from a
// global counter.  Every lane group (G lanes per particle) that has finished
// its solve -- or has none yet -- takes the next particle in the same pass of
// the loop, so the adaptive-step imbalance between particles (step counts
// differ by 2-6x on the reference workloads) no longer leaves lanes idle
// until the slowest particle of their wavefront finishes.  Fetches are
// aggregated per wavefront: one atomicAdd claims a run of indices for all
// idle groups of the wave.
//
// The queue spans a table of requests (flow-map grids with their own window
// t0, T and particle axes), which is how a whole FTLE time series runs in one
// launch.  Each solve is bounded by max_steps and the queue by its total, so
// every wavefront reaches the exit.
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

static __global__ void k_queue_reset(unsigned long long* counter) { *counter = 0ull; }

// Number of compute units of the current device (cached per device).
int device_cu_count();

template <class Flow, class R, int G>
__global__ __launch_bounds__(256, NBX_FLOWMAP_MIN_WAVES) void k_flowmap_queue(
    Flow f, ReqDev single, const ReqDev* __restrict__ table, const int64_t* __restrict__ offsets,
    int nreq, int64_t total, unsigned long long* __restrict__ counter, double rtol, double atol,
    long long nmax) {
    const int lane = static_cast<int>(__lane_id());
    const int glead = lane & ~(G - 1);
    const bool is_lead = (lane & (G - 1)) == 0;
    Dop853State<R> s;
    s.idid = 0;
    int64_t p = 0;
    int r = 0;
    bool active = false, exhausted = false;
    for (;;) {
        // ---- refill (wavefront-uniform control) ----
        const bool want = !active && !exhausted;
        const unsigned long long m = __ballot(want && is_lead);
        if (m != 0ull) {
            const int first = __ffsll(static_cast<long long>(m)) - 1;
            unsigned long long base = 0;
            if (lane == first) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(m)));
            base = __shfl(base, first);
            long long q = -1;
            if (want && is_lead)
                q = static_cast<long long>(base) +
                    __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
            q = __shfl(q, glead);
            if (want) {
                if (q >= 0 && q < total) {
                    if (table == nullptr) {
                        r = 0;
                        p = q;
                    } else {
                        int lo = 0, hi = nreq - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (offsets[mid] <= q) lo = mid;
                            else hi = mid - 1;
                        }
                        r = lo;
                        p = q - offsets[lo];
                    }
                    const ReqDev& Rq = table == nullptr ? single : table[r];
                    if (Rq.mask != nullptr && Rq.mask[p]) {
                        if (is_lead) {
                            Rq.out[p] = make_double2(0.0, 0.0);
                            if (Rq.status) Rq.status[p] = 0;
                            if (Rq.nsteps) Rq.nsteps[p] = 0;
                        }
                    } else {
                        const int64_t i = p / Rq.ny;
                        dop853_init(f, s, Rq.s0, Rq.s1, Rq.xs[i], Rq.ys[p - i * Rq.ny], rtol, atol);
                        active = true;
                    }
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(active || !exhausted) == 0ull) break;
        // ---- one attempted step for every group with a particle ----
        if (active) {
            if (!dop853_step(f, s, rtol, atol, nmax)) {
                if (is_lead) {
                    const ReqDev& Rq = table == nullptr ? single : table[r];
                    Rq.out[p] = make_double2(s.px(), s.py());
                    if (Rq.status) Rq.status[p] = s.idid;
                    if (Rq.nsteps) Rq.nsteps[p] = s.nstep;
                }
                active = false;
            }
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
    const int64_t tid = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
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
