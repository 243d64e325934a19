// Batched FTLE time series: many flow-map requests (each its own window
// t0, T and particle grid) over one flow, integrated by one launch -- the
// persistent queue of particle tiles, or one non-persistent launch over
// every request's tiles (the launch configuration's persistent occupancy 0;
// the autotuner picks the faster) -- followed by one batched FTLE launch over
// every request's grid.  The pair can be captured once into a HIP graph and
// replayed.  Requests run longest window first (the tile launch dispatches
// blocks in order, so the longest solves do not form the launch's tail).
//
// The reference's analogue is a Python loop over t0 calling flowmap_grid_2D
// and ftle_grid_2D per window (examples/time_series/plot_dg_time_series.py:
// 101-124) or a numba prange over t0 (lines 144-185).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "flow_handle.h"
#include "queue.h"
#include "stencil.h"

namespace nbx {
namespace {

struct FtleReqDev {
    const double2* fm;
    double* ftle;
    double2* evec;       // nullable: fused eigen path (FTLE from the eigen-solver)
    const uint8_t* mask; // nullable
    int nx, ny;
    double dx, dy, two_abs_T;
};

struct TileDev {
    int req, i0, j0;
};

__global__ __launch_bounds__(256) void k_ftle_batch(const FtleReqDev* __restrict__ reqs,
                                                    const TileDev* __restrict__ tiles) {
    const TileDev t = tiles[blockIdx.x];
    const FtleReqDev R = reqs[t.req];
    const int j = t.j0 + static_cast<int>(threadIdx.x);
    const int i = t.i0 + static_cast<int>(threadIdx.y);
    if (i >= R.nx || j >= R.ny) return;
    const long k = static_cast<long>(i) * R.ny + j;
    const bool in = i >= 1 && i < R.nx - 1 && j >= 1 && j < R.ny - 1 &&
                    !(R.mask != nullptr && R.mask[k]);
    if (R.evec == nullptr) {
        R.ftle[k] = in ? ftle_point(R.fm, k, R.ny, 1.0 / R.two_abs_T, R.dx, R.dy) : 0.0;
    } else {
        double ev[2] = {0.0, 0.0}, V[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        if (in) ceig_point(R.fm, k, R.ny, R.dx, R.dy, ev, V);
        R.ftle[k] = ev[1] > 1 ? log(ev[1]) / R.two_abs_T : 0.0;
        R.evec[k] = make_double2(V[0][1], V[1][1]);
    }
}

}  // namespace
}  // namespace nbx

struct nbx_batch {
    const nbx_flow* flow = nullptr;
    std::vector<double> params;
    double rtol = 1e-6, atol = 1e-8;
    long long nmax = 100000;
    int precision = NBX_FP64;
    int nreq = 0;
    int64_t total = 0;
    int ntiles = 0;
    nbx::ReqDev* d_table = nullptr;
    int64_t total_slots = 0;
    int64_t* d_slot_off = nullptr;  // first slot of each request (tile order)
    char* d_chunk = nullptr;        // queue launch: ring, counters, solver states
    size_t chunk_bytes = 0;
    nbx::FtleReqDev* d_ftle = nullptr;
    nbx::TileDev* d_tiles = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace nbx {
namespace {

template <class Flow, class R, int G>
int launch_batch_queue(nbx_batch* b, const Flow& f, const LaunchCfg& cfg, hipStream_t s) {
    const int block = cfg.block;
    if (cfg.persistent == 0) {
        if (b->total_slots == 0) return 0;
        k_flowmap_tiles<Flow, R, G><<<blocks_for(b->total_slots * G, block), block, 0, s>>>(
            f, b->d_table, b->d_slot_off, b->nreq, b->total_slots, b->rtol, b->atol, b->nmax);
        NBX_LAUNCHED("flowmap_tiles (batch)");
        return 0;
    }
    const int cus = device_cu_count();
    NBX_REQUIRE(cus > 0, "cannot query the compute-unit count");
    if (b->total_slots == 0) return 0;
    const int64_t ntiles = b->total_slots / (64 / G);
    const unsigned long long ring = chunk_ring(ntiles);
    const size_t state_off = ring * sizeof(unsigned long long) + 256;
    NBX_REQUIRE(b->d_chunk != nullptr &&
                    b->chunk_bytes >= state_off + sizeof(Dop853State<R>) * b->total_slots,
                "the batch's queue buffers are not allocated for this launch configuration");
    char* w = b->d_chunk;
    const ChunkQueue q{reinterpret_cast<unsigned long long*>(w),
                       reinterpret_cast<unsigned long long*>(w + state_off - 256),
                       reinterpret_cast<unsigned long long*>(w + state_off - 128),
                       reinterpret_cast<unsigned long long*>(w + state_off - 64), ring - 1,
                       static_cast<unsigned long long>(ntiles)};
    k_chunk_init<<<blocks_for(static_cast<int64_t>(ring), 256), 256, 0, s>>>(q, ntiles);
    NBX_LAUNCHED("chunk_init (batch)");
    const int64_t want = (ntiles * 64 + block - 1) / block;
    const int64_t grid =
        std::max<int64_t>(1, std::min<int64_t>(want, static_cast<int64_t>(cus) * cfg.persistent));
    k_flowmap_chunks<Flow, R, G><<<static_cast<unsigned>(grid), block, 0, s>>>(
        f, ReqDev{}, b->d_table, b->d_slot_off, b->nreq, q,
        reinterpret_cast<Dop853State<R>*>(w + state_off), b->rtol, b->atol, b->nmax);
    NBX_LAUNCHED("flowmap_chunks (batch)");
    return 0;
}

// Allocate (or grow) the queue launch's buffers for the current launch
// configuration: before a run, and before a capture (no allocation may
// happen inside one).  Sized for the FP64 state, the larger of the two.
int ensure_chunk_buffers(nbx_batch* b) {
    const LaunchCfg cfg = launch_cfg(b->flow->kind, b->precision);
    if (cfg.persistent <= 0 || b->total_slots == 0) return 0;
    const int G = (b->flow->kind == NBX_FLOW_SPLINE3 && b->flow->extrap != 1) ? cfg.lanes : 1;
    const int64_t ntiles = b->total_slots / (64 / G);
    NBX_REQUIRE(ntiles < (int64_t(1) << 31), "too many particle tiles for the queue");
    const size_t need = chunk_ring(ntiles) * sizeof(unsigned long long) + 256 +
                        sizeof(Dop853State<double>) * static_cast<size_t>(b->total_slots);
    if (b->chunk_bytes >= need) return 0;
    if (b->d_chunk) (void)hipFree(b->d_chunk);
    b->d_chunk = nullptr;
    b->chunk_bytes = 0;
    NBX_HIP(hipMalloc(&b->d_chunk, need));
    b->chunk_bytes = need;
    return 0;
}

template <class R>
int launch_batch_flow(nbx_batch* b, hipStream_t s) {
    const nbx_flow* fl = b->flow;
    const double* p = b->params.data();
    const LaunchCfg cfg = launch_cfg(fl->kind, b->precision);
    if (fl->kind == NBX_FLOW_SPLINE3) {
        if (fl->extrap == 1)  // the gradient path, one lane per particle
            return launch_batch_queue<Spline3<R, 1, true>, R, 1>(b, make_spline<R, true>(fl, p[0]),
                                                                 cfg, s);
        const Spline3<R, 1> sp = make_spline<R, false>(fl, p[0]);
        if (cfg.lanes == 4) {
            NBX_REQUIRE(sp.cbytes != 0, "the 4-lane spline kernel needs a coefficient array "
                                        "under 4 GiB (set 1 or 2 lanes per particle)");
            return launch_batch_queue<Spline3<R, 4>, R, 4>(b, regroup<4>(sp), cfg, s);
        }
        if (cfg.lanes == 2) return launch_batch_queue<Spline3<R, 2>, R, 2>(b, regroup<2>(sp), cfg, s);
        return launch_batch_queue<Spline3<R, 1>, R, 1>(b, sp, cfg, s);
    }
    if (fl->kind == NBX_FLOW_LINEAR3)
        return launch_batch_queue<Linear3<R>, R, 1>(b, make_linear<R>(fl, p[0]), cfg, s);
    if (fl->kind == NBX_FLOW_DOUBLE_GYRE) {
        const DoubleGyre<R> dg{p[0], p[1], p[2], p[3], p[4], p[5]};
        return launch_batch_queue<DoubleGyre<R>, R, 1>(b, dg, cfg, s);
    }
    BickleyJet<R> bj;
    for (int k = 0; k < 12; ++k) bj.p[k] = p[k];
    return launch_batch_queue<BickleyJet<R>, R, 1>(b, bj, cfg, s);
}

int enqueue_batch(nbx_batch* b, hipStream_t s) {
    int rc = b->precision == NBX_FP32_DISP ? launch_batch_flow<float>(b, s)
                                           : launch_batch_flow<double>(b, s);
    if (rc) return rc;
    if (b->ntiles > 0) {
        k_ftle_batch<<<b->ntiles, dim3(64, 4), 0, s>>>(b->d_ftle, b->d_tiles);
        NBX_LAUNCHED("ftle_batch");
    }
    return 0;
}

void free_batch(nbx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_table) (void)hipFree(b->d_table);
    if (b->d_slot_off) (void)hipFree(b->d_slot_off);
    if (b->d_chunk) (void)hipFree(b->d_chunk);
    if (b->d_ftle) (void)hipFree(b->d_ftle);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    delete b;
}

}  // namespace
}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_batch_create(const nbx_flow* flow, const nbx_request* reqs, int nreq,
                     const double* params, int nparams, double rtol, double atol,
                     int64_t max_steps, int precision, nbx_batch** out) {
    NBX_REQUIRE(out != nullptr, "out is NULL");
    NBX_REQUIRE(flow != nullptr, "flow handle is NULL");
    NBX_REQUIRE(reqs != nullptr && nreq >= 1, "a batch needs at least one request");
    NBX_REQUIRE(params != nullptr && nparams >= 1, "params must hold at least the direction");
    NBX_REQUIRE(rtol > 0 && atol >= 0, "tolerances must be positive");
    NBX_REQUIRE(max_steps >= 1, "max_steps must be >= 1");
    NBX_REQUIRE(precision == NBX_FP64 || precision == NBX_FP32_DISP, "unknown precision");
    NBX_REQUIRE(flow->kind != NBX_FLOW_DOUBLE_GYRE || nparams >= 6,
                "double_gyre needs 6 parameters");
    NBX_REQUIRE(flow->kind != NBX_FLOW_BICKLEY_JET || nparams >= 12,
                "bickley_jet needs 12 parameters");
    std::vector<ReqDev> table(nreq);
    std::vector<int64_t> offsets(nreq + 1, 0), slot_off(nreq + 1, 0);
    std::vector<FtleReqDev> ftle;
    std::vector<TileDev> tiles;
    // longest window first (stable, so equal windows keep the caller's order)
    std::vector<int> order(nreq);
    for (int r = 0; r < nreq; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return std::fabs(reqs[a].T) > std::fabs(reqs[c].T);
    });
    for (int r = 0; r < nreq; ++r) {
        const nbx_request& q = reqs[order[r]];
        NBX_REQUIRE(q.nx >= 0 && q.ny >= 0, "negative grid size in a request");
        const int64_t P = static_cast<int64_t>(q.nx) * q.ny;
        NBX_REQUIRE(P == 0 || (q.x && q.y && q.flowmap), "NULL array in a request");
        table[r] = ReqDev{params[0] * q.t0, params[0] * (q.t0 + q.T), q.x, q.y, q.nx,
                          q.ny > 0 ? q.ny : 1, q.mask, reinterpret_cast<double2*>(q.flowmap),
                          q.status, q.nsteps};
        offsets[r + 1] = offsets[r] + P;
        slot_off[r + 1] = slot_off[r] + (P > 0 ? tile_slots(q.nx, q.ny) : 0);
        if (q.ftle != nullptr && P > 0) {
            const int id = static_cast<int>(ftle.size());
            ftle.push_back(FtleReqDev{reinterpret_cast<const double2*>(q.flowmap), q.ftle,
                                      reinterpret_cast<double2*>(q.eigvec_max), q.mask, q.nx,
                                      q.ny, q.dx, q.dy, 2 * std::fabs(q.T)});
            for (int i0 = 0; i0 < q.nx; i0 += 4)
                for (int j0 = 0; j0 < q.ny; j0 += 64) tiles.push_back(TileDev{id, i0, j0});
        }
    }
    nbx_batch* b = new nbx_batch();
    b->flow = flow;
    b->params.assign(params, params + nparams);
    b->rtol = rtol;
    b->atol = atol;
    b->nmax = max_steps;
    b->precision = precision;
    b->nreq = nreq;
    b->total = offsets[nreq];
    b->total_slots = slot_off[nreq];
    b->ntiles = static_cast<int>(tiles.size());
    bool ok = hipMalloc(&b->d_table, sizeof(ReqDev) * nreq) == hipSuccess &&
              hipMalloc(&b->d_slot_off, sizeof(int64_t) * (nreq + 1)) == hipSuccess;
    if (ok && !ftle.empty())
        ok = hipMalloc(&b->d_ftle, sizeof(FtleReqDev) * ftle.size()) == hipSuccess &&
             hipMalloc(&b->d_tiles, sizeof(TileDev) * tiles.size()) == hipSuccess;
    if (ok)
        ok = hipMemcpy(b->d_table, table.data(), sizeof(ReqDev) * nreq, hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMemcpy(b->d_slot_off, slot_off.data(), sizeof(int64_t) * (nreq + 1),
                       hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !ftle.empty())
        ok = hipMemcpy(b->d_ftle, ftle.data(), sizeof(FtleReqDev) * ftle.size(),
                       hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(b->d_tiles, tiles.data(), sizeof(TileDev) * tiles.size(),
                       hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        free_batch(b);
        set_error("device allocation or upload of the batch tables failed");
        return -1;
    }
    *out = b;
    return 0;
}

int64_t nbx_batch_particles(const nbx_batch* b) { return b ? b->total : -1; }

int nbx_batch_run(nbx_batch* b, nbx_stream stream) {
    NBX_REQUIRE(b != nullptr, "batch is NULL");
    if (const int rc = ensure_chunk_buffers(b)) return rc;
    return enqueue_batch(b, as_stream(stream));
}

int nbx_batch_capture(nbx_batch* b, nbx_stream stream) {
    NBX_REQUIRE(b != nullptr, "batch is NULL");
    NBX_REQUIRE(stream != nullptr, "graph capture needs a non-default stream");
    hipStream_t s = as_stream(stream);
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    if (const int rc = ensure_chunk_buffers(b)) return rc;
    NBX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_batch(b, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return -1;
    }
    b->graph = g;
    NBX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return 0;
}

int nbx_batch_run_graph(nbx_batch* b, nbx_stream stream) {
    NBX_REQUIRE(b != nullptr && b->exec != nullptr, "batch has no captured graph");
    NBX_HIP(hipGraphLaunch(b->exec, as_stream(stream)));
    return 0;
}

void nbx_batch_destroy(nbx_batch* b) {
    if (b) free_batch(b);
}

}  // extern "C"
