/*
 * numbacs_amd.h -- C ABI of libnumbacs_amd.so, the MI355X (gfx950) FTLE hot
 * path: DOP853 particle flow maps, FTLE / Cauchy-Green eigen stencils and
 * FTLE ridge extraction.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer unless its name ends in _host.
 *   - Arrays use the reference's 'ij' C-order layouts: a flow map is
 *     (nx, ny, 2) float64, an FTLE field (nx, ny) float64, eigenvalues
 *     (nx, ny, 2), eigenvectors (nx, ny, 2, 2) with eigenvectors as columns.
 *   - Every launch is enqueued on the given stream (NULL = default stream);
 *     a call returns once its work is enqueued unless it says otherwise.
 *   - Return value 0 = success, < 0 = error; nbx_last_error() then returns a
 *     thread-local message.
 *   - The percentile, ridge, label and ordered-ridge entry points use a
 *     grow-only device workspace per (device, stream), held until the
 *     process ends and locked for the duration of a call: calls on
 *     different streams run concurrently, calls on one stream from several
 *     host threads take turns.  The flow-map, FTLE and Cauchy-Green entry
 *     points use only their arguments (and, for the persistent launch
 *     shape, a queue counter allocated per call).
 *
 * The reference's own FFI for this path is numbalsoda's
 *   dop853(funcptr, u0, t_eval, rtol, atol, data=params)
 * called once per particle with funcptr = address of a numba cfunc
 *   void rhs(double t, double* y, double* dy, double* p)
 * (integration.py:169-171, flows.py:1146).  Here the cfunc address is
 * replaced by an opaque nbx_flow handle and the per-particle loop by one
 * grid-wide launch.
 */
#ifndef NUMBACS_AMD_H
#define NUMBACS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nbx_flow nbx_flow;
typedef void* nbx_stream; /* a hipStream_t */

enum nbx_flow_kind {
    NBX_FLOW_DOUBLE_GYRE = 0,
    NBX_FLOW_BICKLEY_JET = 1,
    NBX_FLOW_SPLINE3 = 2,
    NBX_FLOW_LINEAR3 = 3
};
enum nbx_extrap { NBX_EXTRAP_CONSTANT = 0, NBX_EXTRAP_LINEAR = 1, NBX_EXTRAP_NEAREST = 2 };
enum nbx_precision { NBX_FP64 = 0, NBX_FP32_DISP = 1 };

/* Per-particle status written by the flow-map launches (Hairer's idid). */
enum nbx_status {
    NBX_ST_MASKED = 0,      /* masked point, not integrated (output 0)          */
    NBX_ST_OK = 1,          /* reached t0+T                                      */
    NBX_ST_MAX_STEPS = -2,  /* step budget exhausted                             */
    NBX_ST_H_SMALL = -3     /* step size underflow (0.1|h| <= |t| * 2.3e-16)     */
};

const char* nbx_last_error(void);
int nbx_version(void);

/* --- flows ------------------------------------------------------------- */

/* An analytic predefined flow (get_predefined_flow, flows.py:1104-1245).
 * Parameters are passed per launch (params[0] = integration direction). */
int nbx_flow_create_analytic(int kind, nbx_flow** out);

/* A gridded flow evaluated by a tricubic natural B-spline (get_flow_2D,
 * flows.py:121-258).  grid9_host = {t0,t1,nt, x0,x1,nx, y0,y1,ny} (the
 * reference's UCGrid tuple), Cu/Cv = DEVICE coefficient arrays of shape
 * (nt+2, nx+2, ny+2) float64 (get_interp_arrays_2D's output layout).
 * The handle keeps its own interleaved (u, v) copy in HBM.  spherical:
 * 0 planar, 1 lon in [-180,180), 2 lon in [0,360) (degrees, km/h velocities
 * converted with radius r).  extrap (nbx_extrap): 0 zero outside the grid
 * box, 1 the value plus gradient times offset at the clamped point, 2 the
 * value at the clamped point. */
int nbx_flow_create_spline3(const double* grid9_host, const double* Cu, const double* Cv,
                            int spherical, int extrap, double r, nbx_flow** out,
                            nbx_stream stream);

/* A gridded flow evaluated by trilinear interpolation of the samples
 * (get_flow_linear_2D, flows.py:418-506): U/V = DEVICE (nt, nx, ny) float64
 * velocity samples on the grid grid9_host; spherical and r as for
 * nbx_flow_create_spline3; extrap: 0 zero outside the grid box (the
 * reference dependency's CONSTANT), 1 continued boundary cell, 2 clamped. */
int nbx_flow_create_linear3(const double* grid9_host, const double* U, const double* V,
                            int spherical, int extrap, double r, nbx_flow** out,
                            nbx_stream stream);

/* Velocity ingest: (nt, ny, nx) 'xy' samples -> (nt, nx, ny) 'ij' times
 * scale (the np.moveaxis(u, 1, 2) * 3.6 of plot_merra_ftle.py:37-41). */
int nbx_ingest_xy(const double* U_xy, int nt, int ny, int nx, double scale, double* U_ij,
                  nbx_stream stream);
void nbx_flow_destroy(nbx_flow* flow);

/* Natural cubic B-spline prefilter (get_interp_arrays_2D, flows.py:39-46):
 * U (nt, nx, ny) -> C (nt+2, nx+2, ny+2). */
int nbx_prefilter3(const double* U, int nt, int nx, int ny, double* C, nbx_stream stream);

/* Velocity samples: out[k] = rhs(t[k], pts[k]) for k < n (pts (n,2)). */
int nbx_flow_eval(const nbx_flow* flow, const double* params_host, int nparams, const double* t,
                  const double* pts, int64_t n, double* out, nbx_stream stream);

/* --- flow maps --------------------------------------------------------- */

/* flowmap_grid_2D (integration.py:123-182) with method="dop853": one
 * adaptive DOP853 solve per (x[i], y[j]) from s0 = params[0]*t0 to
 * s1 = params[0]*(t0+T); out (nx, ny, 2).  mask (nx*ny bytes, nonzero =
 * skip, may be NULL) leaves 0 in out.  status / nsteps (nx*ny int32, may be
 * NULL) receive nbx_status and the attempted step count.  max_steps bounds
 * every solve (Hairer's NMAX).  precision selects FP64 or the
 * FP32-displacement form (FP32 stages, step displacements and spline;
 * FP64 position, time and step control). */
int nbx_flowmap_grid_2d(const nbx_flow* flow, double t0, double T, const double* x, int nx,
                        const double* y, int ny, const double* params_host, int nparams,
                        double rtol, double atol, const uint8_t* mask, int64_t max_steps,
                        int precision, double* out, int32_t* status, int32_t* nsteps,
                        nbx_stream stream);

/* flowmap (integration.py:7-61): pts (npts, 2) -> out (npts, 2). */
int nbx_flowmap_pts(const nbx_flow* flow, double t0, double T, const double* pts, int64_t npts,
                    const double* params_host, int nparams, double rtol, double atol,
                    const uint8_t* mask, int64_t max_steps, int precision, double* out,
                    int32_t* status, int32_t* nsteps, nbx_stream stream);

/* flowmap_aux_grid_2D (integration.py:249-464): n_aux = 5 particles per grid
 * point (eig_main != 0: offsets (h,0), (-h,0), (0,h), (0,-h), (0,0)) or the
 * first 4; out (nx, ny, n_aux, 2).  compute_edge != 0: edge points integrate
 * only the centre particle (eig_main) or all 4; compute_edge == 0: edge
 * points are left out.  Left-out and masked particles (mask (nx*ny)) get 0
 * in out and status 0; status / nsteps hold nx*ny*n_aux entries. */
int nbx_flowmap_aux_grid_2d(const nbx_flow* flow, double t0, double T, const double* x, int nx,
                            const double* y, int ny, double h, int eig_main, int compute_edge,
                            const double* params_host, int nparams, double rtol, double atol,
                            const uint8_t* mask, int64_t max_steps, int precision, double* out,
                            int32_t* status, int32_t* nsteps, nbx_stream stream);

/* flowmap_composition (integration.py:609-644): out (nx, ny, 2) = the chain
 * of nT flow maps on the uniform grid [x0,x1] x [y0,y1] (nx x ny), each
 * sampled bilinearly at the previous one's image, 0 outside the closed box
 * (eval_linear with CONSTANT extrapolation).  flowmaps holds nslots
 * (nx, ny, 2) slots; map k of the chain is slot (first + k) % nslots (the
 * reference's flowmaps[k] with first = 0, nslots = nT).  For nT = 1 the
 * single map is applied twice, as the reference's loop bounds do.  Used by
 * flowmap_composition_initial / _step (integration.py:647-737). */
int nbx_flowmap_compose(const double* flowmaps, int nslots, int first, int nT, double x0,
                        double x1, int nx, double y0, double y1, int ny, double* out,
                        nbx_stream stream);

/* Launch shape of the flow-map kernels for one (flow kind, precision):
 * lanes_per_particle (1, 2 or 4; spline flows only: the lanes of a group
 * split each tricubic gather; spline flows created with linear
 * extrapolation always run one lane per particle, in the kernel that
 * carries the gradient path; 4 lanes need a coefficient array under
 * 4 GiB), threads per block (64, 128 or 256), and persistent_blocks_per_cu
 * (0: one lane group per particle in one launch over the particle tiles;
 * n > 0: a persistent grid of n blocks per compute unit taking wavefront
 * tiles from a device queue, a tile going back to the queue every 16
 * steps while a particle of it is still running; -1: spline coefficients
 * staged through LDS per step, spline flows at 4 lanes and 64-thread
 * blocks; persistent and staged launches serve grid flow maps and batches,
 * other sources run the plain launch).  Process-wide; set by the autotuner
 * (numbacs_amd.autotune).  No reference counterpart (the reference's
 * parallel loop has no tunables, integration.py:166). */
int nbx_set_launch_config(int kind, int precision, int lanes_per_particle, int block_threads,
                          int persistent_blocks_per_cu);
int nbx_get_launch_config(int kind, int precision, int* lanes_per_particle, int* block_threads,
                          int* persistent_blocks_per_cu);

/* Launch order of the particle tiles in the spline flows' grid launch
 * (persistent_blocks_per_cu 0), per precision: edge_first 1 (the default)
 * launches the tiles nearest the velocity box's edge first, by the distance
 * of each tile's 32 x 32-particle block (whole blocks in tile order within a
 * distance, for L2 locality) -- particles leaving the box cross the jump to
 * zero velocity, take the most steps and would otherwise form the launch's
 * tail -- edge_first 0 keeps plain tile order.  Results do not depend on it.  Process-wide; the autotuner times
 * both.  No reference counterpart. */
int nbx_set_tile_order(int precision, int edge_first);
int nbx_get_tile_order(int precision, int* edge_first);

/* --- batched FTLE time series ------------------------------------------ */

/* One flow-map request of a batch: window [t0, t0+T], particle axes x (nx),
 * y (ny), outputs as for nbx_flowmap_grid_2d; with ftle != NULL the batch
 * also computes that request's FTLE field (grid spacing dx, dy), from the
 * closed-form eigenvalue (ftle_grid_2D) or, with eigvec_max != NULL, from
 * the eigen-solver together with eigvec_max (C_eig_2D + ftle_from_eig). */
typedef struct nbx_request {
    double t0, T;
    const double* x;
    int nx;
    const double* y;
    int ny;
    const uint8_t* mask;   /* nullable */
    double* flowmap;       /* (nx, ny, 2) */
    int32_t* status;       /* nullable */
    int32_t* nsteps;       /* nullable */
    double* ftle;          /* nullable, (nx, ny) */
    double* eigvec_max;    /* nullable, (nx, ny, 2) */
    double dx, dy;
} nbx_request;

typedef struct nbx_batch nbx_batch;

/* Build a batch over one flow (all requests share params, tolerances and
 * precision).  Uploads the request table; the output buffers must outlive
 * every run.  No reference counterpart: the reference loops over t0 in
 * Python (examples/time_series/plot_dg_time_series.py:101-185). */
int nbx_batch_create(const nbx_flow* flow, const nbx_request* reqs_host, int nreq,
                     const double* params_host, int nparams, double rtol, double atol,
                     int64_t max_steps, int precision, nbx_batch** out);
int64_t nbx_batch_particles(const nbx_batch* batch);
/* Enqueue the whole batch: one persistent work-queue flow-map launch over
 * every request, then one batched FTLE launch. */
int nbx_batch_run(nbx_batch* batch, nbx_stream stream);
/* Capture nbx_batch_run into a HIP graph (stream must not be the default
 * stream) and replay it with nbx_batch_run_graph. */
int nbx_batch_capture(nbx_batch* batch, nbx_stream stream);
int nbx_batch_run_graph(nbx_batch* batch, nbx_stream stream);
void nbx_batch_destroy(nbx_batch* batch);

/* --- FTLE and Cauchy-Green eigen --------------------------------------- */

/* ftle_grid_2D (diagnostics.py:21-65): interior central differences,
 * borders and masked points 0, ftle = log(lmax)/(2|T|) where lmax > 1. */
int nbx_ftle_grid_2d(const double* fm, int nx, int ny, double T, double dx, double dy,
                     const uint8_t* mask, double* ftle, nbx_stream stream);

/* C_eig_2D (diagnostics.py:200-244): eigvals (nx,ny,2) ascending, eigvecs
 * (nx,ny,2,2) with eigvecs[i,j,:,k] the k-th eigenvector; zeros on the
 * border and at masked points. */
int nbx_C_eig_2d(const double* fm, int nx, int ny, double dx, double dy, const uint8_t* mask,
                 double* eigvals, double* eigvecs, nbx_stream stream);

/* Fused ridge-path stencil: C_eig_2D + ftle_from_eig (diagnostics.py:247-269)
 * in one pass; writes ftle (nx,ny) and eigvec_max (nx,ny,2). */
int nbx_ftle_eig_2d(const double* fm, int nx, int ny, double T, double dx, double dy,
                    const uint8_t* mask, double* ftle, double* eigvec_max, nbx_stream stream);

/* C_tensor_2D (diagnostics.py:67-110): C (nx, ny, 3) = (C11, C12, C22) of the
 * aux-grid gradient (gradF_aux_stencil_2D, utils.py:50-84) on i, j in
 * [2, n-3]; zeros elsewhere and at masked points.  fm_aux (nx, ny, naux, 2),
 * naux 4 or 5. */
int nbx_C_tensor_aux_2d(const double* fm_aux, int nx, int ny, int naux, double h,
                        const uint8_t* mask, double* C, nbx_stream stream);

/* C_eig_aux_2D (diagnostics.py:113-197): eig_main != 0 (naux = 5): on i, j in
 * [2, n-3] eigenvectors of the aux-grid tensor and eigenvalues of the
 * main-grid tensor (gradF_main_stencil_2D, utils.py:88-124); eig_main == 0:
 * both from the aux-grid tensor on i, j in [1, n-2].  Layouts as
 * nbx_C_eig_2d. */
int nbx_C_eig_aux_2d(const double* fm_aux, int nx, int ny, int naux, double dx, double dy,
                     double h, int eig_main, const uint8_t* mask, double* eigvals,
                     double* eigvecs, nbx_stream stream);

/* --- ridges ------------------------------------------------------------ */

/* numpy.percentile(f, q) (linear method, ridges.py:51) of n values; the
 * result is left in DEVICE memory at *result (1 double). */
int nbx_percentile(const double* f, int64_t n, double q, double* result, nbx_stream stream);

/* ftle_ridge_pts (ridges.py:9-76).  dx = x[1]-x[0] and dy = y[1]-y[0] as in
 * the reference.  pts receives the ridge points (capacity nx*ny*2) in raveled
 * (i*ny+j) order; *npts_host their number.  Synchronises the stream. */
int nbx_ridge_pts(const double* f, const double* eigvec_max, const double* x, int nx,
                  const double* y, int ny, double dx, double dy, double sdd_thresh,
                  double percentile, double* pts, int64_t* npts_host, nbx_stream stream);

/* ftle_ridges (ridges.py:186-229): ridge points grouped into 8-connected
 * ridges (labels in raster order of each ridge's first point, points in
 * raveled order, ridges with fewer than min_ridge_pts points dropped).
 * pts (capacity nx*ny*2) holds the kept points ridge after ridge;
 * offsets (capacity nx*ny+1 int64) the start of each ridge plus the end.
 * *nridges_host and *npts_host receive the counts.  Synchronises the
 * stream. */
int nbx_ridges(const double* f, const double* eigvec_max, const double* x, int nx,
               const double* y, int ny, double dx, double dy, double sdd_thresh,
               double percentile, int min_ridge_pts, double* pts, int64_t* offsets,
               int64_t* nridges_host, int64_t* npts_host, nbx_stream stream);

/* ftle_ridges (ridges.py:186-229) without waiting on the stream: the same
 * grouping and outputs as nbx_ridges, but the counts go to DEVICE memory,
 * counts[0] = ridges kept, counts[1] = points kept, counts[2] = ridge points
 * before the min_ridge_pts filter (3 int64), valid once the stream's work
 * is done; nothing is read back, so a caller can queue the next field's
 * flow map behind it or capture it in a graph after a first call (which
 * sizes the stream's workspace).  Without a percentile threshold
 * (percentile = 0) the grouping is sized for the whole grid instead of the
 * ridge-point count.  counts[0] = -1 would report ridge points beyond the
 * grouping capacity (the capacity bounds them). */
int nbx_ridges_async(const double* f, const double* eigvec_max, const double* x, int nx,
                     const double* y, int ny, double dx, double dy, double sdd_thresh,
                     double percentile, int min_ridge_pts, double* pts, int64_t* offsets,
                     int64_t* counts, nbx_stream stream);

/* nbx_ridges without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and errors as nbx_ridges, but the pipeline's scratch is a fixed 64 MiB
 * device array of the library (one per GPU, held until the process ends)
 * instead of an allocated workspace, and the counts are read back without
 * the pinned staging buffer; nothing is allocated.  Grids whose scratch
 * exceeds 64 MiB (about 980,000 points) are rejected with an error.  All
 * calls on one GPU share that array: calls that overlap in time on one GPU
 * (two streams or two host threads) race. */
int nbx_ridges_noalloc(const double* f, const double* eigvec_max, const double* x, int nx,
                       const double* y, int ny, double dx, double dy, double sdd_thresh,
                       double percentile, int min_ridge_pts, double* pts, int64_t* offsets,
                       int64_t* nridges_host, int64_t* npts_host, nbx_stream stream);

/* 8-connected labelling of a byte grid (scipy.ndimage.label with a full
 * 3x3 structure, ridges.py:221-222): labels (nx*ny int32, 0 = background,
 * 1.. in raster order of first pixel); *nlabels_host the count.
 * Synchronises the stream. */
int nbx_label8(const uint8_t* b, int nx, int ny, int32_t* labels, int32_t* nlabels_host,
               nbx_stream stream);

/* ftle_ordered_ridges (ridges.py:720-1054): ridge points (the ftle_ridge_pts
 * predicate, percentile and sdd_thresh as there) linked into ordered curves
 * from the strongest second directional derivative (_linked_ridge_pts,
 * ridges.py:418-603), then curves with endpoints within dist_tol whose
 * tangents agree within ep_tan_ang joined; curves shorter than
 * min_ridge_pts dropped.  The predicate runs on the GPU (device f, evec,
 * x, y); the serial linking runs on the host.  Output to HOST buffers:
 * ridge k is points [offsets[k], offsets[k+1]) of pts (x/y pairs).  Fails
 * if pts_cap (points) or offsets_cap are too small, with *npts_host and
 * *nridges_host set to the sizes needed (nx*ny points always suffice). */
int nbx_ordered_ridges(const double* f, const double* eigvec_max, const double* x, int nx,
                       const double* y, int ny, double dist_tol, double ep_tan_ang,
                       int min_ridge_pts, double sdd_thresh, double percentile, double c,
                       double* pts_host, int64_t pts_cap, int64_t* offsets_host,
                       int64_t offsets_cap, int64_t* nridges_host, int64_t* npts_host,
                       nbx_stream stream);

/* The host half of nbx_ordered_ridges on caller-provided HOST buffers (no
 * GPU needed): pts_dense (nx*ny, 2) ridge points, eigvec_max (nx*ny, 2),
 * sdd (nx*ny) second directional derivative (< 0 exactly at ridge points,
 * which must lie 2+ points inside the grid, 0 elsewhere), h = min(dx, dy).
 * Outputs as nbx_ordered_ridges. */
int nbx_ordered_ridges_link(const double* pts_dense, const double* eigvec_max, const double* sdd,
                            int nx, int ny, double h, double dist_tol, double ep_tan_ang,
                            int min_ridge_pts, double sdd_thresh, double c, double* pts_host,
                            int64_t pts_cap, int64_t* offsets_host, int64_t offsets_cap,
                            int64_t* nridges_host, int64_t* npts_host);

/* --- flow-map surrogate (BF16 MFMA MLP) --------------------------------- */
/* No reference counterpart: kernels of the learned flow-map surrogate that
 * the FTLE path may use with integrator fallback.  bf16 tensors are raw
 * 16-bit patterns (uint16_t); GEMM sizes must be multiples of 64 (M, N) and
 * 32 (K).  mode 0: C16 = bf16(gelu(A B + bias)) and C32 = A B + bias (C32
 * may be NULL), A (M x K), B (K x N), both row-major; mode 1: C16 =
 * bf16(bf16(A B^T) * gelu'(pre)), B stored (N x K), pre (M x N) fp32;
 * mode 2: C32 = A^T B, A stored (K x M), B (K x N); mode 3: C32 = A B. */
int nbx_mlp_gemm(int mode, const uint16_t* A, const uint16_t* B, int M, int N, int K,
                 float* C32, uint16_t* C16, const float* bias, const float* pre,
                 nbx_stream stream);
/* h0 (n x 2*half) = bf16([sin z, cos z]), z = 2 pi ((inp - shift)/scale) Bf,
 * inp (n x 4) fp32, Bf (4 x half) fp32. */
int nbx_mlp_features(const float* inp, int64_t n, const float* shift, const float* scale,
                     const float* Bf, int half, uint16_t* h0, nbx_stream stream);
/* y (n x 2) = h (n x width) W (width x 2) + b, fp32 accumulation. */
int nbx_mlp_out_forward(const uint16_t* h, int64_t n, int width, const uint16_t* W,
                        const float* b, float* y, nbx_stream stream);
/* dy = (y - target)/n, *loss = 0.5 * mean |y - target|^2 (device scalar). */
int nbx_mlp_loss(const float* y, const float* target, int64_t n, float* dy, float* loss,
                 nbx_stream stream);
/* Output layer backward: gW (width x 2), gb (2) from h and dy; da (n x width)
 * = bf16(bf16(dy W^T) * gelu'(pre)). */
int nbx_mlp_out_backward(const float* dy, const uint16_t* h, const float* pre, int64_t n,
                         int width, const uint16_t* W, float* gW, float* gb, uint16_t* da,
                         nbx_stream stream);
/* gb (width) = column sums of da (n x width). */
int nbx_mlp_colsum(const uint16_t* da, int64_t n, int width, float* gb, nbx_stream stream);
/* Adam (bias-corrected, decoupled weight decay) on fp32 master weights w;
 * w16 (nullable) receives the bf16 copy. */
int nbx_mlp_adam(float* w, const float* g, float* m, float* v, uint16_t* w16, int64_t n,
                 float lr, float b1, float b2, float eps, int step, float weight_decay,
                 nbx_stream stream);
/* The same Adam step with the learning rate and the bias corrections in
 * device memory: sched = [lr, 1 - b1^t, 1 - b2^t] (3 floats), so a
 * captured training step replays for every step of a schedule. */
int nbx_mlp_adam_sched(float* w, const float* g, float* m, float* v, uint16_t* w16, int64_t n,
                       const float* sched, float b1, float b2, float eps, float weight_decay,
                       nbx_stream stream);
int nbx_mlp_to_bf16(const float* x, int64_t n, uint16_t* y, nbx_stream stream);

/* The trained network for fused inference: feature shift/scale (4 each) and
 * frequencies Bf (4 x half), hidden layers as transposed bf16 weights WT[l]
 * (width x K_l, k contiguous; K_0 = 2*half, else width) with fp32 biases,
 * the head Wout (width x 2, bf16) with bias bout and the output scaling
 * out_scale (2).  All device pointers. */
typedef struct nbx_mlp_net {
    const float* shift;
    const float* scale;
    const float* Bf;
    int half, width, layers;
    const uint16_t* WT[4];
    const float* b[4];
    const uint16_t* Wout;
    const float* bout;
    const float* out_scale;
} nbx_mlp_net;

/* out (n x 2) = displacement (physical units) for inputs inp (n x 4) =
 * (x0, y0, t0, T): features, hidden layers and head in one launch, equal to
 * the unfused forward pass (nbx_mlp_features, nbx_mlp_gemm mode 0,
 * nbx_mlp_out_forward) times out_scale.  Needs half = 64, width = 256,
 * layers = 4. */
int nbx_mlp_infer(const nbx_mlp_net* net, const float* inp, int64_t n, float* out,
                  nbx_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* NUMBACS_AMD_H */
