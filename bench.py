"""FTLE benchmark on one MI355X: gridded-velocity FTLE + ridges, FP64.

Workload (one step): a seeded 64-mode synthetic velocity field on a
0.5-degree lon/lat grid of 271x101 points with 96 hourly frames, tricubic
spline (spherical=1, km/h), particles on a 0.2-degree grid (676x251),
t0 = 72 h, T = -72 h, DOP853 rtol=1e-6 atol=1e-8 in FP64, then the fused
Cauchy-Green eigen/FTLE stencil and FTLE ridge extraction at the 90th
percentile.  Inputs (spline coefficients, particle axes) are resident in HBM
before timing starts.

Prints one JSON line: grid points/s, the dominant kernel's roofline figures
(its launch duration measured with HIP events on the stream it runs on),
and the CPU oracle timed on this machine's host cores as a reported
baseline, plus the error of the GPU result against that FP64 oracle.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# algorithmic figures per unit of work
SPLINE_BYTES_PER_EVAL = 2 * 64 * 8         # 64 (u, v) coefficient pairs, FP64
EVALS_PER_STEP = 12                        # DOP853 stages per attempted step
BYTES_PER_PARTICLE_STEP = SPLINE_BYTES_PER_EVAL * EVALS_PER_STEP  # 12,288 B
FLOPS_RK_PER_STEP = 350                    # stage sums, solution, error norm, control
FLOPS_SPLINE_RHS = 400                     # weights + 2 x 64 FMA + spherical scaling
FLOPS_PER_PARTICLE_STEP = FLOPS_RK_PER_STEP + EVALS_PER_STEP * FLOPS_SPLINE_RHS


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def build_workload(lib):
    from numbacs_amd.flows import get_flow_2D, prefilter3_device
    from numbacs_amd.workloads import field_sha256, particle_axes, synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity(nt=96)
    grid = ((t[0], t[-1], len(t)), (lon[0], lon[-1], len(lon)), (lat[0], lat[-1], len(lat)))
    Cu = prefilter3_device(U)
    Cv = prefilter3_device(V)
    flow = get_flow_2D(grid, Cu, Cv, spherical=1)
    xp, yp = particle_axes(0.2)
    return dict(t=t, lon=lon, lat=lat, U=U, V=V, grid=grid, Cu=Cu, Cv=Cv, flow=flow, xp=xp, yp=yp,
                sha=field_sha256(U, V))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)

    import torch

    from numbacs_amd import _lib
    from numbacs_amd.diagnostics import ftle_eig_2d_device
    from numbacs_amd.extraction.ridges import ftle_ridges_device
    from numbacs_amd.integration import flowmap_grid_2d_device

    torch.cuda.set_device(0)
    W = build_workload(_lib)
    xp, yp = W["xp"], W["yp"]
    nx, ny = len(xp), len(yp)
    xd, yd = _lib.to_dev(xp), _lib.to_dev(yp)
    params = np.array([-1.0])
    t0, T = 72.0, -72.0
    dx, dy = float(xp[1] - xp[0]), float(yp[1] - yp[0])
    fm = _lib.empty((nx, ny, 2))
    nsteps = _lib.empty((nx, ny), "i32")
    status = _lib.empty((nx, ny), "i32")
    ftle = _lib.empty((nx, ny))
    evec = _lib.empty((nx, ny, 2))

    ev_fm = []

    def step(record=False):
        if record:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        flowmap_grid_2d_device(W["flow"], t0, T, xd, yd, params, 1e-6, 1e-8, out=fm,
                               status=status, nsteps=nsteps)
        if record:
            e1.record()
            ev_fm.append((e0, e1))
        ftle_eig_2d_device(fm, T, dx, dy, ftle=ftle, evec=evec)
        return ftle_ridges_device(ftle, evec, xd, yd, nx, ny, dx, dy, 0.0, 90, 3)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        res = step(record=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_step = 1e3 * wall / args.steps
    fm_ms = float(np.mean([a.elapsed_time(b) for a, b in ev_fm]))

    P = nx * ny
    steps_total = int(nsteps.sum().item())
    st = _lib.to_host(status)
    pts_per_s = P / (wall / args.steps)
    bytes_launch = BYTES_PER_PARTICLE_STEP * steps_total
    flops_launch = FLOPS_PER_PARTICLE_STEP * steps_total
    achieved_gbs = bytes_launch / (fm_ms * 1e-3) / 1e9
    achieved_tflops = flops_launch / (fm_ms * 1e-3) / 1e12

    fm_h = _lib.to_host(fm)
    ftle_h = _lib.to_host(ftle)
    lon, lat = W["lon"], W["lat"]
    outside = ((fm_h[..., 0] < lon[0]) | (fm_h[..., 0] > lon[-1]) | (fm_h[..., 1] < lat[0]) |
               (fm_h[..., 1] > lat[-1]))
    pts_out, offsets, nridges, nridge_pts = res

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "flowmap_spline_gridded_676x251_hbm.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    rec = {
        "metric": "FTLE grid points/s (gridded-velocity FTLE + ridges, plain FP64 DOP853)",
        "value": pts_per_s,
        "unit": "grid points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "gridded-velocity FTLE: 64-mode synthetic field 271x101x96 (0.5 deg, "
                        "hourly), tricubic spline spherical=1, particles 676x251 (0.2 deg), "
                        "t0=72h T=-72h, DOP853 rtol=1e-6 atol=1e-8, FTLE + ridges at p90",
            "particles": P,
            "field_sha256": W["sha"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_flowmap_grid<Spline3<double>>",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel_ms": fm_ms,
            "algorithmic_bytes_per_launch": bytes_launch,
            "fp64_tflops": achieved_tflops,
            "fp64_frac": achieved_tflops / FP64_PEAK_TFLOPS,
        },
        "particle_steps_per_s": steps_total / (fm_ms * 1e-3),
        "mean_steps_per_particle": steps_total / P,
        "status_ok": int((st == 1).sum()),
        "final_outside_velocity_grid": int(outside.sum()),
        "ridges": int(nridges),
        "ridge_points": int(nridge_pts),
    }

    if not args.no_cpu:
        rec["cpu_baseline"], rec["error_vs_fp64_oracle"] = cpu_baseline(W, fm_h, ftle_h, t0, T,
                                                                         args.cpu_reps)
    print(json.dumps(rec))


def cpu_baseline(W, fm_gpu, ftle_gpu, t0, T, reps):
    """The CPU oracle (C restatement of the reference algorithm, OpenMP over
    grid rows as the reference's prange) on the full workload, on all host
    cores this process may use; median of reps runs after one warm-up."""
    import oracle

    cores = len(os.sched_getaffinity(0))
    S = oracle.Spline(W["grid"], _host(W["Cu"]), _host(W["Cv"]), spherical=1)
    xp, yp = W["xp"], W["yp"]
    dx, dy = xp[1] - xp[0], yp[1] - yp[0]
    times = []
    for r in range(reps + 1):
        t1 = time.perf_counter()
        fm, st, _ = oracle.flowmap_grid_2d("spline", np.array([-1.0]), t0, T, xp, yp, spline=S,
                                           nthreads=cores)
        ev, V = oracle.C_eig_2d(fm, dx, dy)
        ftle = np.where(ev[..., 1] > 1, np.log(np.maximum(ev[..., 1], 1e-300)) / (2 * abs(T)), 0)
        oracle.ridges(ftle, V[..., 1], xp, yp, percentile_=90)
        times.append(time.perf_counter() - t1)
    med = float(np.median(times[1:]))
    P = len(xp) * len(yp)
    scale = np.abs(fm).max()
    ferr = np.abs(fm_gpu - fm).max(-1) / scale
    fscale = np.abs(ftle).max()
    ftle_err = np.abs(ftle_gpu - ftle) / fscale
    cpu = {"value": P / med, "unit": "grid points/s", "cores": cores, "kind": "port",
           "sample": f"full workload ({P} particles, flow map + FTLE + ridges), median of {reps}",
           "seconds": med}
    err = {"flowmap_max_rel": float(ferr.max()),
           "flowmap_frac_within_1e-6": float(np.mean(ferr <= 1e-6)),
           "ftle_max_rel": float(ftle_err.max()),
           "ftle_frac_within_1e-6": float(np.mean(ftle_err <= 1e-6))}
    return cpu, err


def _host(t):
    return t.detach().cpu().numpy()


if __name__ == "__main__":
    main()
