"""FTLE benchmark on one MI355X.

Default workload (the one the headline metric is quoted on): gridded-velocity
FTLE + ridges in FP64.  A seeded 64-mode synthetic velocity field on a
0.5-degree lon/lat grid of 271x101 points with 96 hourly frames, tricubic
spline (spherical=1, km/h), particles on a 0.2-degree grid (676x251),
t0 = 72 h, T = -72 h, DOP853 rtol=1e-6 atol=1e-8 in FP64, then the fused
Cauchy-Green eigen/FTLE stencil and FTLE ridge extraction at the 90th
percentile.  One step = flow map + FTLE + ridges with inputs resident in HBM.

The flow-map launch shape is chosen first by the autotuner (timed on this
workload, outside the timed region).  The JSON line carries the dominant
kernel's roofline figures (its duration measured with HIP events on the
stream it runs on) and the CPU oracle timed on this machine's host cores,
with the error of the GPU result against that FP64 oracle.

Other workloads (--workload): double_gyre (401x201, T=-10), timeseries
(48 uneven requests over a 480-frame field in one batched launch, FP64 and
FP32-displacement), precision (double gyre 2001x1001 and Bickley 2001x801 in
FP32-displacement vs FP64), surrogate (MLP surrogate + DOP853 fallback).
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
BF16_PEAK_TFLOPS = 2500.0
# algorithmic figures per unit of work
SPLINE_BYTES_PER_EVAL = 2 * 64 * 8          # 64 (u, v) coefficient pairs, FP64
EVALS_PER_STEP = 12                         # DOP853 stages per attempted step
BYTES_PER_PARTICLE_STEP = SPLINE_BYTES_PER_EVAL * EVALS_PER_STEP  # 12,288 B
FLOPS_RK_PER_STEP = 350                     # stage sums, solution, error norm, control
FLOPS_SPLINE_RHS = 400                      # weights + 2 x 64 FMA + spherical scaling
FLOPS_PER_PARTICLE_STEP = FLOPS_RK_PER_STEP + EVALS_PER_STEP * FLOPS_SPLINE_RHS
FLOPS_DG_RHS = 35                           # 30 flops + 5 transcendentals
MLP_FLOPS_PER_SAMPLE = 2 * (128 * 256 + 3 * 256 * 256 + 256 * 2)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-hbm", action="store_true",
                    help="gridded workload: take the flow map's HBM bytes from the committed "
                         "rocprofv3 record instead of measuring them (child runs under "
                         "rocprofv3 --pmc)")
    ap.add_argument("--tile-order", choices=["edge", "plain"], default=None,
                    help="gridded workload: fixed tile order of the FP64 grid launch")
    ap.add_argument("--no-timeseries", action="store_true",
                    help="default workload: leave out the time-series legs")
    ap.add_argument("--launch", default=None, metavar="LANES,BLOCK,PERSIST",
                    help="gridded workload: fixed flow-map launch shape instead of the "
                         "autotuner (e.g. 4,128,0)")
    ap.add_argument("--workload", default="gridded",
                    choices=["gridded", "double_gyre", "timeseries", "precision", "surrogate",
                             "composition", "next", "copycal", "stencil"])
    ap.add_argument("--surrogate-samples", type=int, default=20_000_000)
    ap.add_argument("--surrogate-t0s", type=int, default=96)
    return ap.parse_args()


def timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def gridded_field(nt=96):
    from numbacs_amd.flows import get_flow_2D, prefilter3_device
    from numbacs_amd.workloads import field_sha256, synthetic_velocity

    t, lon, lat, U, V = synthetic_velocity(nt=nt)
    grid = ((t[0], t[-1], len(t)), (lon[0], lon[-1], len(lon)), (lat[0], lat[-1], len(lat)))
    Cu = prefilter3_device(U)
    Cv = prefilter3_device(V)
    flow = get_flow_2D(grid, Cu, Cv, spherical=1)
    return dict(t=t, lon=lon, lat=lat, grid=grid, Cu=Cu, Cv=Cv, flow=flow, sha=field_sha256(U, V))


# ---------------------------------------------------------------------------
def bench_gridded(args):
    import torch

    from numbacs_amd import _lib, autotune
    from numbacs_amd.diagnostics import ftle_eig_2d_device
    from numbacs_amd.extraction.ridges import ftle_ridges_device, ftle_ridges_device_async
    from numbacs_amd.integration import flowmap_grid_2d_device
    from numbacs_amd.workloads import particle_axes

    W = gridded_field(96)
    xp, yp = particle_axes(0.2)
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

    def flowmap():
        flowmap_grid_2d_device(W["flow"], t0, T, xd, yd, params, 1e-6, 1e-8, out=fm,
                               status=status, nsteps=nsteps)

    if args.tile_order:
        autotune.set_order(_lib.NBX_FP64, args.tile_order == "edge")
    if args.launch:
        autotune.set_config(_lib.NBX_FLOW_SPLINE3, _lib.NBX_FP64,
                            tuple(int(v) for v in args.launch.split(",")))
    if args.launch or args.no_tune:
        tuned, tune_ms = autotune.get(_lib.NBX_FLOW_SPLINE3, _lib.NBX_FP64), None
    else:
        tuned, tune_ms = autotune.tune(_lib.NBX_FLOW_SPLINE3, _lib.NBX_FP64, flowmap, nx * ny,
                                       use_cache=False)

    ev_fm = []
    # Each step's ridges run on a second stream (default priority) behind an
    # event, so they overlap the next step's flow map and no step waits on
    # the host for its counts; the main stream waits for them only before
    # its FTLE kernel rewrites the fields they read.
    main, side = torch.cuda.current_stream(), torch.cuda.Stream()
    pts = _lib.empty((nx * ny, 2))
    offsets = _lib.empty((nx * ny + 1,), "i64")
    counts = _lib.empty((3,), "i64")
    ridges_done = [None]

    def step():
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        flowmap()
        e1.record()
        ev_fm.append((e0, e1))
        if ridges_done[0] is not None:
            main.wait_event(ridges_done[0])
        ftle_eig_2d_device(fm, T, dx, dy, ftle=ftle, evec=evec)
        ready = main.record_event()
        with torch.cuda.stream(side):
            side.wait_event(ready)
            ftle_ridges_device_async(ftle, evec, xd, yd, nx, ny, dx, dy, 0.0, 90, 3, pts,
                                     offsets, counts)
            ridges_done[0] = side.record_event()

    def serial_step():
        flowmap()
        ftle_eig_2d_device(fm, T, dx, dy, ftle=ftle, evec=evec)
        return ftle_ridges_device(ftle, evec, xd, yd, nx, ny, dx, dy, 0.0, 90, 3)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev_fm.clear()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_step = 1e3 * wall / args.steps
    fm_ms = float(np.mean([a.elapsed_time(b) for a, b in ev_fm]))
    nridges, nridge_pts = (int(v) for v in _lib.to_host(counts)[:2])
    # the same step with the ridges on the main stream and their counts read
    # back each step (the host waits): reported beside the pipelined rate
    serial_ms, res = timed(serial_step, args.steps, 1)
    assert (res[2], res[3]) == (nridges, nridge_pts), "pipelined and serial ridges differ"

    P = nx * ny
    steps_total = int(nsteps.sum().item())
    st = _lib.to_host(status)
    fm_h = _lib.to_host(fm)
    ftle_h = _lib.to_host(ftle)
    lon, lat = W["lon"], W["lat"]
    outside = ((fm_h[..., 0] < lon[0]) | (fm_h[..., 0] > lon[-1]) | (fm_h[..., 1] < lat[0]) |
               (fm_h[..., 1] > lat[-1]))
    # Velocity outside the grid box is zero ("constant" extrapolation) and
    # needs no gathers, so only the steps of particles that end inside the
    # box are counted: a lower bound on the gathered bytes and flops.
    steps_inside = int(_lib.to_host(nsteps)[~outside].sum())
    bytes_launch = BYTES_PER_PARTICLE_STEP * steps_inside
    flops_launch = FLOPS_PER_PARTICLE_STEP * steps_inside
    achieved_gbs = bytes_launch / (fm_ms * 1e-3) / 1e9
    achieved_tflops = flops_launch / (fm_ms * 1e-3) / 1e12

    order_edge = bool(autotune.get_order(_lib.NBX_FP64))
    hbm = None
    # (not when this run is itself under rocprofv3, whose launcher sets
    # ROCPROF_* variables for its target: no profiler inside a profiler)
    under_profiler = any(k.startswith("ROCPROF_") for k in os.environ)
    if not args.no_hbm and not under_profiler:
        print("bench: HBM counters (rocprofv3 child runs)", file=sys.stderr, flush=True)
        hbm = hbm_traffic_live(fm_ms, tuned, order_edge)
    if hbm is None:
        hbm = hbm_traffic(fm_ms)
    rec = {
        "metric": "FTLE grid points/s (gridded-velocity FTLE + ridges, plain FP64 DOP853)",
        "value": P / (wall / args.steps),
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
            "pipeline": "each step's ridges on a second stream, overlapping the next step's "
                        "flow map; all steps' work inside the timed region",
            "launch": {"lanes_per_particle": tuned[0], "block": tuned[1],
                       "persistent_blocks_per_cu": tuned[2],
                       "tile_order": ("edge-first" if autotune.get_order(_lib.NBX_FP64)
                                      else "plain") if tuned[2] == 0 else None},
        },
        "roofline": {
            "bound": "fp64",
            "kernel": "k_flowmap<Spline3<double,4>> (DOP853 over the tricubic spline, 4 lanes "
                      "per particle)",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": hbm["bytes_per_launch"] if hbm else None,
            "kernel_ms": fm_ms,
            "algorithmic_flops_per_launch": flops_launch,
            "algorithmic_unit": f"{FLOPS_PER_PARTICLE_STEP} FP64 flop per attempted step "
                                "(350 solver + 12 x 400 per spline evaluation) of the particles "
                                "that end inside the velocity box (a lower bound)",
            "hbm": hbm,
            "gather": {"bytes_per_launch": bytes_launch,
                       "achieved_GBs": achieved_gbs,
                       "unit": "12,288 B of spline coefficients per attempted step, served by "
                               "the lanes' stencil registers, L1 and L2 (not an HBM figure)"},
        },
        "particle_steps_per_s": steps_total / (fm_ms * 1e-3),
        "mean_steps_per_particle": steps_total / P,
        "status_ok": int((st == 1).sum()),
        "final_outside_velocity_grid": int(outside.sum()),
        "serial_ms_per_step": 1e3 * serial_ms,
        "ridges": int(nridges),
        "ridge_points": int(nridge_pts),
        "autotune_ms": tune_ms,
    }
    if not args.no_cpu:
        rec["cpu_baseline"], rec["error_vs_fp64_oracle"] = cpu_baseline(
            W, xp, yp, fm_h, ftle_h, _lib.to_host(evec), t0, T, args.cpu_reps)
    if not args.no_timeseries:
        print("bench: time-series legs", file=sys.stderr, flush=True)
        rec["timeseries"] = timeseries_legs(tune=not args.no_tune)
    return rec


def _pmc_mean(counter, cmd, match, timeout=120):
    """Mean per-dispatch value of one rocprofv3 counter over the dispatches
    whose kernel name contains every string in `match`, from a child run of
    `cmd` under `rocprofv3 --pmc`: (mean, dispatches), or None when
    rocprofv3 is absent or the run fails."""
    import glob
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    d = tempfile.mkdtemp(prefix="nbx_pmc_", dir="/tmp")
    try:
        r = subprocess.run([exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run",
                            "--"] + cmd, cwd=ROOT, env=dict(os.environ, TMPDIR="/tmp"),
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=timeout)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            return None
        return pmc_dispatch_mean(files[0], match)
    except (subprocess.TimeoutExpired, OSError, ValueError, KeyError):
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def pmc_dispatch_mean(path, match):
    """(mean per dispatch, dispatches) of a rocprofv3 counter-collection CSV
    over the dispatches whose kernel name contains every string in `match`
    (a counter's rows of one dispatch are summed), or None if none match."""
    import csv

    per = {}
    with open(path) as fh:
        for row in csv.DictReader(fh):
            if all(m in row["Kernel_Name"] for m in match):
                key = row["Dispatch_Id"]
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    return (sum(per.values()) / len(per), len(per)) if per else None


def hbm_traffic_live(kernel_ms, launch, edge_first):
    """HBM bytes per launch of the gridded spline flow map, measured now: the
    flow map's FETCH_SIZE and WRITE_SIZE (KiB per dispatch) from two child
    runs of this bench under rocprofv3 --pmc at the same launch shape and
    tile order, each scaled by what the counter reads for a 512 MiB device
    copy (two more child runs).  None if any run fails."""
    py = [sys.executable, "-u", os.path.join(ROOT, "bench.py")]
    fm_cmd = py + ["--steps", "2", "--warmup", "1", "--no-cpu", "--no-timeseries", "--no-hbm",
                   "--launch", ",".join(str(int(v)) for v in launch),
                   "--tile-order", "edge" if edge_first else "plain"]
    cal_cmd = py + ["--workload", "copycal"]
    match = ("k_flowmap", "Spline3<double")
    got = {}
    for key, counter, cmd, m in (("fetch", "FETCH_SIZE", fm_cmd, match),
                                 ("write", "WRITE_SIZE", fm_cmd, match),
                                 ("cal_fetch", "FETCH_SIZE", cal_cmd, ("copyBuffer",)),
                                 ("cal_write", "WRITE_SIZE", cal_cmd, ("copyBuffer",))):
        print(f"bench: rocprofv3 --pmc {counter} ({'copy calibration' if 'cal' in key else 'flow map'})",
              file=sys.stderr, flush=True)
        got[key] = _pmc_mean(counter, cmd, m)
        if got[key] is None or got[key][0] <= 0:
            return None
    copy_bytes = 64 * 1024 * 1024 * 8
    f_fetch = copy_bytes / (got["cal_fetch"][0] * 1024)
    f_write = copy_bytes / (got["cal_write"][0] * 1024)
    rd = got["fetch"][0] * 1024 * f_fetch
    wr = got["write"][0] * 1024 * f_write
    b = rd + wr
    gbs = b / (kernel_ms * 1e-3) / 1e9
    return {"bytes_per_launch": b, "read_bytes": rd, "write_bytes": wr, "achieved_GBs": gbs,
            "peak_GBs": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS,
            "fetch_size_kib_per_dispatch": got["fetch"][0],
            "write_size_kib_per_dispatch": got["write"][0],
            "dispatches": [got["fetch"][1], got["write"][1]],
            "calibration": {"copy_bytes": copy_bytes, "fetch_factor": f_fetch,
                            "write_factor": f_write},
            "source": "measured by this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE child "
                      "runs of bench.py at the same launch and tile order (bench.py " +
                      " ".join(fm_cmd[3:]) + "), scaled by a 512 MiB device copy "
                      "(bench.py --workload copycal)"}


def hbm_traffic(kernel_ms):
    """HBM bytes per launch of the spline flow map from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE record (profiles/), scaled by the
    device-copy calibration stored with it; None if the record is absent."""
    tpath = os.path.join(ROOT, "profiles", "flowmap_spline_gridded_676x251_hbm.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as fh:
        rec = json.load(fh)
    b = rec.get("hbm_bytes_per_launch")
    if b is None:
        return None
    gbs = b / (kernel_ms * 1e-3) / 1e9
    return {"bytes_per_launch": b, "achieved_GBs": gbs, "peak_GBs": HBM_PEAK_GBS,
            "frac": gbs / HBM_PEAK_GBS,
            "source": "profiles/flowmap_spline_gridded_676x251_hbm.json (rocprofv3 --pmc "
                      "FETCH_SIZE / WRITE_SIZE of this kernel, not measured in this run)",
            "measured_at_commit": rec.get("measured_at_commit")}


def timeseries_legs(reps=3, tune=True, oracle_rows=16):
    """The FTLE time series (48 uneven requests over the 480-frame field) in
    three forms: the plain FP64 path (one flowmap_grid_2D + ftle_grid_2D per
    request, as the reference's t0 loop), the batched FP64 form and the
    batched FP32-displacement form (one persistent work-queue launch plus one
    batched FTLE launch, HIP graph replay).  FTLE of every request is compared
    with the FP64 oracle on every oracle_rows-th particle row (the oracle
    integrates each sampled row and its two neighbours)."""
    import torch

    import oracle
    from numbacs_amd import _lib, autotune
    from numbacs_amd.batch import FTLEBatch
    from numbacs_amd.diagnostics import ftle_grid_2d_device
    from numbacs_amd.integration import flowmap_grid_2d_device
    from numbacs_amd.workloads import time_series_requests

    W = gridded_field(480)
    reqs = time_series_requests()
    P = sum(len(q["x"]) * len(q["y"]) for q in reqs)
    out = {"workload": "FTLE time series: 48 requests (t0 = 72..119 h hourly, T cycling "
                       "-24/-48/-72 h, seeded grids 201x101..1351x501) over the 64-mode field "
                       "extended to 271x101x480, spherical=1, DOP853 rtol=1e-6 atol=1e-8",
           "particles": P}
    # plain FP64: one single-grid launch + FTLE per request
    plain = []
    for q in reqs:
        nx, ny = len(q["x"]), len(q["y"])
        plain.append(dict(q, xd=_lib.to_dev(q["x"]), yd=_lib.to_dev(q["y"]),
                          fm=_lib.empty((nx, ny, 2)), ftle=_lib.empty((nx, ny)),
                          dx=float(q["x"][1] - q["x"][0]), dy=float(q["y"][1] - q["y"][0])))
    single_cfg = autotune.get(_lib.NBX_FLOW_SPLINE3, _lib.NBX_FP64)

    def run_plain():
        for r in plain:
            flowmap_grid_2d_device(W["flow"], r["t0"], r["T"], r["xd"], r["yd"], [-1.0],
                                   out=r["fm"])
            ftle_grid_2d_device(r["fm"], r["T"], r["dx"], r["dy"], out=r["ftle"])

    sec_plain, _ = timed(run_plain, reps, 1)
    out["plain_fp64"] = {"grid_points_per_s": P / sec_plain, "seconds": sec_plain,
                         "launch": list(single_cfg),
                         "form": "per request: flowmap_grid_2D + ftle_grid_2D"}
    results = {"plain_fp64": [_lib.to_host(r["ftle"]) for r in plain]}
    for prec in ("fp64", "fp32"):
        b = FTLEBatch(W["flow"], reqs, [-1.0], precision=prec)
        pc = _lib.NBX_FP64 if prec == "fp64" else _lib.NBX_FP32_DISP
        if tune:
            cands = [(4, blk, per) for blk in (64, 128, 256) for per in (0, 2, 4, 8)]
            cfg, _ = autotune.tune(_lib.NBX_FLOW_SPLINE3, pc, b.run, P, reps=1,
                                   use_cache=False, cands=cands, order=False)
        else:
            cfg = autotune.get(_lib.NBX_FLOW_SPLINE3, pc)
        b.capture()
        sec, _ = timed(b.run_graph, reps, 1)
        results["batched_" + prec] = [r["ftle"] for r in b.results()]
        out["batched_" + prec] = {"grid_points_per_s": P / sec, "seconds": sec,
                                  "launch": list(cfg), "particle_steps": b.total_steps()}
        b.close()
    autotune.set_config(_lib.NBX_FLOW_SPLINE3, _lib.NBX_FP64, single_cfg)
    out["speedup_batched_fp64_vs_plain"] = sec_plain / out["batched_fp64"]["seconds"]
    out["speedup_batched_fp32_vs_plain"] = sec_plain / out["batched_fp32"]["seconds"]
    # FTLE against the FP64 oracle, every request, sampled rows
    S = oracle.Spline(W["grid"], _host(W["Cu"]), _host(W["Cv"]), spherical=1)
    cores = len(os.sched_getaffinity(0))
    per_req = {k: [] for k in results}
    for k, q in enumerate(reqs):
        nx = len(q["x"])
        centre = np.arange(oracle_rows // 2, nx - 1, oracle_rows)
        sel = np.stack([centre - 1, centre, centre + 1], 1).ravel()
        fm_o, _, _ = oracle.flowmap_grid_2d("spline", np.array([-1.0]), q["t0"], q["T"],
                                            q["x"][sel], q["y"], spline=S, nthreads=cores)
        dx, dy = q["x"][1] - q["x"][0], q["y"][1] - q["y"][0]
        f_o = oracle.ftle_grid_2d(fm_o, q["T"], dx, dy)[1::3]
        scale = np.abs(f_o).max()
        for name in results:
            e = np.abs(results[name][k][centre] - f_o) / scale
            per_req[name].append((float(e.max()), float(np.sqrt(np.mean(e ** 2))),
                                  float(np.mean(e <= 1e-3)), float(np.mean(e <= 1e-6))))
    err = {}
    for name, v in per_req.items():
        v = np.array(v)
        err[name] = {"max_rel": float(v[:, 0].max()),
                     "median_request_max_rel": float(np.median(v[:, 0])),
                     "median_request_rms_rel": float(np.median(v[:, 1])),
                     "frac_within_1e-3": float(v[:, 2].mean()),
                     "frac_within_1e-6": float(v[:, 3].mean()),
                     "requests_all_within_1e-3": int((v[:, 2] == 1.0).sum()),
                     "per_request_max_rel": [round(float(a), 9) for a in v[:, 0]]}
    out["ftle_vs_fp64_oracle"] = dict(
        err, sample=f"every {oracle_rows}th particle row of each of the 48 requests (the oracle "
                    "integrates the row and its two neighbours); error relative to the "
                    "request's largest sampled oracle FTLE")
    return out


def cpu_baseline(W, xp, yp, fm_gpu, ftle_gpu, evec_gpu, t0, T, reps):
    """The CPU oracle (C restatement of the reference algorithm, OpenMP over
    grid rows as the reference's prange) on the full workload, on all host
    cores this process may use; median of reps runs after one warm-up.  Also
    the GPU result's error against it: flow map, FTLE and the p90 ridge sets
    (grid points flagged as ridge points, and the grouped ridges)."""
    import oracle

    cores = len(os.sched_getaffinity(0))
    S = oracle.Spline(W["grid"], _host(W["Cu"]), _host(W["Cv"]), spherical=1)
    dx, dy = xp[1] - xp[0], yp[1] - yp[0]
    parts = []  # (flow map, FTLE, ridges) seconds per run
    for _ in range(reps + 1):
        t1 = time.perf_counter()
        fm, st, _ = oracle.flowmap_grid_2d("spline", np.array([-1.0]), t0, T, xp, yp, spline=S,
                                           nthreads=cores)
        t2 = time.perf_counter()
        ev, V = oracle.C_eig_2d(fm, dx, dy)
        ftle = np.where(ev[..., 1] > 1, np.log(np.maximum(ev[..., 1], 1e-300)) / (2 * abs(T)), 0)
        t3 = time.perf_counter()
        oracle.ridges(ftle, V[..., 1], xp, yp, percentile_=90)
        parts.append((t2 - t1, t3 - t2, time.perf_counter() - t3))
    runs = np.array(parts[1:])
    med = float(np.median(runs.sum(1)))
    P = len(xp) * len(yp)
    ferr = np.abs(fm_gpu - fm).max(-1) / np.abs(fm).max()
    ftle_err = np.abs(ftle_gpu - ftle) / np.abs(ftle).max()
    cpu = {"value": P / med, "unit": "grid points/s", "cores": cores, "kind": "port",
           "sample": f"full workload ({P} particles, flow map + FTLE + ridges), median of {reps} "
                     "runs after one warm-up",
           "seconds": med,
           "seconds_flowmap_ftle_ridges": [float(v) for v in np.median(runs, 0)],
           "cpu_model": _cpu_model()}
    err = {"flowmap_max_rel": float(ferr.max()),
           "flowmap_frac_within_1e-6": float(np.mean(ferr <= 1e-6)),
           "ftle_max_rel": float(ftle_err.max()),
           "ftle_rms_rel": float(np.sqrt(np.mean(ftle_err ** 2))),
           "ftle_frac_within_1e-6": float(np.mean(ftle_err <= 1e-6))}
    _, flag_o = oracle.ridge_predicate(ftle, V[..., 1], xp, yp, percentile_=90)
    _, flag_g = oracle.ridge_predicate(ftle_gpu, evec_gpu, xp, yp, percentile_=90)
    r_o = oracle.ridges(ftle, V[..., 1], xp, yp, percentile_=90)
    r_g = oracle.ridges(ftle_gpu, evec_gpu, xp, yp, percentile_=90)
    same = len(r_o) == len(r_g) and all(a.shape == b.shape for a, b in zip(r_o, r_g))
    err["ridges_p90"] = {"ridge_points_oracle": int(flag_o.sum()),
                         "ridge_points_gpu": int(flag_g.sum()),
                         "flagged_points_differing": int((flag_o ^ flag_g).sum()),
                         "ridges_oracle": len(r_o), "ridges_gpu": len(r_g),
                         "same_ridges_and_sizes": bool(same),
                         "max_point_abs_diff_deg": float(max((np.abs(a - b).max() for a, b in
                                                              zip(r_o, r_g)), default=0.0))
                         if same else None}
    return cpu, err


def _host(t):
    return t.detach().cpu().numpy()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# ---------------------------------------------------------------------------
def bench_double_gyre(args):
    from numbacs_amd import _lib, autotune
    from numbacs_amd.diagnostics import ftle_grid_2d_device
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2d_device

    f, p, _ = get_predefined_flow("double_gyre", int_direction=-1.0)
    x, y = np.linspace(0, 2, 401), np.linspace(0, 1, 201)
    xd, yd = _lib.to_dev(x), _lib.to_dev(y)
    fm = _lib.empty((401, 201, 2))
    ns = _lib.empty((401, 201), "i32")

    def flow():
        flowmap_grid_2d_device(f, 0.0, -10.0, xd, yd, p, out=fm, nsteps=ns)

    tuned, _ = autotune.tune(_lib.NBX_FLOW_DOUBLE_GYRE, _lib.NBX_FP64, flow, 401 * 201,
                             use_cache=False)

    def step():
        flow()
        return ftle_grid_2d_device(fm, -10.0, 0.005, 0.005)

    sec, _ = timed(step, args.steps, args.warmup)
    steps_total = int(ns.sum().item())
    return {"workload": "double gyre 401x201, t0=0, T=-10, DOP853 FP64 + FTLE",
            "grid_points_per_s": 401 * 201 / sec, "ms_per_step": 1e3 * sec,
            "particle_steps_per_s": steps_total / sec, "launch": tuned}


# ---------------------------------------------------------------------------
def bench_timeseries(args):
    return timeseries_legs(reps=max(1, args.steps // 4), tune=not args.no_tune)


# ---------------------------------------------------------------------------
def bench_composition(args):
    """Sliding-window FTLE by flow-map composition against direct integration
    (integration.py:609-737 of the reference): 72-hour backward windows
    stepped by 6 h over the 480-hour field; per window the composer
    integrates one 6-hour map and composes 12, the direct path integrates
    72 hours."""
    import torch

    from numbacs_amd import _lib
    from numbacs_amd.diagnostics import ftle_grid_2d_device
    from numbacs_amd.integration import FlowmapComposer, flowmap_grid_2d_device
    from numbacs_amd.workloads import particle_axes

    W = gridded_field(480)
    xp, yp = particle_axes(0.2)
    nx, ny = len(xp), len(yp)
    dx = float(xp[1] - xp[0])
    xd, yd = _lib.to_dev(xp), _lib.to_dev(yp)
    h, nT, K = -6.0, 12, 24
    t_first = 72.0 + 6.0 * (K + 8)
    # window w covers [t_w, t_w - 72] with t_w = t_first - 6 w
    starts = [t_first + h * w for w in range(K)]
    fm = _lib.empty((nx, ny, 2))
    torch.cuda.synchronize()
    a = time.perf_counter()
    direct_ftle = []
    for t0 in starts:
        flowmap_grid_2d_device(W["flow"], t0, h * nT, xd, yd, [-1.0], out=fm)
        direct_ftle.append(ftle_grid_2d_device(fm, h * nT, dx, dx))
    torch.cuda.synchronize()
    direct_s = (time.perf_counter() - a) / K
    comp = FlowmapComposer(W["flow"], xp, yp, h, nT, [-1.0])
    torch.cuda.synchronize()
    a = time.perf_counter()
    comp.initial(starts[0])
    torch.cuda.synchronize()
    initial_s = time.perf_counter() - a
    errs = []
    cf = _lib.empty((nx, ny, 2))
    a = time.perf_counter()
    comp_ftle = [ftle_grid_2d_device(comp.compose(out=cf), h * nT, dx, dx)]
    for w in range(1, K):
        comp.step(starts[w - 1] + h * nT)
        comp_ftle.append(ftle_grid_2d_device(comp.compose(out=cf), h * nT, dx, dx))
    torch.cuda.synchronize()
    comp_s = (time.perf_counter() - a) / K
    for d, c in zip(direct_ftle, comp_ftle):
        e = (c - d).abs() / d.abs().max()
        errs.append((float(e.max()), float(e.pow(2).mean().sqrt()), float((e <= 1e-2).double()
                                                                           .mean())))
    errs = np.array(errs)
    return {"workload": f"FTLE by flow-map composition: {K} backward windows of 72 h stepped "
                        "by 6 h (t0 = 264 .. 126 h) over the 271x101x480 field, particles "
                        "676x251, 12 maps of 6 h composed bilinearly, FP64",
            "direct_ms_per_window": 1e3 * direct_s,
            "composed_ms_per_window": 1e3 * comp_s,
            "composed_initial_window_ms": 1e3 * initial_s,
            "speedup": direct_s / comp_s,
            "grid_points_per_s_direct": nx * ny / direct_s,
            "grid_points_per_s_composed": nx * ny / comp_s,
            "ftle_vs_direct_max_rel": float(errs[:, 0].max()),
            "ftle_vs_direct_rms_rel_median": float(np.median(errs[:, 1])),
            "ftle_vs_direct_frac_within_1e-2": float(errs[:, 2].mean())}


def bench_next(args):
    """The rows next to the hot path, on the gridded-velocity workload's field
    and particles (676x251, t0 = 72 h, T = -72 h): ordered ridges
    (ridges.py:720-1054), the aux-grid flow map with its Cauchy-Green eigen
    stencils (integration.py:249-464, diagnostics.py:113-197), the trilinear
    flow (flows.py:418-506) and the 'xy' -> 'ij' velocity ingest; GPU times
    with the oracle's results and times beside each."""
    import oracle
    from oracle import ordered as O
    from numbacs_amd import _lib
    from numbacs_amd.diagnostics import ftle_eig_2d_device
    from numbacs_amd.extraction.ridges import ftle_ordered_ridges
    from numbacs_amd.flows import get_flow_linear_2D, ingest_xy
    from numbacs_amd.integration import flowmap_aux_grid_2d_device, flowmap_grid_2d_device
    from numbacs_amd.workloads import particle_axes, synthetic_velocity

    W = gridded_field()
    _, _, _, U, V = synthetic_velocity()
    xp, yp = particle_axes()
    nx, ny = len(xp), len(yp)
    P = nx * ny
    dx = float(xp[1] - xp[0])
    xd, yd = _lib.to_dev(xp), _lib.to_dev(yp)
    t0, T = 72.0, -72.0
    steps = max(3, args.steps // 4)
    # every 40th particle row for the oracle comparisons (the rest masked out)
    sample = np.ones((nx, ny), bool)
    sample[::40] = False
    rows = ~sample[:, 0]
    S = oracle.Spline(W["grid"], _host(W["Cu"]), _host(W["Cv"]), spherical=1)

    def agree(got, want, sens=None):
        scale = np.abs(want).max()
        e = np.abs(got - want).max(-1)
        r = {"frac_within_1e-6": float(np.mean(e <= 1e-6 * scale)),
             "max_rel": float(e.max() / scale)}
        if sens is not None:
            # a rounding-level difference in an accept/reject decision injects
            # an error of order rtol that the trajectory then amplifies: the
            # bound is 1e4 x the response to a 1e-9 relative shift of the start
            r["within_conditioning_bound"] = bool(np.all(e <= np.maximum(1e-6 * scale,
                                                                         1e4 * sens)))
        return r

    def sensitivity(run):
        base = run(xp, yp)
        s1 = np.abs(run(xp + 1e-9 * np.ptp(xp), yp) - base).max(-1)
        s2 = np.abs(run(xp, yp + 1e-9 * np.ptp(yp)) - base).max(-1)
        return base, np.maximum(s1, s2)

    out = {"workload": "rows next to the hot path on the gridded-velocity field (271x101x96, "
                       "spherical=1) and particles (676x251), t0=72 h, T=-72 h, FP64; oracle "
                       "comparisons on every 40th particle row"}

    # ordered ridges on the headline FTLE field (GPU predicate, native host linking)
    fm = _lib.empty((nx, ny, 2))
    flowmap_grid_2d_device(W["flow"], t0, T, xd, yd, [-1.0], out=fm)
    ftle, evec = ftle_eig_2d_device(fm, T, dx, dx)
    tol = 2 * dx
    sec, curves = timed(lambda: ftle_ordered_ridges(ftle, evec, xp, yp, tol, percentile=90,
                                                    min_ridge_pts=5), steps, 1)
    fh, eh = _host(ftle), _host(evec)
    a = time.perf_counter()
    want = O.ordered_ridges(fh, eh, xp, yp, tol, percentile=90, min_ridge_pts=5)
    cpu_s = time.perf_counter() - a
    out["ordered_ridges"] = {
        "what": "ftle_ordered_ridges(ftle, eigvec_max, x, y, dist_tol=0.4, percentile=90, "
                "min_ridge_pts=5) on the step's FTLE field",
        "ms": 1e3 * sec, "curves": len(curves), "points": int(sum(len(c) for c in curves)),
        "oracle_curves": len(want), "oracle_points": int(sum(len(c) for c in want)),
        "equal_to_oracle": bool(len(curves) == len(want) and
                                all(np.array_equal(c, w) for c, w in zip(curves, want))),
        "oracle_python_ms": 1e3 * cpu_s}

    # aux-grid flow map (5 particles per point) + Cauchy-Green eigen from it
    aux = _lib.empty((nx, ny, 5, 2))
    ev = _lib.empty((nx, ny, 2))
    vv = _lib.empty((nx, ny, 2, 2))
    h_aux = 1e-5

    def aux_step():
        flowmap_aux_grid_2d_device(W["flow"], t0, T, xd, yd, [-1.0], h=h_aux, out=aux)
        _lib.check(_lib.load().nbx_C_eig_aux_2d(_lib.ptr(aux), nx, ny, 5, dx, dx, h_aux, 1,
                                                None, _lib.ptr(ev), _lib.ptr(vv),
                                                _lib.stream_ptr()))

    sec, _ = timed(aux_step, steps, 1)
    ah = _host(aux)
    ref, sens = sensitivity(lambda xx, yy: oracle.flowmap_aux_grid_2d(
        "spline", np.array([-1.0]), t0, T, xx, yy, h=h_aux, spline=S, mask=sample))
    out["aux_grid"] = {
        "what": "flowmap_aux_grid_2D (h=1e-5, eig_main) + C_eig_aux_2D",
        "ms": 1e3 * sec, "grid_points_per_s": P / sec, "particles_per_s": 5 * P / sec,
        "centre_vs_flowmap_max_abs": float(np.abs(ah[:, :, 4] - _host(fm)).max()),
        "vs_oracle": agree(ah[rows].reshape(-1, 2), ref[rows].reshape(-1, 2),
                           sens[rows].reshape(-1))}

    # trilinear flow over the same velocity samples
    flin = get_flow_linear_2D(W["grid"], U, V, spherical=1)
    fml = _lib.empty((nx, ny, 2))
    nsl = _lib.empty((nx, ny), "i32")
    sec, _ = timed(lambda: flowmap_grid_2d_device(flin, t0, T, xd, yd, [-1.0], out=fml,
                                                  nsteps=nsl), steps, 1)
    spline_steps = _lib.empty((nx, ny), "i32")
    flowmap_grid_2d_device(W["flow"], t0, T, xd, yd, [-1.0], out=fm, nsteps=spline_steps)
    L = oracle.Linear(W["grid"], U, V, spherical=1)
    ref, sens = sensitivity(lambda xx, yy: oracle.flowmap_grid_2d(
        "linear", np.array([-1.0]), t0, T, xx, yy, spline=L, mask=sample)[0])
    out["trilinear"] = {
        "what": "flowmap_grid_2D over get_flow_linear_2D (trilinear, spherical=1)",
        "ms": 1e3 * sec, "grid_points_per_s": P / sec,
        "mean_steps_per_particle": float(_host(nsl).mean()),
        "spline_mean_steps_per_particle": float(_host(spline_steps).mean()),
        "vs_oracle": agree(_host(fml)[rows], ref[rows], sens[rows])}

    # natural cubic B-spline prefilter (get_interp_arrays_2D) of one component
    from numbacs_amd.flows import prefilter3_device

    pre = {}
    for nt_f in (96, 480):
        _, _, _, Uf, _ = synthetic_velocity(nt_f)
        ud = _lib.to_dev(Uf)
        sec, cd = timed(lambda: prefilter3_device(ud), steps, 1)
        cref = oracle.prefilter3(Uf)
        nb = Uf.nbytes + 8 * cref.size
        pre[f"{nt_f}x271x101"] = {
            "ms": 1e3 * sec, "hbm_gb_per_s_one_pass": nb / sec / 1e9,
            "max_rel_vs_oracle": float(np.abs(_host(cd) - cref).max() / np.abs(cref).max())}
    out["prefilter"] = dict(what="prefilter3 (get_interp_arrays_2D's per-component tricubic "
                                 "natural B-spline prefilter: scatter + three axis solves)", **pre)

    # 'xy' -> 'ij' ingest of the 480-frame field at m/s (x 3.6 -> km/h)
    _, _, _, U4, _ = synthetic_velocity(480)
    uxy = _lib.to_dev(np.ascontiguousarray(np.moveaxis(U4, 1, 2)) / 3.6)
    sec, uij = timed(lambda: ingest_xy(uxy, 3.6), steps, 1)
    nbytes = 2 * U4.nbytes
    exact = np.array_equal(_host(uij), np.moveaxis(_host(uxy), 1, 2) * 3.6)
    out["ingest_xy"] = {
        "what": "ingest_xy of one (480, 101, 271) component, x3.6",
        "ms": 1e3 * sec, "hbm_gb_per_s": nbytes / sec / 1e9, "hbm_frac": nbytes / sec / 8e12,
        "bytes_read_and_written": nbytes, "exact": bool(exact)}
    return out


def bench_precision(args):
    import oracle
    from numbacs_amd import _lib
    from numbacs_amd.diagnostics import ftle_eig_2d_device
    from numbacs_amd.extraction.ridges import ftle_ridge_pts_device
    from numbacs_amd.flows import get_predefined_flow
    from numbacs_amd.integration import flowmap_grid_2d_device
    from numbacs_amd.workloads import bickley_grid, bickley_params, double_gyre_params

    cases = [("double_gyre", 2001, 1001, -10.0), ("double_gyre", 2001, 1001, -20.0),
             ("bickley_jet", 2001, 801, -40.0)]
    out = {"workload": "precision study: FP32-displacement vs FP64 (GPU) and the FP64 oracle "
                       "on sampled rows", "cases": []}
    for name, nx, ny, T in cases:
        if name == "double_gyre":
            f = get_predefined_flow(name, return_default_params=False, return_domain=False)
            p = double_gyre_params(-1.0)
            x, y = np.linspace(0, 2, nx), np.linspace(0, 1, ny)
        else:
            f = get_predefined_flow(name, return_default_params=False, return_domain=False)
            p = bickley_params(-1.0)
            x, y = bickley_grid(nx, ny)
        xd, yd = _lib.to_dev(x), _lib.to_dev(y)
        dx, dy = x[1] - x[0], y[1] - y[0]
        rec = {"flow": name, "grid": [nx, ny], "T": T}
        res = {}
        for prec, code in (("fp64", 0), ("fp32", 1)):
            fm = _lib.empty((nx, ny, 2))

            def run():
                flowmap_grid_2d_device(f, 0.0, T, xd, yd, p, precision=code, out=fm)

            sec, _ = timed(run, 3, 1)
            ftle, ev = ftle_eig_2d_device(fm, T, dx, dy)
            pts, npts = ftle_ridge_pts_device(ftle, ev, xd, yd, nx, ny, dx, dy, 0.0, 90)
            res[prec] = (_lib.to_host(fm), _lib.to_host(ftle), _lib.to_host(pts[:npts]))
            rec[prec + "_grid_points_per_s"] = nx * ny / sec
        e = np.abs(res["fp32"][1] - res["fp64"][1]) / np.abs(res["fp64"][1]).max()
        rec["fp32_ftle_max_rel"] = float(e.max())
        rec["fp32_ftle_rms_rel"] = float(np.sqrt(np.mean(e ** 2)))
        rec["fp32_ftle_frac_within_1e-3"] = float(np.mean(e <= 1e-3))
        a, b = res["fp64"][2], res["fp32"][2]
        rec["ridge_points_fp64"], rec["ridge_points_fp32"] = int(len(a)), int(len(b))
        # oracle on every 50th row (flow map only)
        rows = np.arange(0, nx, 50)
        kind = name
        ref, _, _ = oracle.flowmap_grid_2d(kind, p, 0.0, T, x[rows], y)
        for prec in ("fp64", "fp32"):
            d = np.abs(res[prec][0][rows] - ref).max(-1) / np.abs(ref).max()
            rec[prec + "_vs_oracle_flowmap_max_rel"] = float(d.max())
            rec[prec + "_vs_oracle_frac_within_1e-6"] = float(np.mean(d <= 1e-6))
        out["cases"].append(rec)
    return out


# ---------------------------------------------------------------------------
def bench_surrogate(args):
    import torch

    from numbacs_amd import _lib
    from numbacs_amd import surrogate as S
    from numbacs_amd.diagnostics import ftle_grid_2d_device
    from numbacs_amd.integration import flowmap_grid_2d_device

    W = gridded_field(96)
    box = ((-100.0, 35.0), (-5.0, 45.0))
    per = 100
    nreq = max(1, args.surrogate_samples // (per * per))
    t_gen = time.perf_counter()
    reqs = S.sample_requests(nreq, per, box, (72.0, 95.0), (-72.0, -72.0), seed=5)
    inp, tgt = S.integrate_samples(W["flow"], [-1.0], reqs)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    print(f"surrogate: {int(inp.shape[0])} DOP853 samples in {t_gen:.1f} s", file=sys.stderr,
          flush=True)
    net = S.MLPSurrogate([-32.5, 20.0, 83.5, -72.0], [67.5, 25.0, 11.5, 1.0], [15.0, 15.0],
                         sigma=2.0, seed=7)
    t_train = time.perf_counter()
    losses = S.train(net, inp, tgt, epochs=3, batch=8192, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t_train
    print(f"surrogate: trained in {t_train:.1f} s, loss {losses[0]:.4g} -> {losses[-1]:.4g}",
          file=sys.stderr, flush=True)
    steps = 3 * (int(inp.shape[0]) // 8192)
    train_tflops = 3 * MLP_FLOPS_PER_SAMPLE * 8192 * steps / t_train / 1e12

    nx, ny = 2701, 1001
    x, y = np.linspace(-100, 35, nx), np.linspace(-5, 45, ny)
    dx = x[1] - x[0]
    t0s = np.linspace(72.0, 95.75, 96)[:: max(1, 96 // args.surrogate_t0s)][: args.surrogate_t0s]
    xd, yd = _lib.to_dev(x), _lib.to_dev(y)
    plain_s, hyb_s, infos, errs, mlp_s = 0.0, 0.0, [], [], 0.0
    for t0 in t0s:
        fm_ref = _lib.empty((nx, ny, 2))
        torch.cuda.synchronize()
        a = time.perf_counter()
        flowmap_grid_2d_device(W["flow"], t0, -72.0, xd, yd, [-1.0], out=fm_ref)
        ft_ref = ftle_grid_2d_device(fm_ref, -72.0, dx, dx)
        torch.cuda.synchronize()
        plain_s += time.perf_counter() - a
        a = time.perf_counter()
        inp_g = S.grid_inputs(x, y, t0, -72.0)
        net.predict_device(inp_g)
        torch.cuda.synchronize()
        mlp_s += time.perf_counter() - a
        a = time.perf_counter()
        fm, info = S.flowmap_with_fallback(net, W["flow"], [-1.0], x, y, t0, -72.0, 0.05 * dx)
        ft = ftle_grid_2d_device(fm, -72.0, dx, dx)
        torch.cuda.synchronize()
        hyb_s += time.perf_counter() - a
        infos.append(info)
        e = (ft - ft_ref).abs() / ft_ref.abs().max()
        errs.append(float(e.max()))
        print(f"surrogate: t0={t0:g} fallback {info['fallback_tiles']}/{info['tiles']} tiles",
              file=sys.stderr, flush=True)
    P = nx * ny * len(t0s)
    return {"workload": "learned surrogate: MLP (128 Fourier features, 4x256 GELU, BF16 MFMA, "
                        "FP32 accumulation) trained on DOP853 samples of the gridded field, "
                        f"inference on 2701x1001 at {len(t0s)} of the 96 t0 values, T=-72 h, "
                        "64x64 tiles with 5 probes, fallback when probe error > 0.05 dx",
            "training_samples": int(inp.shape[0]), "sample_generation_s": t_gen,
            "training_s": t_train, "training_tflops_bf16": train_tflops,
            "training_frac_of_bf16_peak": train_tflops / BF16_PEAK_TFLOPS,
            "loss_first_last": [losses[0], losses[-1]],
            "plain_fp64_grid_points_per_s": P / plain_s,
            "surrogate_only_grid_points_per_s": P / mlp_s,
            "surrogate_only_mlp_tflops": MLP_FLOPS_PER_SAMPLE * P / mlp_s / 1e12,
            "hybrid_grid_points_per_s": P / hyb_s,
            "hybrid_speedup_vs_plain": plain_s / hyb_s,
            "fallback_tiles": int(sum(i["fallback_tiles"] for i in infos)),
            "tiles": int(sum(i["tiles"] for i in infos)),
            "probe_err_median_deg": float(np.median([i["probe_err_median"] for i in infos])),
            "hybrid_ftle_max_rel_vs_plain": float(max(errs))}


def bench_stencil(args):
    """The FTLE stencils alone on a large flow map (8001 x 4001 points, a
    smooth synthetic map in HBM): ftle_grid_2D (16 B read + 8 B written per
    point) and the fused C_eig_2D + ftle_from_eig (16 B read + 24 B written),
    timed with HIP events on the launch stream against the 8 TB/s HBM peak."""
    import torch

    from numbacs_amd import _lib
    from numbacs_amd.diagnostics import ftle_eig_2d_device, ftle_grid_2d_device

    nx, ny = 8001, 4001
    t = _lib.torch()
    x = t.linspace(0, 2, nx, dtype=t.float64, device=_lib.device())
    y = t.linspace(0, 1, ny, dtype=t.float64, device=_lib.device())
    X, Y = t.meshgrid(x, y, indexing="ij")
    fm = t.stack([X + 0.1 * t.sin(6 * Y), Y + 0.1 * t.cos(5 * X)], -1).contiguous()
    del X, Y
    dx, dy = 2.0 / (nx - 1), 1.0 / (ny - 1)
    ftle = _lib.empty((nx, ny))
    evec = _lib.empty((nx, ny, 2))
    fm2 = t.empty_like(fm)
    out = {"workload": f"FTLE stencils on a {nx}x{ny} flow map (FP64, HBM-resident); a "
                       "device copy of the flow map as the achievable-bandwidth reference"}
    for name, fn, nbytes in (
            ("copy_reference", lambda: fm2.copy_(fm), 32),
            ("ftle_grid_2d", lambda: ftle_grid_2d_device(fm, -10.0, dx, dy, out=ftle), 24),
            ("ftle_eig_2d", lambda: ftle_eig_2d_device(fm, -10.0, dx, dy, ftle=ftle, evec=evec),
             40)):
        for _ in range(3):
            fn()
        ts = []
        for _ in range(max(5, args.steps)):
            e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            t.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = float(np.median(ts))
        gbs = nbytes * nx * ny / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "algorithmic_bytes": nbytes * nx * ny,
                     "bytes_per_point": nbytes, "achieved_GBs": gbs,
                     "hbm_frac": gbs / HBM_PEAK_GBS}
    return out


def bench_copycal(args):
    """A plain device copy of known size (512 MiB read + 512 MiB written,
    three dispatches): the reference for scaling rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    import torch

    n = 64 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"workload": "device copy calibration", "bytes_per_copy": n * 8, "copies": 3}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)
    import torch

    torch.cuda.set_device(0)
    fn = {"gridded": bench_gridded, "double_gyre": bench_double_gyre,
          "timeseries": bench_timeseries, "precision": bench_precision,
          "surrogate": bench_surrogate, "composition": bench_composition,
          "next": bench_next, "copycal": bench_copycal,
          "stencil": bench_stencil}[args.workload]
    print(json.dumps(fn(args)))


if __name__ == "__main__":
    main()
