"""Batched FTLE time series: many (t0, T, particle grid) requests over one
flow in one persistent work-queue launch plus one batched FTLE launch,
optionally captured once into a HIP graph and replayed.

The reference computes a time series with a Python loop over t0
(examples/time_series/plot_dg_time_series.py:101-124) or a numba prange over
t0 (lines 144-185); each window there is flowmap_grid_2D + ftle_grid_2D.
"""
import ctypes

import numpy as np

from . import _lib
from .flows import FlowHandle

PRECISION = {"fp64": _lib.NBX_FP64, "fp32": _lib.NBX_FP32_DISP}


class FTLEBatch:
    """A set of flow-map + FTLE requests over one flow.

    requests: iterable of dicts with keys t0, T, x, y and optionally mask and
    flowmap (a device tensor the flow map is written to); x, y are the
    particle axes (host arrays or device tensors).  All
    requests share params, tolerances and precision ('fp64' or 'fp32', the
    FP32-displacement form).  With eigvec=True each request's FTLE comes from
    the eigen-solver together with eigvec_max (the ridge path).
    """

    def __init__(self, funcptr, requests, params, rtol=1e-6, atol=1e-8, precision="fp64",
                 ftle=True, eigvec=False, max_steps=100000):
        if not isinstance(funcptr, FlowHandle):
            raise TypeError("funcptr must come from numbacs_amd.flows")
        if precision not in PRECISION:
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.flow = funcptr
        p = np.ascontiguousarray(params, dtype=np.float64).ravel()
        self.params = p
        self.requests = []
        reqs = []
        for q in requests:
            x = _lib.to_dev(q["x"])
            y = _lib.to_dev(q["y"])
            nx, ny = int(x.shape[0]), int(y.shape[0])
            d = dict(t0=float(q["t0"]), T=float(q["T"]), x=x, y=y, nx=nx, ny=ny)
            d["mask"] = None if q.get("mask") is None else _lib.to_dev(
                np.asarray(q["mask"]).astype(np.uint8), np.uint8)
            out = q.get("flowmap")
            if out is not None and (tuple(out.shape) != (nx, ny, 2) or not out.is_contiguous()):
                raise ValueError("a request's flowmap output must be a contiguous (nx, ny, 2) "
                                 "float64 device tensor")
            d["flowmap"] = _lib.empty((nx, ny, 2)) if out is None else out
            d["status"] = _lib.empty((nx, ny), "i32")
            d["nsteps"] = _lib.empty((nx, ny), "i32")
            d["ftle"] = _lib.empty((nx, ny)) if ftle else None
            d["eigvec_max"] = _lib.empty((nx, ny, 2)) if (ftle and eigvec) else None
            xh = _lib.to_host(x)
            yh = _lib.to_host(y)
            d["dx"] = float(xh[1] - xh[0]) if nx > 1 else 1.0
            d["dy"] = float(yh[1] - yh[0]) if ny > 1 else 1.0
            self.requests.append(d)
            reqs.append(_lib.Request(d["t0"], d["T"], _lib.ptr(x), nx, _lib.ptr(y), ny,
                                     _lib.ptr(d["mask"]), _lib.ptr(d["flowmap"]),
                                     _lib.ptr(d["status"]), _lib.ptr(d["nsteps"]),
                                     _lib.ptr(d["ftle"]), _lib.ptr(d["eigvec_max"]), d["dx"],
                                     d["dy"]))
        if not reqs:
            raise ValueError("a batch needs at least one request")
        self._table = (_lib.Request * len(reqs))(*reqs)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().nbx_batch_create(funcptr.handle, self._table, len(reqs),
                                                p.ctypes.data, p.size, float(rtol), float(atol),
                                                int(max_steps), PRECISION[precision],
                                                ctypes.byref(h)))
        self._h = h
        self._graph_stream = None
        self.particles = int(_lib.load().nbx_batch_particles(h))

    def run(self):
        """Enqueue the batch on the current stream."""
        _lib.check(_lib.load().nbx_batch_run(self._h, _lib.stream_ptr()))
        return self

    def capture(self):
        """Capture the batch into a HIP graph (on a private side stream)."""
        t = _lib.torch()
        if self._graph_stream is None:
            self._graph_stream = t.cuda.Stream()
        t.cuda.current_stream().synchronize()
        _lib.check(_lib.load().nbx_batch_capture(self._h, self._graph_stream.cuda_stream))
        return self

    def run_graph(self):
        """Replay the captured graph on the current stream."""
        _lib.check(_lib.load().nbx_batch_run_graph(self._h, _lib.stream_ptr()))
        return self

    def results(self):
        """Per request: dict(flowmap, ftle, eigvec_max, status, nsteps) as numpy."""
        out = []
        for d in self.requests:
            out.append({k: (None if d[k] is None else _lib.to_host(d[k]))
                        for k in ("flowmap", "ftle", "eigvec_max", "status", "nsteps")})
        return out

    def total_steps(self):
        return int(sum(int(d["nsteps"].sum().item()) for d in self.requests))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().nbx_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ftle_time_series(funcptr, t0s, Ts, x, y, params, rtol=1e-6, atol=1e-8, precision="fp64",
                     graph=False):
    """FTLE fields for every (t0, T) pair on one particle grid (x, y):
    returns a list of (nx, ny) arrays, computed as one batch."""
    Ts = np.broadcast_to(np.asarray(Ts, dtype=np.float64), np.shape(t0s))
    reqs = [dict(t0=t0, T=T, x=x, y=y) for t0, T in zip(np.asarray(t0s, dtype=np.float64), Ts)]
    b = FTLEBatch(funcptr, reqs, params, rtol, atol, precision)
    if graph:
        b.capture().run_graph()
    else:
        b.run()
    res = [r["ftle"] for r in b.results()]
    b.close()
    return res
