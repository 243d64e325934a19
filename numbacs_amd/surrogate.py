"""Learned flow-map surrogate with integrator fallback.

An MLP with Fourier-feature input maps (x0, y0, t0, T) to the 2-D
displacement of the flow map.  It is trained on trajectories the DOP853
integrator itself produces and runs on BF16 MFMA kernels with FP32
accumulation (numbacs_amd/csrc/mlp.hip); every kernel call goes through the
C ABI.  At inference the particle grid is cut into 64x64 tiles; a few probe
particles per tile are integrated with DOP853 and any tile whose probe
displacement error exceeds the threshold is integrated with DOP853 in full,
so the integrator stays both the fallback and the arbiter of the surrogate.

Network (defaults): 128 Fourier features (64 frequencies), 4 hidden layers
of 256 with tanh-form GELU, a 2-output linear head.  The numpy restatement
used to check the kernels' gradients is oracle/mlp.py.
"""
import ctypes
import math

import numpy as np

from . import _lib

ROWS = 64  # GEMM row granularity (batches are padded to a multiple)
# probe particles of a 64x64 tile (offsets inside the tile, clipped to the grid)
PROBES = ((8, 8), (8, 55), (55, 8), (55, 55), (32, 32))


def _pad(n):
    return (n + ROWS - 1) // ROWS * ROWS


class MLPSurrogate:
    def __init__(self, shift, scale, out_scale, features=128, width=256, layers=4, sigma=1.0,
                 seed=0):
        t = _lib.torch()
        self.F, self.width, self.layers = int(features), int(width), int(layers)
        rng = np.random.default_rng(seed)
        dev = _lib.device()
        self.shift = t.tensor(np.asarray(shift, np.float32), device=dev)
        self.scale = t.tensor(np.asarray(scale, np.float32), device=dev)
        self.out_scale = t.tensor(np.asarray(out_scale, np.float32), device=dev)
        self.B = t.tensor((rng.standard_normal((4, self.F // 2)) * sigma).astype(np.float32),
                          device=dev)
        dims = [self.F] + [self.width] * self.layers + [2]
        self.W, self.b = [], []
        for l in range(self.layers + 1):
            std = math.sqrt((2.0 if l < self.layers else 1.0) / dims[l])
            self.W.append(t.tensor((rng.standard_normal((dims[l], dims[l + 1])) * std)
                                   .astype(np.float32), device=dev))
            self.b.append(t.zeros(dims[l + 1], dtype=t.float32, device=dev))
        self.W16 = [t.empty(w.shape, dtype=t.bfloat16, device=dev) for w in self.W]
        for w, w16 in zip(self.W, self.W16):
            _lib.check(_lib.load().nbx_mlp_to_bf16(_lib.ptr(w), w.numel(), _lib.ptr(w16),
                                                   _lib.stream_ptr()))
        self.params = self.W + self.b
        self.m = [t.zeros_like(p) for p in self.params]
        self.v = [t.zeros_like(p) for p in self.params]
        self.step = 0

    # ------------------------------------------------------------------ --
    def state(self):
        """Host copy in the layout of oracle.mlp (weights as bf16 values)."""
        return dict(shift=_lib.to_host(self.shift), scale=_lib.to_host(self.scale),
                    B=_lib.to_host(self.B),
                    W=[_lib.to_host(w.float()) for w in self.W16],
                    b=[_lib.to_host(b) for b in self.b])

    def _bufs(self, n, train):
        t = _lib.torch()
        dev = _lib.device()
        h = [t.empty((n, self.F), dtype=t.bfloat16, device=dev)]
        h += [t.empty((n, self.width), dtype=t.bfloat16, device=dev) for _ in range(self.layers)]
        pre = [t.empty((n, self.width), dtype=t.float32, device=dev)
               for _ in range(self.layers)] if train else [None] * self.layers
        y = t.empty((n, 2), dtype=t.float32, device=dev)
        return h, pre, y

    def forward_device(self, inp, train=False):
        """inp: (n, 4) float32 device tensor, n a multiple of 64.  Returns
        (y (n, 2) in normalised displacement units, cache)."""
        L = _lib.load()
        s = _lib.stream_ptr()
        n = int(inp.shape[0])
        h, pre, y = self._bufs(n, train)
        _lib.check(L.nbx_mlp_features(_lib.ptr(inp), n, _lib.ptr(self.shift), _lib.ptr(self.scale),
                                      _lib.ptr(self.B), self.F // 2, _lib.ptr(h[0]), s))
        for l in range(self.layers):
            K = self.F if l == 0 else self.width
            _lib.check(L.nbx_mlp_gemm(0, _lib.ptr(h[l]), _lib.ptr(self.W16[l]), n, self.width, K,
                                      _lib.ptr(pre[l]), _lib.ptr(h[l + 1]), _lib.ptr(self.b[l]),
                                      None, s))
        _lib.check(L.nbx_mlp_out_forward(_lib.ptr(h[-1]), n, self.width, _lib.ptr(self.W16[-1]),
                                         _lib.ptr(self.b[-1]), _lib.ptr(y), s))
        return y, (h, pre)

    def gradients(self, inp, target):
        """Loss and gradients for one batch (target in normalised units):
        returns (loss tensor, [gW...], [gb...])."""
        t = _lib.torch()
        L = _lib.load()
        s = _lib.stream_ptr()
        n = int(inp.shape[0])
        y, (h, pre) = self.forward_device(inp, train=True)
        dev = _lib.device()
        dy = t.empty((n, 2), dtype=t.float32, device=dev)
        loss = t.empty((1,), dtype=t.float32, device=dev)
        _lib.check(L.nbx_mlp_loss(_lib.ptr(y), _lib.ptr(target), n, _lib.ptr(dy), _lib.ptr(loss), s))
        gW = [t.empty_like(w) for w in self.W]
        gb = [t.empty_like(b) for b in self.b]
        da = t.empty((n, self.width), dtype=t.bfloat16, device=dev)
        _lib.check(L.nbx_mlp_out_backward(_lib.ptr(dy), _lib.ptr(h[-1]), _lib.ptr(pre[-1]), n,
                                          self.width, _lib.ptr(self.W16[-1]), _lib.ptr(gW[-1]),
                                          _lib.ptr(gb[-1]), _lib.ptr(da), s))
        for l in range(self.layers - 1, -1, -1):
            K = self.F if l == 0 else self.width
            # weight gradient: gW_l = h_l^T da_l  (K x width)
            _lib.check(L.nbx_mlp_gemm(2, _lib.ptr(h[l]), _lib.ptr(da), K, self.width, n,
                                      _lib.ptr(gW[l]), None, None, None, s))
            _lib.check(L.nbx_mlp_colsum(_lib.ptr(da), n, self.width, _lib.ptr(gb[l]), s))
            if l > 0:
                da_prev = t.empty((n, self.width), dtype=t.bfloat16, device=dev)
                _lib.check(L.nbx_mlp_gemm(1, _lib.ptr(da), _lib.ptr(self.W16[l]), n, self.width,
                                          self.width, None, _lib.ptr(da_prev), None,
                                          _lib.ptr(pre[l - 1]), s))
                da = da_prev
        return loss, gW, gb

    def train_step_sched(self, inp, target, sched, b1=0.9, b2=0.999, eps=1e-8,
                         weight_decay=0.0):
        """One Adam step whose learning rate and bias corrections come from
        the device tensor sched = [lr, 1 - b1^t, 1 - b2^t]: every launch
        reads only device memory, so the step can be captured once and
        replayed for a whole schedule."""
        L = _lib.load()
        s = _lib.stream_ptr()
        loss, gW, gb = self.gradients(inp, target)
        for k, (p, g, m, v) in enumerate(zip(self.params, gW + gb, self.m, self.v)):
            w16 = self.W16[k] if k < len(self.W) else None
            _lib.check(L.nbx_mlp_adam_sched(_lib.ptr(p), _lib.ptr(g), _lib.ptr(m), _lib.ptr(v),
                                            _lib.ptr(w16), p.numel(), _lib.ptr(sched), b1, b2,
                                            eps, weight_decay, s))
        return loss

    def train_step(self, inp, target, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0):
        L = _lib.load()
        s = _lib.stream_ptr()
        loss, gW, gb = self.gradients(inp, target)
        self.step += 1
        grads = gW + gb
        for k, (p, g, m, v) in enumerate(zip(self.params, grads, self.m, self.v)):
            w16 = self.W16[k] if k < len(self.W) else None
            _lib.check(L.nbx_mlp_adam(_lib.ptr(p), _lib.ptr(g), _lib.ptr(m), _lib.ptr(v),
                                      _lib.ptr(w16), p.numel(), lr, b1, b2, eps, self.step,
                                      weight_decay, s))
        return loss

    def _inference_net(self):
        """The ctypes description of the current weights for the fused
        inference kernel (hidden weights transposed, k contiguous), rebuilt
        when training has moved the weights."""
        if getattr(self, "_infer_step", None) != self.step:
            self._WT = [w.t().contiguous() for w in self.W16[:-1]]
            net = _lib.MlpNet()
            net.shift, net.scale, net.Bf = (_lib.ptr(self.shift), _lib.ptr(self.scale),
                                            _lib.ptr(self.B))
            net.half, net.width, net.layers = self.F // 2, self.width, self.layers
            for l in range(self.layers):
                net.WT[l] = _lib.ptr(self._WT[l])
                net.b[l] = _lib.ptr(self.b[l])
            net.Wout, net.bout = _lib.ptr(self.W16[-1]), _lib.ptr(self.b[-1])
            net.out_scale = _lib.ptr(self.out_scale)
            self._net, self._infer_step = net, self.step
        return self._net

    def predict_device(self, inp, out=None):
        """Displacement (n, 2) float32 (physical units) for inputs (n, 4)
        float32, by the fused inference kernel (one launch)."""
        t = _lib.torch()
        n = int(inp.shape[0])
        if out is None:
            out = t.empty((n, 2), dtype=t.float32, device=_lib.device())
        if self.F != 128 or self.width != 256 or self.layers != 4:
            return self.predict_device_unfused(inp)
        _lib.check(_lib.load().nbx_mlp_infer(ctypes.byref(self._inference_net()),
                                             _lib.ptr(inp.contiguous()), n, _lib.ptr(out),
                                             _lib.stream_ptr()))
        return out

    def predict_device_unfused(self, inp, chunk=1 << 20):
        """The same displacement by the layer-by-layer kernels (features,
        one GEMM per layer, head) in chunks of up to chunk particles."""
        t = _lib.torch()
        n = int(inp.shape[0])
        out = t.empty((n, 2), dtype=t.float32, device=_lib.device())
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            m = _pad(b - a)
            x = inp[a:b]
            if m != b - a:
                x = t.cat([x, x[-1:].expand(m - (b - a), 4)], 0)
            y, _ = self.forward_device(x.contiguous())
            out[a:b] = y[: b - a] * self.out_scale
        return out


def grid_inputs(x, y, t0, T):
    """(nx*ny, 4) float32 device tensor of (x0, y0, t0, T) in 'ij' order."""
    t = _lib.torch()
    xd = _lib.to_dev(np.asarray(x, np.float64)).float()
    yd = _lib.to_dev(np.asarray(y, np.float64)).float()
    X, Y = t.meshgrid(xd, yd, indexing="ij")
    n = X.numel()
    return t.stack([X.reshape(n), Y.reshape(n), t.full((n,), float(t0), device=xd.device),
                    t.full((n,), float(T), device=xd.device)], 1).contiguous()


# ---------------------------------------------------------------------------
# Training data from the integrator, training loop, inference with fallback.

def integrate_samples(funcptr, params, requests, rtol=1e-6, atol=1e-8, precision="fp64"):
    """Integrate tensor-product sample grids with the batched DOP853 launch.
    requests: list of dict(t0, T, x, y).  Returns (inputs (n, 4) float32,
    displacements (n, 2) float32) as device tensors, in request order."""
    from .batch import FTLEBatch

    t = _lib.torch()
    b = FTLEBatch(funcptr, requests, params, rtol, atol, precision, ftle=False).run()
    ins, outs = [], []
    for q, d in zip(requests, b.requests):
        X, Y = t.meshgrid(d["x"], d["y"], indexing="ij")
        n = X.numel()
        x0 = t.stack([X.reshape(n), Y.reshape(n)], 1)
        fm = d["flowmap"].reshape(n, 2)
        ins.append(t.cat([x0.float(), t.full((n, 1), float(q["t0"]), device=x0.device),
                          t.full((n, 1), float(q["T"]), device=x0.device)], 1))
        outs.append((fm - x0).float())
    b.close()
    return t.cat(ins, 0).contiguous(), t.cat(outs, 0).contiguous()


def sample_requests(nreq, per_axis, box, t0_range, T_range, seed=0):
    """Random jittered sample grids: per request a (t0, T) and sorted random
    x, y axes over the box ((x0, x1), (y0, y1))."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nreq):
        out.append(dict(t0=float(rng.uniform(*t0_range)), T=float(rng.uniform(*T_range)),
                        x=np.sort(rng.uniform(box[0][0], box[0][1], per_axis)),
                        y=np.sort(rng.uniform(box[1][0], box[1][1], per_axis))))
    return out


def adam_schedule(steps, lr, lr_min, b1=0.9, b2=0.999):
    """Per step k (from 0): [lr_k, 1 - b1^(k+1), 1 - b2^(k+1)] in float32, with
    the cosine learning-rate decay from lr to lr_min."""
    k = np.arange(steps, dtype=np.float64)
    lrs = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * k / max(steps, 1)))
    t = np.arange(1, steps + 1, dtype=np.float32)
    c1 = np.float32(1) - np.power(np.float32(b1), t)
    c2 = np.float32(1) - np.power(np.float32(b2), t)
    return np.stack([lrs.astype(np.float32), c1, c2], axis=1)


def train(net, inputs, targets, epochs=3, batch=8192, lr=2e-3, lr_min=1e-4, seed=0,
          log=None, graph=True):
    """Adam over shuffled mini-batches with a cosine learning-rate decay;
    targets in physical units (divided by net.out_scale here).

    With graph (the default) one training step -- the batch gather, the
    forward and backward passes and Adam, about 30 launches -- is captured
    once into a HIP graph and replayed; per step the host only copies the
    batch's indices and its schedule row.  Without it every launch is
    issued from Python.  The loss and bias-gradient sums add with float
    atomics, so two runs agree to rounding, not bit for bit."""
    t = _lib.torch()
    n = int(inputs.shape[0]) // batch * batch
    tgt = (targets / net.out_scale).contiguous()
    g = t.Generator(device=inputs.device)
    g.manual_seed(seed)
    steps = epochs * (n // batch)
    table = t.from_numpy(adam_schedule(steps, lr, lr_min)).to(inputs.device)
    losses = []
    if not graph:
        k = 0
        for ep in range(epochs):
            perm = t.randperm(int(inputs.shape[0]), device=inputs.device, generator=g)[:n]
            for s in range(n // batch):
                idx = perm[s * batch:(s + 1) * batch]
                loss = net.train_step_sched(inputs[idx].contiguous(), tgt[idx].contiguous(),
                                            table[k])
                k += 1
                net.step += 1
                if s % 200 == 0:
                    losses.append(float(loss.item()))
                    if log:
                        log(f"epoch {ep} step {s} loss {losses[-1]:.3e}")
        return losses
    dev = inputs.device
    idx = t.zeros(batch, dtype=t.long, device=dev)
    inp_s = t.empty((batch, int(inputs.shape[1])), dtype=inputs.dtype, device=dev)
    tgt_s = t.empty((batch, int(tgt.shape[1])), dtype=tgt.dtype, device=dev)
    sched = table[0].clone()
    cg = t.cuda.CUDAGraph()
    with t.cuda.graph(cg):
        t.index_select(inputs, 0, idx, out=inp_s)
        t.index_select(tgt, 0, idx, out=tgt_s)
        loss = net.train_step_sched(inp_s, tgt_s, sched)
    k = 0
    for ep in range(epochs):
        perm = t.randperm(int(inputs.shape[0]), device=dev, generator=g)[:n]
        for s in range(n // batch):
            idx.copy_(perm[s * batch:(s + 1) * batch])
            sched.copy_(table[k])
            cg.replay()
            k += 1
            if s % 200 == 0:
                losses.append(float(loss.item()))
                if log:
                    log(f"epoch {ep} step {s} loss {losses[-1]:.3e}")
    t.cuda.synchronize()
    del cg
    net.step += steps
    return losses


def flowmap_with_fallback(net, funcptr, params, x, y, t0, T, tol, tile=64, probes=PROBES,
                          rtol=1e-6, atol=1e-8):
    """Flow map (nx, ny, 2) device tensor from the surrogate, with per-tile
    probe particles integrated by DOP853 and every tile whose probe
    displacement error exceeds tol integrated by DOP853 in full.  Returns
    (flow map, dict of counts and probe errors)."""
    from .batch import FTLEBatch
    from .integration import flowmap

    t = _lib.torch()
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    nx, ny = len(x), len(y)
    inp = grid_inputs(x, y, t0, T)
    disp = net.predict_device(inp)
    X, Y = t.meshgrid(_lib.to_dev(x), _lib.to_dev(y), indexing="ij")
    fm = t.stack([X, Y], -1) + disp.reshape(nx, ny, 2).double()
    # probes
    tiles = [(i0, j0) for i0 in range(0, nx, tile) for j0 in range(0, ny, tile)]
    pidx = []
    for i0, j0 in tiles:
        for pi, pj in probes:
            pidx.append((min(i0 + pi, nx - 1), min(j0 + pj, ny - 1)))
    pidx = np.array(pidx)
    pts = np.column_stack([x[pidx[:, 0]], y[pidx[:, 1]]])
    exact = flowmap(funcptr, t0, T, pts, params, rtol=rtol, atol=atol)
    pred = _lib.to_host(fm[pidx[:, 0], pidx[:, 1]])
    err = np.abs(pred - exact).max(1).reshape(len(tiles), len(probes)).max(1)
    bad = [tiles[k] for k in np.nonzero(err > tol)[0]]
    if bad:
        reqs = [dict(t0=t0, T=T, x=x[i0:i0 + tile], y=y[j0:j0 + tile]) for i0, j0 in bad]
        b = FTLEBatch(funcptr, reqs, params, rtol, atol, ftle=False).run()
        for (i0, j0), d in zip(bad, b.requests):
            fm[i0:i0 + d["nx"], j0:j0 + d["ny"]] = d["flowmap"]
        b.close()
    info = dict(tiles=len(tiles), fallback_tiles=len(bad), probes=len(pidx),
                probe_err_max=float(err.max()), probe_err_median=float(np.median(err)),
                fallback=bad)
    return fm, info
