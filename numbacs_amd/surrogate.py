"""Learned flow-map surrogate with integrator fallback.

An MLP with Fourier-feature input maps (x0, y0, t0, T) to the 2-D
displacement of the flow map.  It is trained on trajectories the DOP853
integrator itself produces and runs on BF16 MFMA kernels with FP32
accumulation (numbacs_amd/csrc/mlp.hip); every kernel call goes through the
C ABI.  At inference the particle grid is cut into 64x64 tiles; a few probe
particles per tile are integrated with DOP853 and any tile whose probe
displacement error exceeds the threshold is integrated with DOP853 in full,
so the integrator stays both the fallback and the judge of the surrogate.

Network (defaults): 128 Fourier features (64 frequencies), 4 hidden layers
of 256 with tanh-form GELU, a 2-output linear head.  The numpy restatement
used to check the kernels' gradients is oracle/mlp.py.
"""
import math

import numpy as np

from . import _lib

ROWS = 64  # GEMM row granularity (batches are padded to a multiple)


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

    def predict_device(self, inp, chunk=1 << 20):
        """Displacement (n, 2) float32 (physical units) for inputs (n, 4)."""
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
