"""CPU oracle of the flow-map surrogate MLP -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the network the HIP kernels implement (no reference
counterpart: NumbaCS has no learned surrogate):

  u   = (inputs - shift) / scale                       (x0, y0, t0, T), fp32
  z   = 2*pi * u @ B                                    B: (4, F/2) fp32
  h0  = bf16([sin z, cos z])                            (N, F)
  a_l = h_{l-1} @ W_l + b_l   (fp32 accumulation)       W_l bf16, b_l fp32
  h_l = bf16(gelu(a_l))       l = 1..L (tanh-form GELU)
  y   = h_L @ W_out + b_out   (fp32)                    (N, 2)

and its backward pass for the loss 0.5 * mean over samples of |y - target|^2
summed over the two outputs, with bf16 rounding of the same tensors the GPU
rounds (weights, activations, and the activation gradients fed to the next
GEMM).  Used by the tests as the gradient checker of the GPU kernels.
"""
import numpy as np

# working float type of the restatement (float32 like the GPU; the
# finite-difference check switches it to float64)
FT = np.float32
GELU_C = 0.7978845608028654   # sqrt(2/pi)
GELU_A = 0.044715


def bf16(x):
    """Round float32 to bfloat16 (round to nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    b = (b + 0x7FFF + lsb) & 0xFFFF0000
    out = b.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(x), x, out)


def gelu(a):
    a = np.asarray(a).astype(FT)
    c, k = FT(GELU_C), FT(GELU_A)
    t = np.tanh(c * (a + k * a * a * a))
    return FT(0.5) * a * (FT(1) + t)


def gelu_grad(a):
    a = np.asarray(a).astype(FT)
    c, k = FT(GELU_C), FT(GELU_A)
    inner = c * (a + k * a * a * a)
    t = np.tanh(inner)
    dinner = c * (FT(1) + FT(3) * k * a * a)
    return FT(0.5) * (FT(1) + t) + FT(0.5) * a * (FT(1) - t * t) * dinner


def features(inp, shift, scale, B):
    u = ((inp.astype(FT) - shift.astype(FT)) / scale.astype(FT))
    z = FT(2 * np.pi) * (u @ B.astype(FT))
    return bf16(np.concatenate([np.sin(z), np.cos(z)], axis=1))


def forward(inp, net):
    """net: dict(shift, scale, B, W: [W_1..W_L, W_out] (bf16-valued fp32),
    b: [b_1..b_L, b_out]).  Returns (y, cache)."""
    h = features(inp, net["shift"], net["scale"], net["B"])
    hs, pre = [h], []
    W, b = net["W"], net["b"]
    for l in range(len(W) - 1):
        a = (h.astype(np.float64) @ W[l].astype(np.float64)).astype(FT) + b[l]
        pre.append(a)
        h = bf16(gelu(a))
        hs.append(h)
    y = (h.astype(np.float64) @ W[-1].astype(np.float64)).astype(FT) + b[-1]
    return y, (hs, pre)


def loss_and_grads(inp, target, net):
    """0.5 * mean_n sum_c (y - target)^2 and its gradients w.r.t. W, b."""
    y, (hs, pre) = forward(inp, net)
    N = inp.shape[0]
    d = (y - target.astype(FT))
    loss = 0.5 * float(np.mean(np.sum(d.astype(np.float64) ** 2, axis=1)))
    dy = (d / FT(N)).astype(FT)
    W = net["W"]
    gW = [None] * len(W)
    gb = [None] * len(W)
    gW[-1] = (hs[-1].astype(np.float64).T @ dy.astype(np.float64)).astype(FT)
    gb[-1] = dy.sum(0, dtype=np.float64).astype(FT)
    dh = bf16((dy.astype(np.float64) @ W[-1].astype(np.float64).T).astype(FT))
    for l in range(len(W) - 2, -1, -1):
        da = bf16(dh * gelu_grad(pre[l]))
        gW[l] = (hs[l].astype(np.float64).T @ da.astype(np.float64)).astype(FT)
        gb[l] = da.sum(0, dtype=np.float64).astype(FT)
        if l > 0:
            dh = bf16((da.astype(np.float64) @ W[l].astype(np.float64).T).astype(FT))
    return loss, gW, gb


def init_net(rng, F=128, width=256, layers=4, sigma=1.0, shift=None, scale=None):
    B = (rng.standard_normal((4, F // 2)) * sigma).astype(np.float32)
    dims = [F] + [width] * layers
    W, b = [], []
    for l in range(layers):
        W.append(bf16(rng.standard_normal((dims[l], dims[l + 1])) * np.sqrt(2.0 / dims[l])))
        b.append(np.zeros(dims[l + 1], np.float32))
    W.append(bf16(rng.standard_normal((width, 2)) * np.sqrt(1.0 / width)))
    b.append(np.zeros(2, np.float32))
    return dict(shift=np.zeros(4, np.float32) if shift is None else shift,
                scale=np.ones(4, np.float32) if scale is None else scale, B=B, W=W, b=b)
