// BF16 MFMA kernels of the flow-map surrogate (an MLP mapping
// (x0, y0, t0, T) to the 2-D displacement): Fourier features, hidden layers
// as BF16 GEMMs with FP32 accumulation on v_mfma_f32_16x16x32_bf16 with
// fused bias/GELU epilogues, their backward GEMMs, the small output layer,
// the loss and an Adam update.  No reference counterpart: the surrogate is
// an addition on top of the reference's integrator path.
//
// GEMM tiling: a 256-thread block computes a 64x64 tile of C = A * B; each of
// its four wavefronts owns a 32x32 quarter as 2x2 MFMA 16x16 tiles.  K runs
// in steps of 32: the block stages a 64x32 slice of A and a 32x64 slice of B
// in LDS with k contiguous (rows padded to 40 elements against bank
// conflicts), transposing on the way in when the operand is stored with the
// other dimension contiguous, so forward (X W), backward-data (dY W^T) and
// backward-weight (X^T dY) products all use the same kernel.
//
// MFMA operand layout (v_mfma_f32_16x16x32_bf16, wave64): lane l supplies
// A[l % 16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l % 16]; it receives
// D[4*(l/16) + i][l % 16], i = 0..3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.h"

namespace nbx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

// bf16 as raw 16-bit patterns in memory; round to nearest even.
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    const uint32_t lsb = (u >> 16) & 1u;
    return static_cast<uint16_t>((u + 0x7fffu + lsb) >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGeluA = 0.044715f;

__device__ __forceinline__ float gelu(float a) {
    const float t = tanhf(kGeluC * (a + kGeluA * a * a * a));
    return 0.5f * a * (1.0f + t);
}

__device__ __forceinline__ float gelu_grad(float a) {
    const float inner = kGeluC * (a + kGeluA * a * a * a);
    const float t = tanhf(inner);
    const float dinner = kGeluC * (1.0f + 3.0f * kGeluA * a * a);
    return 0.5f * (1.0f + t) + 0.5f * a * (1.0f - t * t) * dinner;
}

// EPI_F32_ADD: split-K partial sums added into C32 (zeroed beforehand)
enum Epilogue { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_GELU_GRAD = 2, EPI_F32_ADD = 3 };

// C(m, n) = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn].  AK: sak == 1 (else sam == 1); BKC: sbk == 1
// (else sbn == 1).  M, N multiples of 64, K a multiple of 32.  Block z
// covers k in [z*ksplit, min((z+1)*ksplit, K)) (ksplit a multiple of 32;
// ksplit = K without a split).
template <bool AK, bool BKC, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const uint16_t* __restrict__ A, long sam, long sak,
                                              const uint16_t* __restrict__ B, long sbk, long sbn,
                                              int M, int N, int K, float* __restrict__ C32,
                                              long ldc, uint16_t* __restrict__ C16, long ldc16,
                                              const float* __restrict__ bias,
                                              const float* __restrict__ pre, long ldp,
                                              int ksplit) {
    __shared__ __attribute__((aligned(16))) uint16_t As[BM][LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN][LDK];
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const long m0 = static_cast<long>(blockIdx.y) * BM, n0 = static_cast<long>(blockIdx.x) * BN;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int kb = static_cast<int>(blockIdx.z) * ksplit;
    const int ke = kb + ksplit < K ? kb + ksplit : K;
    for (int k0 = kb; k0 < ke; k0 += BK) {
        // ---- stage A (64 x 32) and B (32 x 64) with k contiguous ----
        if (AK) {
            const int row = t >> 2, kq = (t & 3) * 8;
            const u16x8 v = *reinterpret_cast<const u16x8*>(A + (m0 + row) * sam + (k0 + kq));
            *reinterpret_cast<u16x8*>(&As[row][kq]) = v;
        } else {
            const int k = t >> 3, mq = (t & 7) * 8;
            const u16x8 v = *reinterpret_cast<const u16x8*>(A + (m0 + mq) + (k0 + k) * sak);
#pragma unroll
            for (int i = 0; i < 8; ++i) As[mq + i][k] = v[i];
        }
        if (BKC) {
            const int n = t >> 2, kq = (t & 3) * 8;
            const u16x8 v = *reinterpret_cast<const u16x8*>(B + (k0 + kq) + (n0 + n) * sbn);
            *reinterpret_cast<u16x8*>(&Bs[n][kq]) = v;
        } else {
            const int k = t >> 3, nq = (t & 7) * 8;
            const u16x8 v = *reinterpret_cast<const u16x8*>(B + (k0 + k) * sbk + (n0 + nq));
#pragma unroll
            for (int i = 0; i < 8; ++i) Bs[nq + i][k] = v[i];
        }
        __syncthreads();
        const int r16 = lane & 15, kk = 8 * (lane >> 4);
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            af[i] = *reinterpret_cast<const bf16x8*>(&As[32 * wm + 16 * i + r16][kk]);
            bfr[i] = *reinterpret_cast<const bf16x8*>(&Bs[32 * wn + 16 * i + r16][kk]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    const int c16 = lane & 15, rq = 4 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const long n = n0 + 32 * wn + 16 * j + c16;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const long m = m0 + 32 * wm + 16 * i + rq + q;
                const float v = acc[i][j][q];
                if (EPI == EPI_F32) {
                    C32[m * ldc + n] = v;
                } else if (EPI == EPI_F32_ADD) {
                    atomicAdd(&C32[m * ldc + n], v);
                } else if (EPI == EPI_BIAS_GELU) {
                    const float a = v + bias[n];
                    if (C32) C32[m * ldc + n] = a;
                    C16[m * ldc16 + n] = f2bf(gelu(a));
                } else {
                    const float dh = bf2f(f2bf(v));
                    C16[m * ldc16 + n] = f2bf(dh * gelu_grad(pre[m * ldp + n]));
                }
            }
        }
    }
}

// Fourier features: u = (inp - shift)/scale, z = 2 pi u B, h0 = [sin z, cos z].
__global__ void k_features(const float* __restrict__ inp, int64_t n, const float* __restrict__ shift,
                           const float* __restrict__ scale, const float* __restrict__ Bf, int half,
                           uint16_t* __restrict__ h0) {
    const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (idx >= n * half) return;
    const int64_t r = idx / half;
    const int c = static_cast<int>(idx - r * half);
    float z = 0.f;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const float u = (inp[4 * r + d] - shift[d]) / scale[d];
        z += u * Bf[d * half + c];
    }
    z *= 6.283185307179586f;
    h0[r * 2 * half + c] = f2bf(sinf(z));
    h0[r * 2 * half + half + c] = f2bf(cosf(z));
}

// y = h W_out + b_out for a 2-column output layer (one row per thread).
__global__ void k_out_fwd(const uint16_t* __restrict__ h, int64_t n, int width,
                          const uint16_t* __restrict__ W, const float* __restrict__ b,
                          float* __restrict__ y) {
    const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (r >= n) return;
    float s0 = 0.f, s1 = 0.f;
    const uint16_t* hr = h + r * width;
    // 16-byte loads of the row (width a multiple of 8), sums in k order
    for (int k8 = 0; k8 < width; k8 += 8) {
        const u16x8 hv = *reinterpret_cast<const u16x8*>(hr + k8);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k8 + q;
            const float a = bf2f(hv[q]);
            s0 += a * bf2f(W[2 * k]);
            s1 += a * bf2f(W[2 * k + 1]);
        }
    }
    y[2 * r] = s0 + b[0];
    y[2 * r + 1] = s1 + b[1];
}

// d = y - target; dy = d / n; loss += 0.5 * |d|^2 / n (block-reduced).
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ target, int64_t n,
                       float* __restrict__ dy, float* __restrict__ loss) {
    __shared__ float part[256];
    const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    float l = 0.f;
    if (r < n) {
        const float d0 = y[2 * r] - target[2 * r], d1 = y[2 * r + 1] - target[2 * r + 1];
        dy[2 * r] = d0 / static_cast<float>(n);
        dy[2 * r + 1] = d1 / static_cast<float>(n);
        l = 0.5f * (d0 * d0 + d1 * d1) / static_cast<float>(n);
    }
    part[threadIdx.x] = l;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// da = bf16(bf16(dy W_out^T) * gelu'(pre)) for the last hidden layer.
__global__ void k_out_bwd_data(const float* __restrict__ dy, int64_t n, int width,
                               const uint16_t* __restrict__ W, const float* __restrict__ pre,
                               uint16_t* __restrict__ da) {
    const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (idx >= n * width) return;
    const int64_t r = idx / width;
    const int k = static_cast<int>(idx - r * width);
    const float dh = bf2f(f2bf(dy[2 * r] * bf2f(W[2 * k]) + dy[2 * r + 1] * bf2f(W[2 * k + 1])));
    da[idx] = f2bf(dh * gelu_grad(pre[idx]));
}

// Rows per block of the column-sum kernels (k_out_bwd_weight, k_colsum):
// enough blocks to fill the GPU at a training batch of a few thousand rows.
constexpr int kSumRows = 32;

// gW_out[k][c] += sum_r h[r][k] dy[r][c], gb_out[c] += sum_r dy[r][c]:
// one block per kSumRows-row chunk, one thread per k, atomics across chunks.
__global__ void k_out_bwd_weight(const uint16_t* __restrict__ h, const float* __restrict__ dy,
                                 int64_t n, int width, float* __restrict__ gW,
                                 float* __restrict__ gb) {
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kSumRows;
    const int64_t r1 = r0 + kSumRows < n ? r0 + kSumRows : n;
    for (int k = threadIdx.x; k < width; k += blockDim.x) {
        float s0 = 0.f, s1 = 0.f;
        for (int64_t r = r0; r < r1; ++r) {
            const float a = bf2f(h[r * width + k]);
            s0 += a * dy[2 * r];
            s1 += a * dy[2 * r + 1];
        }
        atomicAdd(&gW[2 * k], s0);
        atomicAdd(&gW[2 * k + 1], s1);
    }
    if (threadIdx.x < 2) {
        float s = 0.f;
        for (int64_t r = r0; r < r1; ++r) s += dy[2 * r + threadIdx.x];
        atomicAdd(&gb[threadIdx.x], s);
    }
}

// gb[c] = sum_r da[r][c] (bf16 rows), one block per kSumRows-row chunk.
__global__ void k_colsum(const uint16_t* __restrict__ da, int64_t n, int width,
                         float* __restrict__ gb) {
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kSumRows;
    const int64_t r1 = r0 + kSumRows < n ? r0 + kSumRows : n;
    for (int c = threadIdx.x; c < width; c += blockDim.x) {
        float s = 0.f;
        for (int64_t r = r0; r < r1; ++r) s += bf2f(da[r * width + c]);
        atomicAdd(&gb[c], s);
    }
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) p[i] = v;
}

// Adam with bias correction on FP32 master weights; writes the BF16 copy.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ w16, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2, float wd) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = (mi / c1) / (sqrtf(vi / c2) + eps) + wd * w[i];
    const float wi = w[i] - lr * upd;
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

// Adam with the learning rate and both bias corrections read from device
// memory (sched = [lr, 1 - b1^t, 1 - b2^t]), so one captured training step
// can be replayed for every step of a schedule.
__global__ void k_adam_sched(float* __restrict__ w, const float* __restrict__ g,
                             float* __restrict__ m, float* __restrict__ v,
                             uint16_t* __restrict__ w16, int64_t n,
                             const float* __restrict__ sched, float b1, float b2, float eps,
                             float wd) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const float lr = sched[0], c1 = sched[1], c2 = sched[2];
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = (mi / c1) / (sqrtf(vi / c2) + eps) + wd * w[i];
    const float wi = w[i] - lr * upd;
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------
// Fused inference: Fourier features, the four hidden layers and the head in
// one launch.  A 256-thread block takes 64 particles; their activations stay
// in LDS (two 64 x 256 bf16 buffers, rows padded to 264 elements), each
// wavefront computes a 64 x 64 quarter of a layer's output as 4 x 4 MFMA
// 16x16x32 tiles, reading its weight fragments straight from global memory
// (weights stored transposed, k contiguous, so a lane's 8 k-values are one
// 16-byte load; the next k-step's fragments are loaded before the current
// step's MFMAs).  Every step is the unfused path's arithmetic in its order:
// the features of k_features, per layer the same k-steps of 32 accumulated
// in order then bias + GELU + bf16 (k_gemm's epilogue), the head's
// sequential fp32 sums of k_out_fwd and the output scaling -- so the result
// equals the unfused forward pass bit for bit.
constexpr int kInferRows = 64, kInferWidth = 256, kInferHalf = 64, kActLd = kInferWidth + 8;

struct MlpNetDev {
    const float* shift;
    const float* scale;
    const float* Bf;
    const uint16_t* WT[4];
    const float* b[4];
    const uint16_t* Wout;
    const float* bout;
    const float* out_scale;
};

__global__ __launch_bounds__(256) void k_mlp_infer(MlpNetDev net, const float* __restrict__ inp,
                                                   int64_t n, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t act[2][kInferRows][kActLd];
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kInferRows;
    // ---- Fourier features (k_features): thread -> row t/4, 16 frequencies
    {
        const int r = t >> 2, c0 = (t & 3) * 16;
        const int64_t gr = r0 + r < n ? r0 + r : n - 1;
        float u[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) u[d] = (inp[4 * gr + d] - net.shift[d]) / net.scale[d];
#pragma unroll 4
        for (int i = 0; i < 16; ++i) {
            const int c = c0 + i;
            float z = 0.f;
#pragma unroll
            for (int d = 0; d < 4; ++d) z += u[d] * net.Bf[d * kInferHalf + c];
            z *= 6.283185307179586f;
            act[0][r][c] = f2bf(sinf(z));
            act[0][r][kInferHalf + c] = f2bf(cosf(z));
        }
    }
    __syncthreads();
    // ---- hidden layers
    const int r16 = lane & 15, kk = 8 * (lane >> 4);
    int cur = 0;
#pragma unroll 1
    for (int l = 0; l < 4; ++l) {
        const int K = l == 0 ? 2 * kInferHalf : kInferWidth;
        const uint16_t* __restrict__ WT = net.WT[l];
        f32x4 acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        bf16x8 bnext[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bnext[j] = *reinterpret_cast<const bf16x8*>(
                WT + static_cast<long>(64 * w + 16 * j + r16) * K + kk);
#pragma unroll 1
        for (int k0 = 0; k0 < K; k0 += 32) {
            bf16x8 bcur[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) bcur[j] = bnext[j];
            if (k0 + 32 < K) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    bnext[j] = *reinterpret_cast<const bf16x8*>(
                        WT + static_cast<long>(64 * w + 16 * j + r16) * K + k0 + 32 + kk);
            }
            bf16x8 af[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *reinterpret_cast<const bf16x8*>(&act[cur][16 * i + r16][k0 + kk]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bcur[j], acc[i][j],
                                                                        0, 0, 0);
        }
        const int nxt = cur ^ 1;
        const int rq = 4 * (lane >> 4);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int nc = 64 * w + 16 * j + r16;
            const float bias = net.b[l][nc];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    act[nxt][16 * i + rq + q][nc] = f2bf(gelu(acc[i][j][q] + bias));
        }
        __syncthreads();
        cur = nxt;
    }
    // ---- head (k_out_fwd's sequential sums) and output scaling
    if (t < 2 * kInferRows) {
        const int r = t >> 1, c = t & 1;
        if (r0 + r < n) {
            float sum = 0.f;
            for (int k = 0; k < kInferWidth; ++k) sum += bf2f(act[cur][r][k]) * bf2f(net.Wout[2 * k + c]);
            out[2 * (r0 + r) + c] = (sum + net.bout[c]) * net.out_scale[c];
        }
    }
}

int check_gemm_shape(int M, int N, int K) {
    NBX_REQUIRE(M > 0 && N > 0 && K > 0, "empty GEMM");
    NBX_REQUIRE(M % BM == 0 && N % BN == 0 && K % BK == 0,
                "GEMM sizes must be multiples of 64 (M, N) and 32 (K)");
    return 0;
}

}  // namespace
}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_mlp_gemm(int mode, const uint16_t* A, const uint16_t* B, int M, int N, int K,
                 float* C32, uint16_t* C16, const float* bias, const float* pre,
                 nbx_stream stream) {
    if (int rc = check_gemm_shape(M, N, K)) return rc;
    NBX_REQUIRE(A != nullptr && B != nullptr, "NULL operand");
    hipStream_t s = as_stream(stream);
    const dim3 grid(N / BN, M / BM);
    switch (mode) {
        case 0:  // forward: C16 = gelu(A (MxK) @ B (KxN) + bias), C32 = pre-activation
            NBX_REQUIRE(C16 != nullptr && bias != nullptr, "forward needs C16 and bias");
            k_gemm<true, false, EPI_BIAS_GELU><<<grid, 256, 0, s>>>(A, K, 1, B, N, 1, M, N, K, C32,
                                                                    N, C16, N, bias, nullptr, 0,
                                                                    K);
            break;
        case 1:  // backward data: C16 = bf16(bf16(A (MxK) @ B^T) * gelu'(pre)), B stored (N x K)
            NBX_REQUIRE(C16 != nullptr && pre != nullptr, "backward-data needs C16 and pre");
            k_gemm<true, true, EPI_GELU_GRAD><<<grid, 256, 0, s>>>(A, K, 1, B, 1, K, M, N, K,
                                                                   nullptr, 0, C16, N, nullptr,
                                                                   pre, N, K);
            break;
        case 2: {  // backward weight: C32 = A^T @ B with A stored (K x M), B stored (K x N)
            NBX_REQUIRE(C32 != nullptr, "backward-weight needs C32");
            // K is the batch: split it over blocks when the output has few
            // tiles (a 256 x 256 weight is 16 tiles on 256 CUs), each block
            // adding its partial tile into C32
            const int tiles = (M / BM) * (N / BN);
            int splits = 256 / tiles;
            if (splits > K / 256) splits = K / 256;
            if (splits <= 1) {
                k_gemm<false, false, EPI_F32><<<grid, 256, 0, s>>>(A, 1, M, B, N, 1, M, N, K, C32,
                                                                   N, nullptr, 0, nullptr, nullptr,
                                                                   0, K);
            } else {
                const int ksplit = (K / splits + BK - 1) / BK * BK;
                const dim3 g3(N / BN, M / BM, (K + ksplit - 1) / ksplit);
                k_fill_f32<<<blocks_for(static_cast<int64_t>(M) * N, 256), 256, 0, s>>>(
                    C32, static_cast<int64_t>(M) * N, 0.f);
                k_gemm<false, false, EPI_F32_ADD><<<g3, 256, 0, s>>>(A, 1, M, B, N, 1, M, N, K,
                                                                     C32, N, nullptr, 0, nullptr,
                                                                     nullptr, 0, ksplit);
            }
            break;
        }
        case 3:  // plain: C32 = A (MxK) @ B (KxN)
            NBX_REQUIRE(C32 != nullptr, "plain GEMM needs C32");
            k_gemm<true, false, EPI_F32><<<grid, 256, 0, s>>>(A, K, 1, B, N, 1, M, N, K, C32, N,
                                                              nullptr, 0, nullptr, nullptr, 0, K);
            break;
        default:
            set_error("unknown GEMM mode");
            return -2;
    }
    NBX_LAUNCHED("mlp_gemm");
    return 0;
}

int nbx_mlp_features(const float* inp, int64_t n, const float* shift, const float* scale,
                     const float* Bf, int half, uint16_t* h0, nbx_stream stream) {
    NBX_REQUIRE(n >= 0 && half > 0, "bad feature shape");
    if (n == 0) return 0;
    k_features<<<blocks_for(n * half, 256), 256, 0, as_stream(stream)>>>(inp, n, shift, scale, Bf,
                                                                        half, h0);
    NBX_LAUNCHED("mlp_features");
    return 0;
}

int nbx_mlp_out_forward(const uint16_t* h, int64_t n, int width, const uint16_t* W,
                        const float* b, float* y, nbx_stream stream) {
    if (n == 0) return 0;
    k_out_fwd<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(h, n, width, W, b, y);
    NBX_LAUNCHED("mlp_out_fwd");
    return 0;
}

int nbx_mlp_loss(const float* y, const float* target, int64_t n, float* dy, float* loss,
                 nbx_stream stream) {
    NBX_REQUIRE(n > 0, "empty batch");
    hipStream_t s = as_stream(stream);
    k_fill_f32<<<1, 1, 0, s>>>(loss, 1, 0.f);
    k_loss<<<blocks_for(n, 256), 256, 0, s>>>(y, target, n, dy, loss);
    NBX_LAUNCHED("mlp_loss");
    return 0;
}

int nbx_mlp_out_backward(const float* dy, const uint16_t* h, const float* pre, int64_t n,
                         int width, const uint16_t* W, float* gW, float* gb, uint16_t* da,
                         nbx_stream stream) {
    NBX_REQUIRE(n > 0, "empty batch");
    hipStream_t s = as_stream(stream);
    k_fill_f32<<<blocks_for(2 * width, 256), 256, 0, s>>>(gW, 2 * width, 0.f);
    k_fill_f32<<<1, 2, 0, s>>>(gb, 2, 0.f);
    k_out_bwd_weight<<<static_cast<unsigned>((n + kSumRows - 1) / kSumRows), 256, 0, s>>>(
        h, dy, n, width, gW, gb);
    k_out_bwd_data<<<blocks_for(n * width, 256), 256, 0, s>>>(dy, n, width, W, pre, da);
    NBX_LAUNCHED("mlp_out_bwd");
    return 0;
}

int nbx_mlp_colsum(const uint16_t* da, int64_t n, int width, float* gb, nbx_stream stream) {
    NBX_REQUIRE(n > 0, "empty batch");
    hipStream_t s = as_stream(stream);
    k_fill_f32<<<blocks_for(width, 256), 256, 0, s>>>(gb, width, 0.f);
    k_colsum<<<static_cast<unsigned>((n + kSumRows - 1) / kSumRows), 256, 0, s>>>(da, n, width,
                                                                                  gb);
    NBX_LAUNCHED("mlp_colsum");
    return 0;
}

int nbx_mlp_adam(float* w, const float* g, float* m, float* v, uint16_t* w16, int64_t n, float lr,
                 float b1, float b2, float eps, int step, float weight_decay, nbx_stream stream) {
    NBX_REQUIRE(step >= 1, "Adam steps count from 1");
    if (n == 0) return 0;
    const float c1 = 1.f - powf(b1, static_cast<float>(step));
    const float c2 = 1.f - powf(b2, static_cast<float>(step));
    k_adam<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(w, g, m, v, w16, n, lr, b1, b2, eps,
                                                              c1, c2, weight_decay);
    NBX_LAUNCHED("mlp_adam");
    return 0;
}

int nbx_mlp_adam_sched(float* w, const float* g, float* m, float* v, uint16_t* w16, int64_t n,
                       const float* sched, float b1, float b2, float eps, float weight_decay,
                       nbx_stream stream) {
    NBX_REQUIRE(sched != nullptr, "NULL schedule");
    if (n == 0) return 0;
    k_adam_sched<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(w, g, m, v, w16, n, sched,
                                                                    b1, b2, eps, weight_decay);
    NBX_LAUNCHED("mlp_adam_sched");
    return 0;
}

int nbx_mlp_infer(const nbx_mlp_net* net, const float* inp, int64_t n, float* out,
                  nbx_stream stream) {
    NBX_REQUIRE(net != nullptr, "NULL network");
    NBX_REQUIRE(n >= 0, "negative particle count");
    NBX_REQUIRE(net->half == kInferHalf && net->width == kInferWidth && net->layers == 4,
                "the fused inference kernel runs the 128-feature, 4 x 256 network");
    if (n == 0) return 0;
    NBX_REQUIRE(inp != nullptr && out != nullptr, "NULL array");
    MlpNetDev d{net->shift, net->scale, net->Bf, {}, {}, net->Wout, net->bout, net->out_scale};
    for (int l = 0; l < 4; ++l) {
        NBX_REQUIRE(net->WT[l] != nullptr && net->b[l] != nullptr, "NULL layer");
        d.WT[l] = net->WT[l];
        d.b[l] = net->b[l];
    }
    NBX_REQUIRE(d.shift && d.scale && d.Bf && d.Wout && d.bout && d.out_scale, "NULL parameter");
    const int64_t blocks = (n + kInferRows - 1) / kInferRows;
    NBX_REQUIRE(blocks < (static_cast<int64_t>(1) << 31), "too many particles for one launch");
    k_mlp_infer<<<static_cast<unsigned>(blocks), 256, 0, as_stream(stream)>>>(d, inp, n, out);
    NBX_LAUNCHED("mlp_infer");
    return 0;
}

int nbx_mlp_to_bf16(const float* x, int64_t n, uint16_t* y, nbx_stream stream) {
    if (n == 0) return 0;
    k_to_bf16<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(x, n, y);
    NBX_LAUNCHED("mlp_to_bf16");
    return 0;
}

}  // extern "C"
