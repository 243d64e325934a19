// Device velocity fields ("flows").  Each flow is a small trivially-copyable
// functor passed by value in the kernel arguments and inlined into the
// integrator; it evaluates dy = p0 * v(p0 * t, y), the sign-flipped
// right-hand side convention of the reference (params[0] = integration
// direction, flows.py:1146-1158, docs userguide.rst:216-227).
//
// Call form: f(t, px, py, f0, f1) with time and position in double and the
// velocity returned in the stage precision R of the integrator.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace nbx {

constexpr double kPi = 3.141592653589793;

// Double gyre, flows.py:1146-1158.  p = [dir, A, eps, alpha, omega, psi].
// The float form evaluates the same expression in float.
template <class R>
struct DoubleGyre {
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ void prepare_factors() {}
    double p0, A, eps, alpha, omega, psi;

    __device__ __forceinline__ void operator()(double t, double y0d, double y1d, R& f0,
                                               R& f1) const {
        if constexpr (sizeof(R) == sizeof(double)) {
            const double y0 = y0d, y1 = y1d;
            const double tt = p0 * t;
            const double a = eps * sin(omega * tt + psi);
            const double b = 1 - 2 * a;
            const double f = a * (y0 * y0) + b * y0;
            const double df = 2 * a * y0 + b;
            double sf, cf, sy, cy;
            sincos(kPi * f, &sf, &cf);
            sincos(kPi * y1, &sy, &cy);
            f0 = p0 * (-kPi * A * sf * cy - alpha * y0);
            f1 = p0 * (kPi * A * cf * sy * df - alpha * y1);
        } else {
            // time enters through sin(omega*tt + psi): keep that argument in
            // double, the rest in float
            const float y0 = static_cast<float>(y0d), y1 = static_cast<float>(y1d);
            const double tt = p0 * t;
            const float a = static_cast<float>(eps * sin(omega * tt + psi));
            const float b = 1.0f - 2.0f * a;
            const float f = a * (y0 * y0) + b * y0;
            const float df = 2.0f * a * y0 + b;
            const float pi = static_cast<float>(kPi);
            float sf, cf, sy, cy;
            sincosf(pi * f, &sf, &cf);
            sincosf(pi * y1, &sy, &cy);
            const float q0 = static_cast<float>(p0), qA = static_cast<float>(A),
                        qal = static_cast<float>(alpha);
            f0 = q0 * (-pi * qA * sf * cy - qal * y0);
            f1 = q0 * (pi * qA * cf * sy * df - qal * y1);
        }
    }
};

// Bickley jet with three travelling modes, flows.py:1182-1213.
// p = [dir, U0, L, A1, A2, A3, k1, k2, k3, c1, c2, c3].
template <class R>
struct BickleyJet {
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ void prepare_factors() {}
    double p[12];

    __device__ __forceinline__ void operator()(double t, double y0, double y1, R& f0,
                                               R& f1) const {
        const double tt = p[0] * t;
        if constexpr (sizeof(R) == sizeof(double)) {
            const double Y = y1 / p[2];
            const double ch = cosh(Y);
            const double sech2 = 1 / (ch * ch);
            double s1, c1, s2, c2, s3, c3;
            sincos(p[6] * (y0 - p[9] * tt), &s1, &c1);
            sincos(p[7] * (y0 - p[10] * tt), &s2, &c2);
            sincos(p[8] * (y0 - p[11] * tt), &s3, &c3);
            f0 = p[0] * (p[1] * sech2 +
                         2 * p[1] * tanh(Y) * sech2 * (p[3] * c1 + p[4] * c2 + p[5] * c3));
            f1 = -p[0] * (p[1] * p[2] * sech2 *
                          (p[3] * p[6] * s1 + p[4] * p[7] * s2 + p[5] * p[8] * s3));
        } else {
            // phase arguments k (x - c t) reduced in double, the rest in float
            const float Y = static_cast<float>(y1 / p[2]);
            const float ch = coshf(Y);
            const float sech2 = 1.0f / (ch * ch);
            float s1, c1, s2, c2, s3, c3;
            sincosf(static_cast<float>(fmod(p[6] * (y0 - p[9] * tt), 2 * kPi)), &s1, &c1);
            sincosf(static_cast<float>(fmod(p[7] * (y0 - p[10] * tt), 2 * kPi)), &s2, &c2);
            sincosf(static_cast<float>(fmod(p[8] * (y0 - p[11] * tt), 2 * kPi)), &s3, &c3);
            const float q0 = static_cast<float>(p[0]), U0 = static_cast<float>(p[1]),
                        L = static_cast<float>(p[2]);
            const float A1 = static_cast<float>(p[3]), A2 = static_cast<float>(p[4]),
                        A3 = static_cast<float>(p[5]);
            const float K1 = static_cast<float>(p[6]), K2 = static_cast<float>(p[7]),
                        K3 = static_cast<float>(p[8]);
            f0 = q0 * (U0 * sech2 + 2.0f * U0 * tanhf(Y) * sech2 * (A1 * c1 + A2 * c2 + A3 * c3));
            f1 = -q0 * (U0 * L * sech2 * (A1 * K1 * s1 + A2 * K2 * s2 + A3 * K3 * s3));
        }
    }
};

// Uniform cubic B-spline basis at local coordinate s in [0, 1].
// The 1/6 scaling is a multiplication by the rounded reciprocal (within one
// rounding of the quotient), which spares twelve IEEE divisions per
// evaluation; the 1/2 of the derivative weights is exact either way.
template <class R>
__device__ __forceinline__ void bspline_weights(R s, R w[4]) {
    constexpr R sixth = R(1) / R(6);
    const R s2 = s * s, s3 = s2 * s, u = R(1) - s;
    w[0] = u * u * u * sixth;
    w[1] = (R(3) * s3 - R(6) * s2 + R(4)) * sixth;
    w[2] = (R(-3) * s3 + R(3) * s2 + R(3) * s + R(1)) * sixth;
    w[3] = s3 * sixth;
}

// The same weights times 6 (no scaling: the 4-lane gather applies the
// three axes' 1/6 once, as 1/216 on its y-weight)
template <class R>
__device__ __forceinline__ void bspline_weights6(R s, R w[4]) {
    const R s2 = s * s, s3 = s2 * s, u = R(1) - s;
    w[0] = u * u * u;
    w[1] = R(3) * s3 - R(6) * s2 + R(4);
    w[2] = R(-3) * s3 + R(3) * s2 + R(3) * s + R(1);
    w[3] = s3;
}

template <class R>
__device__ __forceinline__ void bspline_dweights(R s, R dw[4]) {
    const R s2 = s * s, u = R(1) - s;
    dw[0] = -u * u * R(0.5);
    dw[1] = (R(3) * s2 - R(4) * s) * R(0.5);
    dw[2] = (R(-3) * s2 + R(2) * s + R(1)) * R(0.5);
    dw[3] = s2 * R(0.5);
}

// w[i] for a lane-dependent i without dynamic register indexing
template <class R>
__device__ __forceinline__ R pick4(const R w[4], int i) {
    return i == 0 ? w[0] : (i == 1 ? w[1] : (i == 2 ? w[2] : w[3]));
}

// Python's float modulo (result takes the sign of the divisor).
__device__ __forceinline__ double py_mod(double a, double m) {
    // fmod(a, m) == a exactly when |a| < |m|: the library call only runs for
    // arguments a full period or more away
    double r = fabs(a) < fabs(m) ? a : fmod(a, m);
    if (r != 0.0) {
        if ((m < 0) != (r < 0)) r += m;
    } else {
        r = copysign(0.0, m);
    }
    return r;
}

// Degrees to radians as the reference writes it, (deg * pi) / 180: the
// quotient by 180 is the product with the rounded reciprocal plus one FMA
// correction (Markstein), which is the correctly rounded quotient -- equal
// to the IEEE division on 2e8 sampled latitudes and random doubles -- at
// four instructions instead of a dozen for the division sequence.
__device__ __forceinline__ double deg_to_rad(double deg) {
    constexpr double inv180 = 1.0 / 180.0;
    const double a = deg * kPi;
    const double q = a * inv180;
    return fma(fma(-q, 180.0, a), inv180, q);
}

// Cosine of a latitude x in radians as sin(pi/2 - |x|): pi/2 in two parts
// (Cody-Waite) keeps the reduced argument u accurate to the last bits, and
// sin(u) for |u| <= pi/2 is its Taylor series to u^25 (truncation below
// 3e-23) by Horner in u^2: within 2.2e-16 relative of the exact cosine on
// all latitudes (5e7 samples against long-double cos; the library cosine
// agrees with it bit for bit on 71%), without the library's range reduction
// (the spherical RHS takes one cosine per evaluation; the library's took ~50
// instructions of it).  Valid for |x| <= pi; larger arguments (no latitude)
// take the library cosine.
// fma(a, b, c) for a coefficient c, taken as a scalar-register operand of
// v_fma_f64.  Written as fma(), the compiler picks v_fmac_f64, whose addend
// is also its destination, and first copies each 64-bit coefficient into
// vector registers: two extra vector instructions per Horner step.  The
// result is the same fused multiply-add.
__device__ __forceinline__ double fma_coef(double a, double b, double c) {
#if __HIP_DEVICE_COMPILE__
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
#else
    return fma(a, b, c);
#endif
}

__device__ __forceinline__ double cos_lat(double x) {
    constexpr double pio2_hi = 1.5707963267948966;      // double nearest pi/2
    constexpr double pio2_lo = 6.123233995736766e-17;   // pi/2 - pio2_hi
    const double ax = fabs(x);
    if (ax > 2 * pio2_hi) return cos(x);
    const double u = (pio2_hi - ax) + pio2_lo;
    const double s = u * u;
    double p = 1.0 / 15511210043330985984000000.0;      // 1/25!
    p = fma_coef(p, s, -1.0 / 25852016738884976640000.0);    // -1/23!
    p = fma_coef(p, s, 1.0 / 51090942171709440000.0);        // 1/21!
    p = fma_coef(p, s, -1.0 / 121645100408832000.0);         // -1/19!
    p = fma_coef(p, s, 1.0 / 355687428096000.0);             // 1/17!
    p = fma_coef(p, s, -1.0 / 1307674368000.0);              // -1/15!
    p = fma_coef(p, s, 1.0 / 6227020800.0);                  // 1/13!
    p = fma_coef(p, s, -1.0 / 39916800.0);                   // -1/11!
    p = fma_coef(p, s, 1.0 / 362880.0);                      // 1/9!
    p = fma_coef(p, s, -1.0 / 5040.0);                       // -1/7!
    p = fma_coef(p, s, 1.0 / 120.0);                         // 1/5!
    p = fma_coef(p, s, -1.0 / 6.0);                          // -1/3!
    return fma(u * s, p, u);
}

// The float form's latitude cosine: the same reduction in float, Taylor to
// u^13 (truncation below 7e-10), within 1.3e-7 relative.
__device__ __forceinline__ float cos_lat(float x) {
    constexpr float pio2_hi = 1.57079637f;              // float nearest pi/2
    constexpr float pio2_lo = -4.37113900e-8f;          // pi/2 - pio2_hi
    const float ax = fabsf(x);
    if (ax > 2 * pio2_hi) return cosf(x);
    const float u = (pio2_hi - ax) + pio2_lo;
    const float s = u * u;
    float p = 1.0f / 6227020800.0f;                     // 1/13!
    p = fmaf(p, s, -1.0f / 39916800.0f);
    p = fmaf(p, s, 1.0f / 362880.0f);
    p = fmaf(p, s, -1.0f / 5040.0f);
    p = fmaf(p, s, 1.0f / 120.0f);
    p = fmaf(p, s, -1.0f / 6.0f);
    return fmaf(u * s, p, u);
}

// Coefficient storage element.  FP64: one (u, v) pair, so one gather serves
// both velocity components.  FP32: (u, v) at time index t and at t + 1 in one
// 16-byte element, so one gather serves two time slices of the stencil and a
// stencil is 8 rows of 4 elements instead of 16 rows (the gathers are bound
// by distinct cache lines, not bytes).
template <class R> struct CoefPair;
template <> struct CoefPair<double> { using type = double2; };
template <> struct CoefPair<float> { using type = float4; };

// Sum of a value over the G lanes of an aligned lane group, with a butterfly
// of DPP quad permutations, so every lane of the group ends with the same
// bits (the additions are commutative pairwise, and each level combines the
// same two partial sums in every lane).  The empty asm materialises v first:
// otherwise the compiler may fuse the multiply that produced v into the
// first add (fma(a, b, other lane's rounded a*b)), and the lanes of a group
// would then round differently and take different step-size decisions.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    asm volatile("" : "+v"(v));
    if constexpr (G >= 2)
        v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
    if constexpr (G >= 4)
        v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
    return v;
}

template <int G>
__device__ __forceinline__ double group_sum(double v) {
    auto xchg = [](double a, int ctrl) {
        const long long b = __double_as_longlong(a);
        int lo = static_cast<int>(b & 0xffffffffll), hi = static_cast<int>(b >> 32);
        if (ctrl == 0xB1) {
            lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);
            hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
        } else {
            lo = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xF, 0xF, false);
            hi = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xF, 0xF, false);
        }
        return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                    static_cast<unsigned int>(lo));
    };
    asm volatile("" : "+v"(v));
    if constexpr (G >= 2) v += xchg(v, 0xB1);
    if constexpr (G >= 4) v += xchg(v, 0x4E);
    return v;
}

// Tricubic natural B-spline velocity on a uniform (t, x, y) grid (the
// reference's get_flow_2D over interpolation.eval_spline, flows.py:121-258).
// C holds (nt+2, nx+2, ny+2) (u, v) pairs, y fastest.  R is the coefficient
// and evaluation precision (double for the plain path, float for the
// FP32-displacement form).  G lanes (1, 2 or 4, aligned in the wavefront)
// share one evaluation and combine partial sums with group_sum:
//   G = 2: lane g gathers the t-slices a with a % 2 == g (whole 4-pair rows);
//   G = 4: lane g gathers y-coefficient g of all 16 rows, so one load
//          instruction covers one contiguous row per particle -- a third of
//          the distinct cache lines per instruction of the row-per-lane
//          forms, and the vector L1's tag rate (one line per clock) is what
//          bounds these gathers.
//
// Out-of-box queries: extrap 0 ("constant") is zero outside the grid box,
// the rule of the reference dependency's eval_linear (pinned by the
// reference's composition fixture) carried over to its spline evaluation
// (inferred: no reference test leaves the grid); extrap 2 ("nearest") clamps
// the query into the box per axis; extrap 1 ("linear") adds the gradient at
// the clamped point times the outward offset.  L says whether that gradient
// path is compiled in: it is the register-heaviest code of the functor, and
// kernels built without it (L = false) serve modes 0 and 2 with fewer
// registers (FP32 4-lane flow map: 167 instead of 234 VGPRs, 3 waves per
// SIMD instead of 2; FP64: 221 instead of 256).
// A uniform double moved into a vector register pair.  The flow kernels hold
// the spline's box and spacing for their whole solve; in scalar registers
// (next to the solver's constants and the kernel's pointers) they spill to
// vector-register lanes and every evaluation reloads them one lane read at
// a time.  The copy cannot be folded back into the scalar value.
__device__ __forceinline__ double in_vgpr(double x) {
    double y;
    asm volatile("v_mov_b64 %0, %1" : "=v"(y) : "v"(x));
    return y;
}

// A box of coefficient elements copied to LDS by a one-wavefront block for
// the stencils of one DOP853 step (cells it0.., ix0.., iy0.. of the
// coefficient array, bt x bx x by elements, y fastest): a stencil whose 4 x 4
// x 4 elements all lie inside is read from LDS, any other from HBM/L2.
template <class P>
struct StageBox {
    const P* lds = nullptr;
    int it0 = 0, ix0 = 0, iy0 = 0, bt = 0, bx = 0, by = 0;
    bool on = false;
    __device__ __forceinline__ bool covers(int it, int ix, int iy) const {
        return on && static_cast<unsigned>(it - it0) <= static_cast<unsigned>(bt - 4) &&
               static_cast<unsigned>(ix - ix0) <= static_cast<unsigned>(bx - 4) &&
               static_cast<unsigned>(iy - iy0) <= static_cast<unsigned>(by - 4);
    }
};
struct NoStageBox {};

template <class R, int G = 1, bool L = false, bool S = false>
struct Spline3 {
    using P = typename CoefPair<R>::type;
    const P* __restrict__ C;
    // S (4-lane form only): stencils inside the step's LDS box are read there
    std::conditional_t<S, StageBox<P>, NoStageBox> box;
    double lo[3], hi[3], dinv[3];
    int nm2[3];   // n - 2 per axis: the largest cell index
    int sx, st;   // strides (in pairs) of the x and t axes
    int spherical, extrap;
    double r;     // sphere radius (km) for spherical = 1, 2
    double p0;    // integration direction
    double kdeg;  // p0 * (180 / (pi r)): km/h -> degrees/h, set with p0 and r
    // 4-lane unstaged form: the lane's column of the last stencil it
    // gathered and that stencil's origin (element offset), so a query in
    // the same cell as the particle's previous one reads registers instead
    // of the cache hierarchy (the kernel's own copy of the functor holds it)
    static constexpr bool kReuse = (G == 4) && !S;
    static constexpr int kRows = sizeof(R) == sizeof(double) ? 16 : 8;
    mutable P cst[kReuse ? kRows : 1];
    mutable long ckey = -1;
    // bytes of C (the 4-lane form's buffer loads: under 4 GiB, checked by
    // the dispatcher)
    unsigned cbytes = 0;

    // at the start of a kernel, on its own copy of the functor
    __device__ __forceinline__ void prepare() {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = in_vgpr(lo[d]);
            hi[d] = in_vgpr(hi[d]);
            dinv[d] = in_vgpr(dinv[d]);
        }
        prepare_factors();
    }
    // the direction and the degree factor only (for kernels with no
    // vector registers to spare for the box)
    __device__ __forceinline__ void prepare_factors() {
        p0 = in_vgpr(p0);
        kdeg = in_vgpr(kdeg);
    }

    __device__ __forceinline__ void cell(int d, double q, int& i, R& s) const {
        const double u = (q - lo[d]) * dinv[d];
        int k = static_cast<int>(floor(u));
        k = k > nm2[d] ? nm2[d] : k;
        k = k < 0 ? 0 : k;
        i = k;
        s = static_cast<R>(u - k);
    }

    // (u, v) of stencil time slice a (0..3), x row b, y column c, relative
    // to the stencil origin base
    __device__ __forceinline__ void coef(const P* base, int a, int b, int c, R& u, R& v) const {
        if constexpr (sizeof(R) == sizeof(double)) {
            const P e = base[a * st + b * sx + c];
            u = e.x;
            v = e.y;
        } else {
            const P e = base[(a & ~1) * st + b * sx + c];
            u = (a & 1) ? e.z : e.x;
            v = (a & 1) ? e.w : e.y;
        }
    }

    // y-weighted sum of one stencil row (slice a, x row b)
    __device__ __forceinline__ void row_sum(const P* base, int a, int b, const R wy[4], R& ru,
                                            R& rv) const {
        R cu, cv;
        coef(base, a, b, 0, cu, cv);
        ru = wy[0] * cu;
        rv = wy[0] * cv;
#pragma unroll
        for (int c = 1; c < 4; ++c) {
            coef(base, a, b, c, cu, cv);
            ru = fma(wy[c], cu, ru);
            rv = fma(wy[c], cv, rv);
        }
    }

    // sum over the 4 x 4 (t, x) rows of one y-column starting at lbase (time
    // and x strides ts, xs in elements) weighted by wt, wx -- the 4-lane
    // gather of one lane, from HBM/L2 or from the LDS box
    __device__ __forceinline__ static void column_sums(const P* lbase, int ts, int xs,
                                                       const R wt[4], const R wx[4], R& au,
                                                       R& av) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            R ru = 0, rv = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                R cu, cv;
                if constexpr (sizeof(R) == sizeof(double)) {
                    const P e = lbase[a * ts + b * xs];
                    cu = e.x;
                    cv = e.y;
                } else {
                    const P e = lbase[(a & ~1) * ts + b * xs];
                    cu = (a & 1) ? e.z : e.x;
                    cv = (a & 1) ? e.w : e.y;
                }
                ru = fma(wx[b], cu, ru);
                rv = fma(wx[b], cv, rv);
            }
            au = fma(wt[a], ru, au);
            av = fma(wt[a], rv, av);
        }
    }

    // (u, v) at an in-box point
    __device__ __forceinline__ void eval_in(double qt, double qx, double qy, R& su,
                                            R& sv) const {
        int it, ix, iy;
        R s0, s1, s2;
        cell(0, qt, it, s0);
        cell(1, qx, ix, s1);
        cell(2, qy, iy, s2);
        R wt[4], wx[4], wy[4];
        if constexpr (G == 4) {
            bspline_weights6(s0, wt);
            bspline_weights6(s1, wx);
            bspline_weights6(s2, wy);
        } else {
            bspline_weights(s0, wt);
            bspline_weights(s1, wx);
            bspline_weights(s2, wy);
        }
        const P* base = C + (static_cast<long>(it) * st + static_cast<long>(ix) * sx + iy);
        // explicit fma: the accumulation order and rounding do not depend on
        // how the compiler contracts each kernel instantiation
        R au = 0, av = 0;
        if constexpr (G == 1) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    R ru, rv;
                    row_sum(base, a, b, wy, ru, rv);
                    const R wab = wt[a] * wx[b];
                    au = fma(wab, ru, au);
                    av = fma(wab, rv, av);
                }
            }
        } else if constexpr (G == 4) {
            // lane g of the group takes y-coefficient g of all 16 rows, so
            // the group's four lanes read each 4-pair row in one load; the
            // y-weight is applied last, once per lane
            const int g = static_cast<int>(__lane_id() & 3);
            bool staged = false;
            if constexpr (S) staged = box.covers(it, ix, iy);
            if (staged) {
                if constexpr (S) {
                    const P* lb = box.lds + ((it - box.it0) * box.bx + (ix - box.ix0)) * box.by +
                                  (iy - box.iy0) + g;
                    column_sums(lb, box.bx * box.by, box.by, wt, wx, au, av);
                }
            } else if constexpr (kReuse) {
                const long key = static_cast<long>(it) * st + static_cast<long>(ix) * sx + iy;
                if (key != ckey) {
                    ckey = key;
                    // buffer loads: the lane's origin is the vector offset,
                    // each row's offset a scalar one (no per-load address
                    // arithmetic, and no 64-bit row pointers held in
                    // scalar registers)
                    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                        const_cast<P*>(C), 0, static_cast<int>(cbytes), 0x00020000);
                    // (time slices step the vector offset, x rows are the
                    // scalar offsets: three live scalars, not sixteen)
                    const int vo = static_cast<int>((key + g) * static_cast<long>(sizeof(P)));
                    const int tstep = (sizeof(R) == sizeof(double) ? 1 : 2) * st *
                                      static_cast<int>(sizeof(P));
                    const int xstep = sx * static_cast<int>(sizeof(P));
#pragma unroll
                    for (int a = 0; a < kRows / 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            const auto raw = __builtin_amdgcn_raw_buffer_load_b128(
                                rsrc, vo + a * tstep, b * xstep, 0);
                            __builtin_memcpy(&cst[a * 4 + b], &raw, sizeof(P));
                        }
                }
                // the same sums as column_sums over the held rows
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    R ru = 0, rv = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        R cu, cv;
                        if constexpr (sizeof(R) == sizeof(double)) {
                            const P e = cst[a * 4 + b];
                            cu = e.x;
                            cv = e.y;
                        } else {
                            const P e = cst[(a >> 1) * 4 + b];
                            cu = (a & 1) ? e.z : e.x;
                            cv = (a & 1) ? e.w : e.y;
                        }
                        ru = fma(wx[b], cu, ru);
                        rv = fma(wx[b], cv, rv);
                    }
                    au = fma(wt[a], ru, au);
                    av = fma(wt[a], rv, av);
                }
            } else {
                column_sums(base + g, st, sx, wt, wx, au, av);
            }
            const R wc = pick4(wy, g) * R(1.0 / 216.0);
            au *= wc;
            av *= wc;
        } else {
            static_assert(G == 2, "lane groups of 1, 2 or 4");
            // lane g of the group takes the t-slices a = 2g, 2g+1 (one element
            // row pair in FP32): each lane reads whole 4-element rows
            const int g = static_cast<int>(__lane_id() & (G - 1));
#pragma unroll
            for (int aa = 0; aa < 2; ++aa) {
                const int a = 2 * g + aa;
                const R wa = pick4(wt, a);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    R ru, rv;
                    row_sum(base, a, b, wy, ru, rv);
                    const R wab = wa * wx[b];
                    au = fma(wab, ru, au);
                    av = fma(wab, rv, av);
                }
            }
        }
        su = group_sum<G>(au);
        sv = group_sum<G>(av);
    }

    // value and gradient (d/dt, d/dx, d/dy) at an in-box point; every lane
    // of the group evaluates the whole stencil (rare: out-of-box queries
    // with linear extrapolation only)
    __device__ __forceinline__ void eval_grad(double qt, double qx, double qy, R& su, R& sv,
                              R g[3][2]) const {
        int it, ix, iy;
        R s0, s1, s2;
        cell(0, qt, it, s0);
        cell(1, qx, ix, s1);
        cell(2, qy, iy, s2);
        R wt[4], wx[4], wy[4], dt[4], dx[4], dy[4];
        bspline_weights(s0, wt);
        bspline_weights(s1, wx);
        bspline_weights(s2, wy);
        bspline_dweights(s0, dt);
        bspline_dweights(s1, dx);
        bspline_dweights(s2, dy);
        const P* base = C + (static_cast<long>(it) * st + static_cast<long>(ix) * sx + iy);
        R au = 0, av = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) g[d][0] = g[d][1] = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                R ru = 0, rv = 0, du = 0, dv = 0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    R cu, cv;
                    coef(base, a, b, c, cu, cv);
                    ru += wy[c] * cu;
                    rv += wy[c] * cv;
                    du += dy[c] * cu;
                    dv += dy[c] * cv;
                }
                au += wt[a] * wx[b] * ru;
                av += wt[a] * wx[b] * rv;
                g[0][0] += dt[a] * wx[b] * ru;
                g[0][1] += dt[a] * wx[b] * rv;
                g[1][0] += wt[a] * dx[b] * ru;
                g[1][1] += wt[a] * dx[b] * rv;
                g[2][0] += wt[a] * wx[b] * du;
                g[2][1] += wt[a] * wx[b] * dv;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            g[d][0] *= static_cast<R>(dinv[d]);
            g[d][1] *= static_cast<R>(dinv[d]);
        }
        su = au;
        sv = av;
    }

    __device__ __forceinline__ void eval(double qt, double qx, double qy, R& su, R& sv) const {
        const double ct = fmin(fmax(qt, lo[0]), hi[0]);
        const double cx = fmin(fmax(qx, lo[1]), hi[1]);
        const double cy = fmin(fmax(qy, lo[2]), hi[2]);
        const bool inside = (ct == qt) & (cx == qx) & (cy == qy);
        if (inside || extrap == 2) {
            eval_in(ct, cx, cy, su, sv);
        } else if (!L || extrap == 0) {
            // without the gradient path only modes 0 and 2 are dispatched
            su = 0;
            sv = 0;
        } else if constexpr (L) {
            R g[3][2];
            eval_grad(ct, cx, cy, su, sv, g);
            su += g[0][0] * static_cast<R>(qt - ct) + g[1][0] * static_cast<R>(qx - cx) +
                  g[2][0] * static_cast<R>(qy - cy);
            sv += g[0][1] * static_cast<R>(qt - ct) + g[1][1] * static_cast<R>(qx - cx) +
                  g[2][1] * static_cast<R>(qy - cy);
        }
    }

    // the x lookup coordinate of a position: the longitude wrapped into
    // [-180, 180) (spherical=1) or [0, 360) (spherical=2), else itself
    __device__ __forceinline__ double lookup_x(double y0) const {
        if (spherical == 1) return py_mod(y0 - 180, 360) - 180;
        if (spherical != 0) return py_mod(y0, 360);
        return y0;
    }

    __device__ __forceinline__ void operator()(double t, double y0, double y1, R& f0,
                                               R& f1) const {
        const double tt = p0 * t;
        // one inlined spline evaluation per call for every coordinate mode
        const double xx = lookup_x(y0);
        R su, sv;
        eval(tt, xx, y1, su, sv);
        if (spherical == 0) {
            f0 = static_cast<R>(p0 * static_cast<double>(su));
            f1 = static_cast<R>(p0 * static_cast<double>(sv));
        } else if constexpr (sizeof(R) == sizeof(double)) {
            // km/h -> degrees/h with one factor p0 * 180 / (pi r) per flow,
            // computed on the host (left to the device, the compiler
            // recomputed it, a division included, at every evaluation), so
            // the meridional component needs no division: within a few ulps
            // of dividing each by pi r
            f0 = (kdeg * su) / cos_lat(deg_to_rad(y1));
            f1 = kdeg * sv;
        } else {
            // the float form converts km/h to degrees/h in float too: the
            // velocity is already float, and the cosine of a float latitude
            // is good to ~1e-7, below the form's own rounding
            const float k = static_cast<float>(kdeg);
            f0 = (k * su) / cos_lat(static_cast<float>(y1) * static_cast<float>(kPi / 180));
            f1 = k * sv;
        }
    }
};

// Trilinear velocity on a uniform (t, x, y) grid (the reference's
// get_flow_linear_2D over interpolation.eval_linear, flows.py:418-506).
// V holds the (nt, nx, ny) samples as (u, v) pairs, y fastest.  Cell index
// floor((q - lo)/step) clamped to [0, n-2], weights q/step - index, the
// three linear stages in source order with contraction off.  Out-of-box
// queries: extrap 0 ("constant") is zero, pinned for eval_linear by the
// reference's composition fixture; 2 ("nearest") clamps the query into the
// box; 1 ("linear") continues the boundary cell's trilinear function.  The
// evaluation is FP64 in both precision forms (8 gathers of 16 B: not the
// bottleneck), rounded to R at the end.
template <class R>
struct Linear3 {
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ void prepare_factors() {}
    const double2* __restrict__ V;
    double lo[3], hi[3], step[3];
    int n[3];
    int sx, st;  // strides (in pairs) of the x and t axes
    int spherical, extrap;
    double r, p0;

    __device__ __forceinline__ void eval(double qt, double qx, double qy, double& su,
                                         double& sv) const {
#pragma clang fp contract(off)
        const double q[3] = {qt, qx, qy};
        bool out = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) out |= (q[d] < lo[d]) | (q[d] > hi[d]);
        if (out && extrap == 0) {
            su = 0.0;
            sv = 0.0;
            return;
        }
        int i[3];
        double w[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double qq = extrap == 2 ? fmin(fmax(q[d], lo[d]), hi[d]) : q[d];
            const double u = (qq - lo[d]) / step[d];
            int k = static_cast<int>(floor(u));
            k = k > n[d] - 2 ? n[d] - 2 : k;
            k = k < 0 ? 0 : k;
            i[d] = k;
            w[d] = u - k;
        }
        const double2* c = V + (static_cast<long>(i[0]) * st + static_cast<long>(i[1]) * sx + i[2]);
        const double2 a00 = c[0], a01 = c[1], a10 = c[sx], a11 = c[sx + 1];
        const double2 b00 = c[st], b01 = c[st + 1], b10 = c[st + sx], b11 = c[st + sx + 1];
        const double w1 = w[1], w2 = w[2], w0 = w[0];
        const double bu0 = (1 - w1) * ((1 - w2) * a00.x + w2 * a01.x) + w1 * ((1 - w2) * a10.x + w2 * a11.x);
        const double bv0 = (1 - w1) * ((1 - w2) * a00.y + w2 * a01.y) + w1 * ((1 - w2) * a10.y + w2 * a11.y);
        const double bu1 = (1 - w1) * ((1 - w2) * b00.x + w2 * b01.x) + w1 * ((1 - w2) * b10.x + w2 * b11.x);
        const double bv1 = (1 - w1) * ((1 - w2) * b00.y + w2 * b01.y) + w1 * ((1 - w2) * b10.y + w2 * b11.y);
        su = (1 - w0) * bu0 + w0 * bu1;
        sv = (1 - w0) * bv0 + w0 * bv1;
    }

    __device__ __forceinline__ void operator()(double t, double y0, double y1, R& f0,
                                               R& f1) const {
        const double tt = p0 * t;
        double xx = y0;
        if (spherical == 1) xx = py_mod(y0 - 180, 360) - 180;
        else if (spherical != 0) xx = py_mod(y0, 360);
        double su, sv;
        eval(tt, xx, y1, su, sv);
        if (spherical == 0) {
            f0 = static_cast<R>(p0 * su);
            f1 = static_cast<R>(p0 * sv);
        } else {
            f0 = static_cast<R>(((p0 * su) * 180) / (kPi * r * cos_lat(deg_to_rad(y1))));
            f1 = static_cast<R>(((p0 * sv) * 180) / (kPi * r));
        }
    }
};

}  // namespace nbx
