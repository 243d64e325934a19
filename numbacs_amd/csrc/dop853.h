// DOP853 for a two-component state, one solve per GPU thread (or per group
// of G cooperating lanes, see flows.h), written as an init + step state
// machine so that a persistent kernel can refill finished lanes.
//
// Method: Dormand-Prince 8(5,3) as published by Hairer, Norsett & Wanner
// ("Solving Ordinary Differential Equations I", 2nd ed., sec. II.10), with
// the default step-size control of Hairer's dop853 code -- the integrator
// behind numbalsoda.dop853 (integration.py:169-171):
//   safety 0.9, fac1 0.333, fac2 6, beta 0, initial step from HINIT,
//   hmax = |xend - x|, last step clipped to xend, blended err5/err3 norm.
// Stiffness detection (every 1000 accepted steps) and dense output are not
// part of the flow-map path and are omitted.
//
// Precision forms (R = stage precision):
//   R = double: the state is the absolute position (anchor 0), every
//               operation in the order of Hairer's code;
//   R = float : the stored stages (velocities) and the velocity evaluation
//               (spline gathers and sums) are float; the position is a double
//               anchor, moved to each accepted point, plus the step's
//               displacement h * sum(b k) formed from the float stages; time,
//               step size, the stage sums, the error norm and the step-size
//               control stay double.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace nbx {
namespace dop {

constexpr double c2 = 0.526001519587677318785587544488e-01;
constexpr double c3 = 0.789002279381515978178381316732e-01;
constexpr double c4 = 0.118350341907227396726757197510e+00;
constexpr double c5 = 0.281649658092772603273242802490e+00;
constexpr double c6 = 0.333333333333333333333333333333e+00;
constexpr double c7 = 0.25e+00;
constexpr double c8 = 0.307692307692307692307692307692e+00;
constexpr double c9 = 0.651282051282051282051282051282e+00;
constexpr double c10 = 0.6e+00;
constexpr double c11 = 0.857142857142857142857142857142e+00;

constexpr double a21 = 5.26001519587677318785587544488e-2;
constexpr double a31 = 1.97250569845378994544595329183e-2;
constexpr double a32 = 5.91751709536136983633785987549e-2;
constexpr double a41 = 2.95875854768068491816892993775e-2;
constexpr double a43 = 8.87627564304205475450678981324e-2;
constexpr double a51 = 2.41365134159266685502369798665e-1;
constexpr double a53 = -8.84549479328286085344864962717e-1;
constexpr double a54 = 9.24834003261792003115737966543e-1;
constexpr double a61 = 3.7037037037037037037037037037e-2;
constexpr double a64 = 1.70828608729473871279604482173e-1;
constexpr double a65 = 1.25467687566822425016691814123e-1;
constexpr double a71 = 3.7109375e-2;
constexpr double a74 = 1.70252211019544039314978060272e-1;
constexpr double a75 = 6.02165389804559606850219397283e-2;
constexpr double a76 = -1.7578125e-2;
constexpr double a81 = 3.70920001185047927108779319836e-2;
constexpr double a84 = 1.70383925712239993810214054705e-1;
constexpr double a85 = 1.07262030446373284651809199168e-1;
constexpr double a86 = -1.53194377486244017527936158236e-2;
constexpr double a87 = 8.27378916381402288758473766002e-3;
constexpr double a91 = 6.24110958716075717114429577812e-1;
constexpr double a94 = -3.36089262944694129406857109825e+0;
constexpr double a95 = -8.68219346841726006818189891453e-1;
constexpr double a96 = 2.75920996994467083049415600797e+1;
constexpr double a97 = 2.01540675504778934086186788979e+1;
constexpr double a98 = -4.34898841810699588477366255144e+1;
constexpr double a101 = 4.77662536438264365890433908527e-1;
constexpr double a104 = -2.48811461997166764192642586468e+0;
constexpr double a105 = -5.90290826836842996371446475743e-1;
constexpr double a106 = 2.12300514481811942347288949897e+1;
constexpr double a107 = 1.52792336328824235832596922938e+1;
constexpr double a108 = -3.32882109689848629194453265587e+1;
constexpr double a109 = -2.03312017085086261358222928593e-2;
constexpr double a111 = -9.3714243008598732571704021658e-1;
constexpr double a114 = 5.18637242884406370830023853209e+0;
constexpr double a115 = 1.09143734899672957818500254654e+0;
constexpr double a116 = -8.14978701074692612513997267357e+0;
constexpr double a117 = -1.85200656599969598641566180701e+1;
constexpr double a118 = 2.27394870993505042818970056734e+1;
constexpr double a119 = 2.49360555267965238987089396762e+0;
constexpr double a1110 = -3.0467644718982195003823669022e+0;
constexpr double a121 = 2.27331014751653820792359768449e+0;
constexpr double a124 = -1.05344954667372501984066689879e+1;
constexpr double a125 = -2.00087205822486249909675718444e+0;
constexpr double a126 = -1.79589318631187989172765950534e+1;
constexpr double a127 = 2.79488845294199600508499808837e+1;
constexpr double a128 = -2.85899827713502369474065508674e+0;
constexpr double a129 = -8.87285693353062954433549289258e+0;
constexpr double a1210 = 1.23605671757943030647266201528e+1;
constexpr double a1211 = 6.43392746015763530355970484046e-1;

constexpr double b1 = 5.42937341165687622380535766363e-2;
constexpr double b6 = 4.45031289275240888144113950566e+0;
constexpr double b7 = 1.89151789931450038304281599044e+0;
constexpr double b8 = -5.8012039600105847814672114227e+0;
constexpr double b9 = 3.1116436695781989440891606237e-1;
constexpr double b10 = -1.52160949662516078556178806805e-1;
constexpr double b11 = 2.01365400804030348374776537501e-1;
constexpr double b12 = 4.47106157277725905176885569043e-2;

constexpr double bhh1 = 0.244094488188976377952755905512e+00;
constexpr double bhh2 = 0.733846688281611857341361741547e+00;
constexpr double bhh3 = 0.220588235294117647058823529412e-01;

constexpr double er1 = 0.1312004499419488073250102996e-01;
constexpr double er6 = -0.1225156446376204440720569753e+01;
constexpr double er7 = -0.4957589496572501915214079952e+00;
constexpr double er8 = 0.1664377182454986536961530415e+01;
constexpr double er9 = -0.3503288487499736816886487290e+00;
constexpr double er10 = 0.3341791187130174790297318841e+00;
constexpr double er11 = 0.8192320648511571246570742613e-01;
constexpr double er12 = -0.2235530786388629525884427845e-01;

constexpr double uround = 2.3e-16;
constexpr double safe = 0.9;
constexpr double facc1 = 1.0 / 0.333;
constexpr double facc2 = 1.0 / 6.0;
// x^(1/8) (x >= 0) for the step-size factor err^expo1, expo1 = 1/8, and the
// initial step's (0.01/der12)^(1/8): three correctly rounded square roots,
// within about one ulp of the exact root, as pow is.  pow's inlined log/exp
// polynomials kept ~37 VGPRs of hoisted constants live through the whole
// solve (4-lane FP64 spline kernel: 221 -> 184 VGPRs; 3-8% faster across the
// flow-map kernels).
__device__ __forceinline__ double root8(double x) { return sqrt(sqrt(sqrt(x))); }

}  // namespace dop

// Per-solve state.  Position = (ax, ay) + (dx, dy).
template <class R>
struct Dop853State {
    double x, xend, h, hmax, posneg, facold;
    double ax, ay;
    R dx, dy;
    R k1x, k1y;
    int nstep;
    bool last, reject;
    int idid;  // 0 while running; 1 ok, -2 step budget exhausted, -3 step underflow

    __device__ __forceinline__ double px() const { return ax + static_cast<double>(dx); }
    __device__ __forceinline__ double py() const { return ay + static_cast<double>(dy); }
};

// Start a solve at (x0, y0) from time x to xend: first derivative and the
// initial step (Hairer's HINIT, order 8).
template <class R, class Flow>
__device__ __forceinline__ void dop853_init(const Flow& f, Dop853State<R>& s, double x,
                                            double xend, double p0x, double p0y, double rtol,
                                            double atol) {
    s.x = x;
    s.xend = xend;
    s.posneg = (xend - x) >= 0 ? 1.0 : -1.0;
    s.hmax = fabs(xend - x);
    s.facold = 1.0e-4;
    s.last = false;
    s.reject = false;
    s.nstep = 0;
    s.idid = 0;
    if constexpr (sizeof(R) == sizeof(double)) {
        s.ax = 0.0;
        s.ay = 0.0;
        s.dx = p0x;
        s.dy = p0y;
    } else {
        s.ax = p0x;
        s.ay = p0y;
        s.dx = R(0);
        s.dy = R(0);
    }
    const double y0 = s.px(), y1 = s.py();
    f(x, y0, y1, s.k1x, s.k1y);
    const double f00 = static_cast<double>(s.k1x), f01 = static_cast<double>(s.k1y);
    const double sk0 = atol + rtol * fabs(y0);
    const double sk1 = atol + rtol * fabs(y1);
    const double dnf = (f00 / sk0) * (f00 / sk0) + (f01 / sk1) * (f01 / sk1);
    const double dny = (y0 / sk0) * (y0 / sk0) + (y1 / sk1) * (y1 / sk1);
    double h = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
    h = fmin(h, s.hmax);
    h = s.posneg >= 0 ? fabs(h) : -fabs(h);
    R f10, f11;
    f(x + h, y0 + h * f00, y1 + h * f01, f10, f11);
    const double d0 = (static_cast<double>(f10) - f00) / sk0;
    const double d1 = (static_cast<double>(f11) - f01) / sk1;
    const double der2 = sqrt(d0 * d0 + d1 * d1) / h;
    const double der12 = fmax(fabs(der2), sqrt(dnf));
    const double h1 =
        der12 <= 1.0e-15 ? fmax(1.0e-6, fabs(h) * 1.0e-3) : dop::root8(0.01 / der12);
    h = fmin(fmin(100 * fabs(h), h1), s.hmax);
    s.h = s.posneg >= 0 ? h : -h;
}

// The head of one attempted step: the step budget and underflow checks and
// the clipping of the last step to xend.  Sets s.idid != 0 and returns false
// when the solve ends here; otherwise s.h is the step about to be taken.
template <class R>
__device__ __forceinline__ bool dop853_step_begin(Dop853State<R>& s, long long nmax) {
    using namespace dop;
    if (s.nstep > nmax) {
        s.idid = -2;
        return false;
    }
    if (0.1 * fabs(s.h) <= fabs(s.x) * uround) {
        s.idid = -3;
        return false;
    }
    if ((s.x + 1.01 * s.h - s.xend) * s.posneg > 0.0) {
        s.h = s.xend - s.x;
        s.last = true;
    }
    return true;
}

// The stages, error estimate and step-size control of the attempted step
// that dop853_step_begin prepared.
template <class R, class Flow>
__device__ __forceinline__ bool dop853_step_body(const Flow& f, Dop853State<R>& s, double rtol,
                                                 double atol) {
    using namespace dop;
    ++s.nstep;
    const double x = s.x, h = s.h;
    // Stage sums in double for both forms: DOP853's coefficients are large
    // and cancel (|a98| = 43.5), so float sums would cost ~1e-6 relative per
    // stage, the size of the tolerance itself.  In the float form only the
    // stored stages, the stored displacement and the velocity evaluation are
    // float.
    auto D = [](R v) { return static_cast<double>(v); };
    const double ax = s.ax, ay = s.ay;
    const double y0 = ax + D(s.dx), y1 = ay + D(s.dy);  // = s.dx, s.dy exactly for double
    const double K1x = D(s.k1x), K1y = D(s.k1y);
    R k2x, k2y, k3x, k3y, k4x, k4y, k5x, k5y, k6x, k6y, k7x, k7y, k8x, k8y, k9x, k9y, k10x, k10y;
#define NBX_K(n) const double K##n##x = D(k##n##x), K##n##y = D(k##n##y)
    f(x + c2 * h, y0 + h * a21 * K1x, y1 + h * a21 * K1y, k2x, k2y);
    NBX_K(2);
    f(x + c3 * h, y0 + h * (a31 * K1x + a32 * K2x), y1 + h * (a31 * K1y + a32 * K2y), k3x, k3y);
    NBX_K(3);
    f(x + c4 * h, y0 + h * (a41 * K1x + a43 * K3x), y1 + h * (a41 * K1y + a43 * K3y), k4x, k4y);
    NBX_K(4);
    f(x + c5 * h, y0 + h * (a51 * K1x + a53 * K3x + a54 * K4x),
      y1 + h * (a51 * K1y + a53 * K3y + a54 * K4y), k5x, k5y);
    NBX_K(5);
    f(x + c6 * h, y0 + h * (a61 * K1x + a64 * K4x + a65 * K5x),
      y1 + h * (a61 * K1y + a64 * K4y + a65 * K5y), k6x, k6y);
    NBX_K(6);
    f(x + c7 * h, y0 + h * (a71 * K1x + a74 * K4x + a75 * K5x + a76 * K6x),
      y1 + h * (a71 * K1y + a74 * K4y + a75 * K5y + a76 * K6y), k7x, k7y);
    NBX_K(7);
    f(x + c8 * h, y0 + h * (a81 * K1x + a84 * K4x + a85 * K5x + a86 * K6x + a87 * K7x),
      y1 + h * (a81 * K1y + a84 * K4y + a85 * K5y + a86 * K6y + a87 * K7y), k8x, k8y);
    NBX_K(8);
    f(x + c9 * h,
      y0 + h * (a91 * K1x + a94 * K4x + a95 * K5x + a96 * K6x + a97 * K7x + a98 * K8x),
      y1 + h * (a91 * K1y + a94 * K4y + a95 * K5y + a96 * K6y + a97 * K7y + a98 * K8y), k9x,
      k9y);
    NBX_K(9);
    f(x + c10 * h,
      y0 + h * (a101 * K1x + a104 * K4x + a105 * K5x + a106 * K6x + a107 * K7x + a108 * K8x +
                a109 * K9x),
      y1 + h * (a101 * K1y + a104 * K4y + a105 * K5y + a106 * K6y + a107 * K7y + a108 * K8y +
                a109 * K9y),
      k10x, k10y);
    NBX_K(10);
    // stage 11 reuses k2, stage 12 reuses k3 (as in Hairer's code)
    f(x + c11 * h,
      y0 + h * (a111 * K1x + a114 * K4x + a115 * K5x + a116 * K6x + a117 * K7x + a118 * K8x +
                a119 * K9x + a1110 * K10x),
      y1 + h * (a111 * K1y + a114 * K4y + a115 * K5y + a116 * K6y + a117 * K7y + a118 * K8y +
                a119 * K9y + a1110 * K10y),
      k2x, k2y);
    const double L2x = D(k2x), L2y = D(k2y);
    const double xph = x + h;
    f(xph,
      y0 + h * (a121 * K1x + a124 * K4x + a125 * K5x + a126 * K6x + a127 * K7x + a128 * K8x +
                a129 * K9x + a1210 * K10x + a1211 * L2x),
      y1 + h * (a121 * K1y + a124 * K4y + a125 * K5y + a126 * K6y + a127 * K7y + a128 * K8y +
                a129 * K9y + a1210 * K10y + a1211 * L2y),
      k3x, k3y);
    const double L3x = D(k3x), L3y = D(k3y);
#undef NBX_K
    const double k4nx = b1 * K1x + b6 * K6x + b7 * K7x + b8 * K8x + b9 * K9x + b10 * K10x +
                        b11 * L2x + b12 * L3x;
    const double k4ny = b1 * K1y + b6 * K6y + b7 * K7y + b8 * K8y + b9 * K9y + b10 * K10y +
                        b11 * L2y + b12 * L3y;
    const double pnx = y0 + h * k4nx;
    const double pny = y1 + h * k4ny;
    // error estimate: blended 5th/3rd order norms
    const double skx = atol + rtol * fmax(fabs(y0), fabs(pnx));
    const double sky = atol + rtol * fmax(fabs(y1), fabs(pny));
    const double e2x = (k4nx - bhh1 * K1x - bhh2 * K9x - bhh3 * L3x) / skx;
    const double e2y = (k4ny - bhh1 * K1y - bhh2 * K9y - bhh3 * L3y) / sky;
    const double e5x = (er1 * K1x + er6 * K6x + er7 * K7x + er8 * K8x + er9 * K9x +
                        er10 * K10x + er11 * L2x + er12 * L3x) / skx;
    const double e5y = (er1 * K1y + er6 * K6y + er7 * K7y + er8 * K8y + er9 * K9y +
                        er10 * K10y + er11 * L2y + er12 * L3y) / sky;
    const double err2 = e2x * e2x + e2y * e2y;
    double err = e5x * e5x + e5y * e5y;
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (2 * deno));
    const double fac11 = dop::root8(err);
    const double fac = fmax(facc2, fmin(facc1, fac11 / safe));
    double hnew = h / fac;
    if (err <= 1.0) {
        s.facold = fmax(err, 1.0e-4);
        // new state, then the derivative at the stored point.  The float
        // form moves its double anchor to the accepted position: a float
        // displacement from the initial point loses increments below its
        // resolution (~1e-7 of the distance travelled), and a particle whose
        // steps shrink at a velocity discontinuity (the grid edge under zero
        // extrapolation) would stop moving while time crawls on.
        if constexpr (sizeof(R) == sizeof(double)) {
            s.dx = pnx;
            s.dy = pny;
        } else {
            s.ax = pnx;
            s.ay = pny;
            s.dx = R(0);
            s.dy = R(0);
        }
        f(xph, s.px(), s.py(), s.k1x, s.k1y);
        s.x = xph;
        if (s.last) {
            s.idid = 1;
            return false;
        }
        if (fabs(hnew) > s.hmax) hnew = s.posneg * s.hmax;
        if (s.reject) hnew = s.posneg * fmin(fabs(hnew), fabs(h));
        s.reject = false;
    } else {
        hnew = h / fmin(facc1, fac11 / safe);
        s.reject = true;
        s.last = false;
    }
    s.h = hnew;
    return true;
}

// One attempted step (accepted or rejected).  Sets s.idid != 0 when the
// solve has ended; returns true while it is still running.
template <class R, class Flow>
__device__ __forceinline__ bool dop853_step(const Flow& f, Dop853State<R>& s, double rtol,
                                            double atol, long long nmax) {
    if (!dop853_step_begin(s, nmax)) return false;
    return dop853_step_body(f, s, rtol, atol);
}

}  // namespace nbx
