// DOP853 for a two-component state, one solve per GPU thread.
//
// Method: Dormand-Prince 8(5,3) as published by Hairer, Norsett & Wanner
// ("Solving Ordinary Differential Equations I", 2nd ed., sec. II.10), with
// the default step-size control of Hairer's dop853 code -- the integrator
// behind numbalsoda.dop853 (integration.py:169-171):
//   safety 0.9, fac1 0.333, fac2 6, beta 0, initial step from HINIT,
//   hmax = |xend - x|, last step clipped to xend, blended err5/err3 norm.
// Stiffness detection (every 1000 accepted steps) and dense output are not
// part of the flow-map path and are omitted.
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
constexpr double expo1 = 1.0 / 8.0;

}  // namespace dop

struct SolveResult {
    int idid;     // 1 ok, -2 step budget exhausted, -3 step size underflow
    int nstep;    // attempted steps
};

// Initial step size (Hairer's HINIT, order 8).
template <class Flow>
__device__ __forceinline__ double dop853_hinit(const Flow& f, double x, double y0, double y1,
                                               double posneg, double f00, double f01,
                                               double hmax, double rtol, double atol) {
    double sk0 = atol + rtol * fabs(y0);
    double sk1 = atol + rtol * fabs(y1);
    const double dnf = (f00 / sk0) * (f00 / sk0) + (f01 / sk1) * (f01 / sk1);
    const double dny = (y0 / sk0) * (y0 / sk0) + (y1 / sk1) * (y1 / sk1);
    double h = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
    h = fmin(h, hmax);
    h = posneg >= 0 ? fabs(h) : -fabs(h);
    double f10, f11;
    f(x + h, y0 + h * f00, y1 + h * f01, f10, f11);
    const double d0 = (f10 - f00) / sk0;
    const double d1 = (f11 - f01) / sk1;
    double der2 = sqrt(d0 * d0 + d1 * d1) / h;
    const double der12 = fmax(fabs(der2), sqrt(dnf));
    const double h1 = der12 <= 1.0e-15 ? fmax(1.0e-6, fabs(h) * 1.0e-3) : pow(0.01 / der12, 0.125);
    h = fmin(fmin(100 * fabs(h), h1), hmax);
    return posneg >= 0 ? h : -h;
}

// Integrate y' = f(x, y) from x to xend; (y0, y1) in/out.
template <class Flow>
__device__ SolveResult dop853_solve(const Flow& f, double x, const double xend, double& Y0,
                                    double& Y1, const double rtol, const double atol,
                                    const long long nmax) {
    using namespace dop;
    const double posneg = (xend - x) >= 0 ? 1.0 : -1.0;
    const double hmax = fabs(xend - x);
    double facold = 1.0e-4;
    bool last = false, reject = false;
    long long nstep = 0;
    int idid = 1;
    double y0 = Y0, y1 = Y1;
    double k1x, k1y;
    f(x, y0, y1, k1x, k1y);
    double h = dop853_hinit(f, x, y0, y1, posneg, k1x, k1y, hmax, rtol, atol);

    for (;;) {
        if (nstep > nmax) { idid = -2; break; }
        if (0.1 * fabs(h) <= fabs(x) * uround) { idid = -3; break; }
        if ((x + 1.01 * h - xend) * posneg > 0.0) {
            h = xend - x;
            last = true;
        }
        ++nstep;
        double k2x, k2y, k3x, k3y, k4x, k4y, k5x, k5y, k6x, k6y, k7x, k7y, k8x, k8y, k9x, k9y,
            k10x, k10y;
        f(x + c2 * h, y0 + h * a21 * k1x, y1 + h * a21 * k1y, k2x, k2y);
        f(x + c3 * h, y0 + h * (a31 * k1x + a32 * k2x), y1 + h * (a31 * k1y + a32 * k2y), k3x,
          k3y);
        f(x + c4 * h, y0 + h * (a41 * k1x + a43 * k3x), y1 + h * (a41 * k1y + a43 * k3y), k4x,
          k4y);
        f(x + c5 * h, y0 + h * (a51 * k1x + a53 * k3x + a54 * k4x),
          y1 + h * (a51 * k1y + a53 * k3y + a54 * k4y), k5x, k5y);
        f(x + c6 * h, y0 + h * (a61 * k1x + a64 * k4x + a65 * k5x),
          y1 + h * (a61 * k1y + a64 * k4y + a65 * k5y), k6x, k6y);
        f(x + c7 * h, y0 + h * (a71 * k1x + a74 * k4x + a75 * k5x + a76 * k6x),
          y1 + h * (a71 * k1y + a74 * k4y + a75 * k5y + a76 * k6y), k7x, k7y);
        f(x + c8 * h, y0 + h * (a81 * k1x + a84 * k4x + a85 * k5x + a86 * k6x + a87 * k7x),
          y1 + h * (a81 * k1y + a84 * k4y + a85 * k5y + a86 * k6y + a87 * k7y), k8x, k8y);
        f(x + c9 * h,
          y0 + h * (a91 * k1x + a94 * k4x + a95 * k5x + a96 * k6x + a97 * k7x + a98 * k8x),
          y1 + h * (a91 * k1y + a94 * k4y + a95 * k5y + a96 * k6y + a97 * k7y + a98 * k8y), k9x,
          k9y);
        f(x + c10 * h,
          y0 + h * (a101 * k1x + a104 * k4x + a105 * k5x + a106 * k6x + a107 * k7x +
                    a108 * k8x + a109 * k9x),
          y1 + h * (a101 * k1y + a104 * k4y + a105 * k5y + a106 * k6y + a107 * k7y +
                    a108 * k8y + a109 * k9y),
          k10x, k10y);
        // stage 11 reuses k2, stage 12 reuses k3 (as in Hairer's code)
        f(x + c11 * h,
          y0 + h * (a111 * k1x + a114 * k4x + a115 * k5x + a116 * k6x + a117 * k7x +
                    a118 * k8x + a119 * k9x + a1110 * k10x),
          y1 + h * (a111 * k1y + a114 * k4y + a115 * k5y + a116 * k6y + a117 * k7y +
                    a118 * k8y + a119 * k9y + a1110 * k10y),
          k2x, k2y);
        const double xph = x + h;
        f(xph,
          y0 + h * (a121 * k1x + a124 * k4x + a125 * k5x + a126 * k6x + a127 * k7x +
                    a128 * k8x + a129 * k9x + a1210 * k10x + a1211 * k2x),
          y1 + h * (a121 * k1y + a124 * k4y + a125 * k5y + a126 * k6y + a127 * k7y +
                    a128 * k8y + a129 * k9y + a1210 * k10y + a1211 * k2y),
          k3x, k3y);
        const double k4nx = b1 * k1x + b6 * k6x + b7 * k7x + b8 * k8x + b9 * k9x + b10 * k10x +
                            b11 * k2x + b12 * k3x;
        const double k4ny = b1 * k1y + b6 * k6y + b7 * k7y + b8 * k8y + b9 * k9y + b10 * k10y +
                            b11 * k2y + b12 * k3y;
        const double k5nx = y0 + h * k4nx;
        const double k5ny = y1 + h * k4ny;
        // error estimate: blended 5th/3rd order norms
        const double skx = atol + rtol * fmax(fabs(y0), fabs(k5nx));
        const double sky = atol + rtol * fmax(fabs(y1), fabs(k5ny));
        const double e2x = (k4nx - bhh1 * k1x - bhh2 * k9x - bhh3 * k3x) / skx;
        const double e2y = (k4ny - bhh1 * k1y - bhh2 * k9y - bhh3 * k3y) / sky;
        const double e5x = (er1 * k1x + er6 * k6x + er7 * k7x + er8 * k8x + er9 * k9x +
                            er10 * k10x + er11 * k2x + er12 * k3x) / skx;
        const double e5y = (er1 * k1y + er6 * k6y + er7 * k7y + er8 * k8y + er9 * k9y +
                            er10 * k10y + er11 * k2y + er12 * k3y) / sky;
        const double err2 = e2x * e2x + e2y * e2y;
        double err = e5x * e5x + e5y * e5y;
        double deno = err + 0.01 * err2;
        if (deno <= 0.0) deno = 1.0;
        err = fabs(h) * err * sqrt(1.0 / (2 * deno));
        const double fac11 = pow(err, expo1);
        const double fac = fmax(facc2, fmin(facc1, fac11 / safe));
        double hnew = h / fac;
        if (err <= 1.0) {
            facold = fmax(err, 1.0e-4);
            f(xph, k5nx, k5ny, k1x, k1y);  // FSAL: derivative at the new point
            y0 = k5nx;
            y1 = k5ny;
            x = xph;
            if (last) break;
            if (fabs(hnew) > hmax) hnew = posneg * hmax;
            if (reject) hnew = posneg * fmin(fabs(hnew), fabs(h));
            reject = false;
        } else {
            hnew = h / fmin(facc1, fac11 / safe);
            reject = true;
            last = false;
        }
        h = hnew;
    }
    (void)facold;
    Y0 = y0;
    Y1 = y1;
    return SolveResult{idid, static_cast<int>(nstep)};
}

}  // namespace nbx
