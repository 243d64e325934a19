/*
 * nbx_oracle.c -- CPU oracle for the FTLE hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C, IEEE-ordered (no FP contraction, no fast-math) restatement of the
 * reference algorithm:
 *   - DOP853 (Hairer, Norsett & Wanner, "Solving ODEs I", 2nd ed., sec. II.10;
 *     default step control safe=0.9, fac1=0.333, fac2=6, beta=0, HINIT start,
 *     hmax=|xend-x|, last step clipped to xend) -- the integrator that
 *     numbalsoda.dop853 wraps (integration.py:169-171);
 *   - the double-gyre and Bickley-jet right-hand sides (flows.py:1146-1158,
 *     flows.py:1182-1213);
 *   - the natural cubic B-spline prefilter and tricubic evaluator behind
 *     get_interp_arrays_2D / get_flow_2D (flows.py:9-46, flows.py:121-258);
 *   - flowmap_grid_2D (integration.py:123-182), ftle_grid_2D
 *     (diagnostics.py:21-65, utils.py:9-46, utils.py:168-189), C_eig_2D
 *     (diagnostics.py:200-244), the ridge predicate of ftle_ridge_pts /
 *     _ftle_ridges (extraction/ridges.py:9-154), numpy's linear percentile
 *     (ridges.py:48-51) and 8-connected labelling in raster order
 *     (ridges.py:217-229, scipy.ndimage.label).
 *
 * This file is the checker for the HIP product path and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never calls it.
 *
 * Parity pins (see tests/test_oracle.py):
 *   dop853 + double gyre : tests/golden/fm.npy (reference fixture) and
 *                          scipy DOP853 golden vectors (tests/golden, .npz files);
 *   bickley              : 27-point velocity literals of the reference tests
 *                          and scipy DOP853 golden vectors;
 *   spline prefilter/eval: literal arrays of the reference tests;
 *   ftle / C_eig         : ftle.npy, Cevals.npy, Cevecs.npy fixtures;
 *   ridge predicate      : ridge_pts.npy fixture;
 *   percentile           : numpy.percentile; labelling: scipy.ndimage.label.
 *   Out-of-grid spline extrapolation: "constant" is zero outside the grid
 *   box, the rule the same interpolation package's eval_linear follows
 *   (pinned by the reference's composition fixture fm_ci.npy); for the
 *   spline itself this is inferred, not pinned (no reference test leaves the
 *   grid).  "nearest" clamps the query into the box, "linear" continues
 *   with the gradient at the clamped point.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* DOP853 tableau: Hairer's published coefficients (readable copy:
 * scipy/integrate/_ivp/dop853_coefficients.py).                            */
static const double c2 = 0.526001519587677318785587544488e-01;
static const double c3 = 0.789002279381515978178381316732e-01;
static const double c4 = 0.118350341907227396726757197510e+00;
static const double c5 = 0.281649658092772603273242802490e+00;
static const double c6 = 0.333333333333333333333333333333e+00;
static const double c7 = 0.25e+00;
static const double c8 = 0.307692307692307692307692307692e+00;
static const double c9 = 0.651282051282051282051282051282e+00;
static const double c10 = 0.6e+00;
static const double c11 = 0.857142857142857142857142857142e+00;

static const double a21 = 5.26001519587677318785587544488e-2;
static const double a31 = 1.97250569845378994544595329183e-2;
static const double a32 = 5.91751709536136983633785987549e-2;
static const double a41 = 2.95875854768068491816892993775e-2;
static const double a43 = 8.87627564304205475450678981324e-2;
static const double a51 = 2.41365134159266685502369798665e-1;
static const double a53 = -8.84549479328286085344864962717e-1;
static const double a54 = 9.24834003261792003115737966543e-1;
static const double a61 = 3.7037037037037037037037037037e-2;
static const double a64 = 1.70828608729473871279604482173e-1;
static const double a65 = 1.25467687566822425016691814123e-1;
static const double a71 = 3.7109375e-2;
static const double a74 = 1.70252211019544039314978060272e-1;
static const double a75 = 6.02165389804559606850219397283e-2;
static const double a76 = -1.7578125e-2;
static const double a81 = 3.70920001185047927108779319836e-2;
static const double a84 = 1.70383925712239993810214054705e-1;
static const double a85 = 1.07262030446373284651809199168e-1;
static const double a86 = -1.53194377486244017527936158236e-2;
static const double a87 = 8.27378916381402288758473766002e-3;
static const double a91 = 6.24110958716075717114429577812e-1;
static const double a94 = -3.36089262944694129406857109825e+0;
static const double a95 = -8.68219346841726006818189891453e-1;
static const double a96 = 2.75920996994467083049415600797e+1;
static const double a97 = 2.01540675504778934086186788979e+1;
static const double a98 = -4.34898841810699588477366255144e+1;
static const double a101 = 4.77662536438264365890433908527e-1;
static const double a104 = -2.48811461997166764192642586468e+0;
static const double a105 = -5.90290826836842996371446475743e-1;
static const double a106 = 2.12300514481811942347288949897e+1;
static const double a107 = 1.52792336328824235832596922938e+1;
static const double a108 = -3.32882109689848629194453265587e+1;
static const double a109 = -2.03312017085086261358222928593e-2;
static const double a111 = -9.3714243008598732571704021658e-1;
static const double a114 = 5.18637242884406370830023853209e+0;
static const double a115 = 1.09143734899672957818500254654e+0;
static const double a116 = -8.14978701074692612513997267357e+0;
static const double a117 = -1.85200656599969598641566180701e+1;
static const double a118 = 2.27394870993505042818970056734e+1;
static const double a119 = 2.49360555267965238987089396762e+0;
static const double a1110 = -3.0467644718982195003823669022e+0;
static const double a121 = 2.27331014751653820792359768449e+0;
static const double a124 = -1.05344954667372501984066689879e+1;
static const double a125 = -2.00087205822486249909675718444e+0;
static const double a126 = -1.79589318631187989172765950534e+1;
static const double a127 = 2.79488845294199600508499808837e+1;
static const double a128 = -2.85899827713502369474065508674e+0;
static const double a129 = -8.87285693353062954433549289258e+0;
static const double a1210 = 1.23605671757943030647266201528e+1;
static const double a1211 = 6.43392746015763530355970484046e-1;

static const double b1 = 5.42937341165687622380535766363e-2;
static const double b6 = 4.45031289275240888144113950566e+0;
static const double b7 = 1.89151789931450038304281599044e+0;
static const double b8 = -5.8012039600105847814672114227e+0;
static const double b9 = 3.1116436695781989440891606237e-1;
static const double b10 = -1.52160949662516078556178806805e-1;
static const double b11 = 2.01365400804030348374776537501e-1;
static const double b12 = 4.47106157277725905176885569043e-2;

static const double bhh1 = 0.244094488188976377952755905512e+00;
static const double bhh2 = 0.733846688281611857341361741547e+00;
static const double bhh3 = 0.220588235294117647058823529412e-01;

static const double er1 = 0.1312004499419488073250102996e-01;
static const double er6 = -0.1225156446376204440720569753e+01;
static const double er7 = -0.4957589496572501915214079952e+00;
static const double er8 = 0.1664377182454986536961530415e+01;
static const double er9 = -0.3503288487499736816886487290e+00;
static const double er10 = 0.3341791187130174790297318841e+00;
static const double er11 = 0.8192320648511571246570742613e-01;
static const double er12 = -0.2235530786388629525884427845e-01;

/* ------------------------------------------------------------------------ */
/* Right-hand sides.  Signature follows the reference's cfunc ABI
 * void rhs(double t, double* y, double* dy, double* p) (flows.py:1146). */
typedef void (*orc_rhs_fn)(const void* ctx, double t, const double* y, double* dy);

/* flows.py:1152-1158 */
static void rhs_double_gyre(const void* ctx, double t, const double* y, double* dy)
{
    const double* p = (const double*)ctx;
    double tt = p[0] * t;
    double a = p[2] * sin(p[4] * tt + p[5]);
    double b = 1 - 2 * a;
    double f = a * (y[0] * y[0]) + b * y[0];
    double df = 2 * a * y[0] + b;
    dy[0] = p[0] * (-M_PI * p[1] * sin(M_PI * f) * cos(M_PI * y[1]) - p[3] * y[0]);
    dy[1] = p[0] * (M_PI * p[1] * cos(M_PI * f) * sin(M_PI * y[1]) * df - p[3] * y[1]);
}

/* flows.py:1189-1213 */
static void rhs_bickley(const void* ctx, double t, const double* y, double* dy)
{
    const double* p = (const double*)ctx;
    double tt = p[0] * t;
    double Y = y[1] / p[2];
    double ch = cosh(Y);
    double sech2 = 1 / (ch * ch);
    dy[0] = p[0] * (p[1] * sech2 + 2 * p[1] * tanh(Y) * sech2 *
                    (p[3] * cos(p[6] * (y[0] - p[9] * tt)) +
                     p[4] * cos(p[7] * (y[0] - p[10] * tt)) +
                     p[5] * cos(p[8] * (y[0] - p[11] * tt))));
    dy[1] = -p[0] * (p[1] * p[2] * sech2 *
                     (p[3] * p[6] * sin(p[6] * (y[0] - p[9] * tt)) +
                      p[4] * p[7] * sin(p[7] * (y[0] - p[10] * tt)) +
                      p[5] * p[8] * sin(p[8] * (y[0] - p[11] * tt))));
}

/* Tricubic spline and trilinear interpolant over a uniform (t, x, y) grid.
 * Out-of-box queries by extrap mode:
 *   0 "constant": 0 (the interpolation package's CONSTANT rule, pinned for
 *                 eval_linear by the composition fixture fm_ci.npy);
 *   1 "linear"  : spline: value at the clamped point plus the gradient there
 *                 times the outward offset; trilinear: the boundary cell's
 *                 trilinear function continued (weights outside [0, 1]);
 *   2 "nearest" : the query clamped into the box per axis.
 * Modes 1 and 2 are not pinned by the reference (no test leaves the grid). */
typedef struct {
    double lo[3], hi[3];
    int n[3]; /* data points per axis (coefficients are n+2) */
    const double* cu;
    const double* cv;
    int spherical;
    int extrap;
    double r;
    double p0;
} orc_spline;

/* cubic B-spline weights and their derivatives for local coordinate s in [0,1] */
static void bspline_w(double s, double w[4], double dw[4])
{
    double s2 = s * s, s3 = s2 * s, u = 1 - s;
    w[0] = u * u * u / 6.0;
    w[1] = (3 * s3 - 6 * s2 + 4) / 6.0;
    w[2] = (-3 * s3 + 3 * s2 + 3 * s + 1) / 6.0;
    w[3] = s3 / 6.0;
    if (dw) {
        dw[0] = -u * u / 2.0;
        dw[1] = (3 * s2 - 4 * s) / 2.0;
        dw[2] = (-3 * s2 + 2 * s + 1) / 2.0;
        dw[3] = s2 / 2.0;
    }
}

/* in-grid cell and local coordinate (index clamped to [0, n-2]) */
static int cell_of(double v, double lo, double hi, int n, double* s, double* dinv_out)
{
    double dinv = (n - 1) / (hi - lo);
    double u = (v - lo) * dinv;
    int i = (int)floor(u);
    if (i > n - 2) i = n - 2;
    if (i < 0) i = 0;
    *s = u - i;
    if (dinv_out) *dinv_out = dinv;
    return i;
}

/* evaluate value (and optionally gradient) of one coefficient array at an
 * in-box point */
static double spline3_eval_in(const orc_spline* g, const double* C, const double q[3], double grad[3])
{
    double s[3], dinv[3], w[3][4], dw[3][4];
    int i[3];
    for (int d = 0; d < 3; ++d) {
        i[d] = cell_of(q[d], g->lo[d], g->hi[d], g->n[d], &s[d], &dinv[d]);
        bspline_w(s[d], w[d], grad ? dw[d] : NULL);
    }
    const long sx = g->n[2] + 2, st = (long)(g->n[1] + 2) * sx;
    double v = 0, gt = 0, gx = 0, gy = 0;
    for (int a = 0; a < 4; ++a) {
        for (int b = 0; b < 4; ++b) {
            const double* row = C + (i[0] + a) * st + (i[1] + b) * sx + i[2];
            double r = 0, rd = 0;
            for (int c = 0; c < 4; ++c) {
                r += w[2][c] * row[c];
                if (grad) rd += dw[2][c] * row[c];
            }
            v += w[0][a] * w[1][b] * r;
            if (grad) {
                gt += dw[0][a] * w[1][b] * r;
                gx += w[0][a] * dw[1][b] * r;
                gy += w[0][a] * w[1][b] * rd;
            }
        }
    }
    if (grad) {
        grad[0] = gt * dinv[0];
        grad[1] = gx * dinv[1];
        grad[2] = gy * dinv[2];
    }
    return v;
}

double orc_spline3_eval(const orc_spline* g, const double* C, double t, double x, double y)
{
    double q[3] = {t, x, y}, qc[3];
    int out = 0;
    for (int d = 0; d < 3; ++d) {
        qc[d] = q[d] < g->lo[d] ? g->lo[d] : (q[d] > g->hi[d] ? g->hi[d] : q[d]);
        out |= (qc[d] != q[d]);
    }
    if (!out) return spline3_eval_in(g, C, q, NULL);
    if (g->extrap == 0) return 0.0;
    if (g->extrap == 1) {
        double gr[3];
        double v = spline3_eval_in(g, C, qc, gr);
        for (int d = 0; d < 3; ++d) v += gr[d] * (q[d] - qc[d]);
        return v;
    }
    return spline3_eval_in(g, C, qc, NULL);
}

/* Python's float modulo (result has the sign of the divisor) */
static double py_fmod(double a, double m)
{
    double r = fmod(a, m);
    if (r != 0) {
        if ((m < 0) != (r < 0)) r += m;
    } else {
        r = copysign(0.0, m);
    }
    return r;
}

/* eval_linear on a uniform 3-D grid (values (n0, n1, n2), last index
 * fastest): cell index floor((q - lo)/step) clamped to [0, n-2], weights
 * q/step - index; out-of-box queries by extrap mode (see above).          */
double orc_linear3_eval(const orc_spline* g, const double* V, double t, double x, double y)
{
    double q[3] = {t, x, y};
    int out = 0;
    for (int d = 0; d < 3; ++d) out |= (q[d] < g->lo[d] || q[d] > g->hi[d]);
    if (out && g->extrap == 0) return 0.0;
    int i[3];
    double w[3];
    for (int d = 0; d < 3; ++d) {
        double qq = q[d];
        if (g->extrap == 2) qq = qq < g->lo[d] ? g->lo[d] : (qq > g->hi[d] ? g->hi[d] : qq);
        double u = (qq - g->lo[d]) / ((g->hi[d] - g->lo[d]) / (g->n[d] - 1));
        int k = (int)floor(u);
        if (k > g->n[d] - 2) k = g->n[d] - 2;
        if (k < 0) k = 0;
        i[d] = k;
        w[d] = u - k;
    }
    const long s1 = g->n[2], s0 = (long)g->n[1] * g->n[2];
    const double* c = V + i[0] * s0 + i[1] * s1 + i[2];
    double b0 = (1 - w[1]) * ((1 - w[2]) * c[0] + w[2] * c[1]) +
                w[1] * ((1 - w[2]) * c[s1] + w[2] * c[s1 + 1]);
    c += s0;
    double b1 = (1 - w[1]) * ((1 - w[2]) * c[0] + w[2] * c[1]) +
                w[1] * ((1 - w[2]) * c[s1] + w[2] * c[s1 + 1]);
    return (1 - w[0]) * b0 + w[0] * b1;
}

/* flows.py:154-254 (spline, kind 2) and flows.py:451-506 (trilinear, kind 3) */
static void rhs_gridded(const orc_spline* g, int linear, double t, const double* y, double* dy)
{
    double p0 = g->p0;
    double tt = p0 * t;
    double xx = y[0], yy = y[1];
    if (g->spherical == 1) xx = py_fmod(y[0] - 180, 360) - 180;
    else if (g->spherical == 2) xx = py_fmod(y[0], 360);
    double su = linear ? orc_linear3_eval(g, g->cu, tt, xx, yy)
                       : orc_spline3_eval(g, g->cu, tt, xx, yy);
    double sv = linear ? orc_linear3_eval(g, g->cv, tt, xx, yy)
                       : orc_spline3_eval(g, g->cv, tt, xx, yy);
    if (g->spherical == 1 || g->spherical == 2) {
        dy[0] = ((p0 * su) * 180) / (M_PI * g->r * cos(yy * M_PI / 180));
        dy[1] = ((p0 * sv) * 180) / (M_PI * g->r);
    } else {
        dy[0] = p0 * su;
        dy[1] = p0 * sv;
    }
}

static void rhs_spline(const void* ctx, double t, const double* y, double* dy)
{
    rhs_gridded((const orc_spline*)ctx, 0, t, y, dy);
}

static void rhs_linear(const void* ctx, double t, const double* y, double* dy)
{
    rhs_gridded((const orc_spline*)ctx, 1, t, y, dy);
}

/* ------------------------------------------------------------------------ */
/* DOP853 for a 2-component state.  Returns Hairer's idid: 1 = success,
 * -2 = step budget (nmax) exhausted, -3 = step size too small.             */
static double sgn(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }

static double hinit2(orc_rhs_fn f, const void* ctx, double x, const double* y, double posneg,
                     const double* f0, double hmax, double rtol, double atol)
{
    const int iord = 8;
    double dnf = 0, dny = 0, sk, h, h1, der2 = 0, der12;
    double y1[2], f1[2];
    for (int i = 0; i < 2; ++i) {
        sk = atol + rtol * fabs(y[i]);
        dnf += (f0[i] / sk) * (f0[i] / sk);
        dny += (y[i] / sk) * (y[i] / sk);
    }
    if (dnf <= 1.0e-10 || dny <= 1.0e-10)
        h = 1.0e-6;
    else
        h = sqrt(dny / dnf) * 0.01;
    h = fmin(h, hmax);
    h = sgn(h, posneg);
    for (int i = 0; i < 2; ++i) y1[i] = y[i] + h * f0[i];
    f(ctx, x + h, y1, f1);
    for (int i = 0; i < 2; ++i) {
        sk = atol + rtol * fabs(y[i]);
        double d = (f1[i] - f0[i]) / sk;
        der2 += d * d;
    }
    der2 = sqrt(der2) / h;
    der12 = fmax(fabs(der2), sqrt(dnf));
    if (der12 <= 1.0e-15)
        h1 = fmax(1.0e-6, fabs(h) * 1.0e-3);
    else
        h1 = pow(0.01 / der12, 1.0 / iord);
    h = fmin(fmin(100 * fabs(h), h1), hmax);
    return sgn(h, posneg);
}

int orc_dop853_2d(orc_rhs_fn f, const void* ctx, double x, double xend, double* y, double rtol,
                  double atol, long nmax, long* nstep_out, long* naccpt_out)
{
    const double uround = 2.3e-16, safe = 0.9, fac1 = 0.333, fac2 = 6.0, beta = 0.0;
    const double expo1 = 1.0 / 8.0 - beta * 0.2;
    const double facc1 = 1.0 / fac1, facc2 = 1.0 / fac2;
    double posneg = sgn(1.0, xend - x);
    double hmax = fabs(xend - x);
    double facold = 1.0e-4;
    int last = 0, reject = 0, idid = 1;
    long nstep = 0, naccpt = 0;
    double k1[2], k2[2], k3[2], k4[2], k5[2], k6[2], k7[2], k8[2], k9[2], k10[2], y1[2];
    double h, hnew, xph, err, err2, sk, erri, deno, fac11, fac;

    f(ctx, x, y, k1);
    h = hinit2(f, ctx, x, y, posneg, k1, hmax, rtol, atol);

    for (;;) {
        if (nstep > nmax) { idid = -2; break; }
        if (0.1 * fabs(h) <= fabs(x) * uround) { idid = -3; break; }
        if ((x + 1.01 * h - xend) * posneg > 0.0) {
            h = xend - x;
            last = 1;
        }
        nstep++;
        for (int i = 0; i < 2; ++i) y1[i] = y[i] + h * a21 * k1[i];
        f(ctx, x + c2 * h, y1, k2);
        for (int i = 0; i < 2; ++i) y1[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
        f(ctx, x + c3 * h, y1, k3);
        for (int i = 0; i < 2; ++i) y1[i] = y[i] + h * (a41 * k1[i] + a43 * k3[i]);
        f(ctx, x + c4 * h, y1, k4);
        for (int i = 0; i < 2; ++i) y1[i] = y[i] + h * (a51 * k1[i] + a53 * k3[i] + a54 * k4[i]);
        f(ctx, x + c5 * h, y1, k5);
        for (int i = 0; i < 2; ++i) y1[i] = y[i] + h * (a61 * k1[i] + a64 * k4[i] + a65 * k5[i]);
        f(ctx, x + c6 * h, y1, k6);
        for (int i = 0; i < 2; ++i)
            y1[i] = y[i] + h * (a71 * k1[i] + a74 * k4[i] + a75 * k5[i] + a76 * k6[i]);
        f(ctx, x + c7 * h, y1, k7);
        for (int i = 0; i < 2; ++i)
            y1[i] = y[i] + h * (a81 * k1[i] + a84 * k4[i] + a85 * k5[i] + a86 * k6[i] + a87 * k7[i]);
        f(ctx, x + c8 * h, y1, k8);
        for (int i = 0; i < 2; ++i)
            y1[i] = y[i] + h * (a91 * k1[i] + a94 * k4[i] + a95 * k5[i] + a96 * k6[i] +
                                a97 * k7[i] + a98 * k8[i]);
        f(ctx, x + c9 * h, y1, k9);
        for (int i = 0; i < 2; ++i)
            y1[i] = y[i] + h * (a101 * k1[i] + a104 * k4[i] + a105 * k5[i] + a106 * k6[i] +
                                a107 * k7[i] + a108 * k8[i] + a109 * k9[i]);
        f(ctx, x + c10 * h, y1, k10);
        for (int i = 0; i < 2; ++i)
            y1[i] = y[i] + h * (a111 * k1[i] + a114 * k4[i] + a115 * k5[i] + a116 * k6[i] +
                                a117 * k7[i] + a118 * k8[i] + a119 * k9[i] + a1110 * k10[i]);
        f(ctx, x + c11 * h, y1, k2);
        xph = x + h;
        for (int i = 0; i < 2; ++i)
            y1[i] = y[i] + h * (a121 * k1[i] + a124 * k4[i] + a125 * k5[i] + a126 * k6[i] +
                                a127 * k7[i] + a128 * k8[i] + a129 * k9[i] + a1210 * k10[i] +
                                a1211 * k2[i]);
        f(ctx, xph, y1, k3);
        for (int i = 0; i < 2; ++i) {
            k4[i] = b1 * k1[i] + b6 * k6[i] + b7 * k7[i] + b8 * k8[i] + b9 * k9[i] +
                    b10 * k10[i] + b11 * k2[i] + b12 * k3[i];
            k5[i] = y[i] + h * k4[i];
        }
        err = 0.0;
        err2 = 0.0;
        for (int i = 0; i < 2; ++i) {
            sk = atol + rtol * fmax(fabs(y[i]), fabs(k5[i]));
            erri = k4[i] - bhh1 * k1[i] - bhh2 * k9[i] - bhh3 * k3[i];
            err2 += (erri / sk) * (erri / sk);
            erri = er1 * k1[i] + er6 * k6[i] + er7 * k7[i] + er8 * k8[i] + er9 * k9[i] +
                   er10 * k10[i] + er11 * k2[i] + er12 * k3[i];
            err += (erri / sk) * (erri / sk);
        }
        deno = err + 0.01 * err2;
        if (deno <= 0.0) deno = 1.0;
        err = fabs(h) * err * sqrt(1.0 / (2 * deno));
        fac11 = pow(err, expo1);
        fac = fac11 / pow(facold, beta);
        fac = fmax(facc2, fmin(facc1, fac / safe));
        hnew = h / fac;
        if (err <= 1.0) {
            facold = fmax(err, 1.0e-4);
            naccpt++;
            f(ctx, xph, k5, k4);
            k1[0] = k4[0];
            k1[1] = k4[1];
            y[0] = k5[0];
            y[1] = k5[1];
            x = xph;
            if (last) break;
            if (fabs(hnew) > hmax) hnew = posneg * hmax;
            if (reject) hnew = posneg * fmin(fabs(hnew), fabs(h));
            reject = 0;
        } else {
            hnew = h / fmin(facc1, fac11 / safe);
            reject = 1;
            last = 0;
        }
        h = hnew;
    }
    if (nstep_out) *nstep_out = nstep;
    if (naccpt_out) *naccpt_out = naccpt;
    return idid;
}

/* ------------------------------------------------------------------------ */
/* Flow-map drivers.  kind: 0 = double gyre, 1 = Bickley jet, 2 = spline.   */
typedef struct {
    int kind;
    const double* params;
    orc_spline spl;
} orc_flow;

static void flow_select(const orc_flow* fl, orc_rhs_fn* f, const void** ctx)
{
    if (fl->kind == 0) { *f = rhs_double_gyre; *ctx = fl->params; }
    else if (fl->kind == 1) { *f = rhs_bickley; *ctx = fl->params; }
    else if (fl->kind == 2) { *f = rhs_spline; *ctx = &fl->spl; }
    else { *f = rhs_linear; *ctx = &fl->spl; }
}

static void flow_init(orc_flow* fl, int kind, const double* params, const double* grid9,
                      const double* cu, const double* cv, int spherical, int extrap, double r)
{
    memset(fl, 0, sizeof(*fl));
    fl->kind = kind;
    fl->params = params;
    if (kind == 2 || kind == 3) {
        for (int d = 0; d < 3; ++d) {
            fl->spl.lo[d] = grid9[3 * d + 0];
            fl->spl.hi[d] = grid9[3 * d + 1];
            fl->spl.n[d] = (int)grid9[3 * d + 2];
        }
        fl->spl.cu = cu;
        fl->spl.cv = cv;
        fl->spl.spherical = spherical;
        fl->spl.extrap = extrap;
        fl->spl.r = r;
        fl->spl.p0 = params[0];
    }
}

/* integration.py:162-172: t_eval = params[0]*linspace(t0, t0+T, 2); one
 * solve per (i, j) in 'ij' order; masked points stay 0.  status[k] holds the
 * idid of each solve (0 for masked points); nsteps[k] the attempted steps.
 * nthreads <= 0 uses the OpenMP default.                                    */
void orc_flowmap_grid_2d(int kind, const double* params, const double* grid9, const double* cu,
                         const double* cv, int spherical, int extrap, double r, double t0,
                         double T, const double* x, int nx, const double* y, int ny, double rtol,
                         double atol, const uint8_t* mask, long nmax, int row_lo, int row_hi,
                         double* out, int32_t* status, int32_t* nsteps, int nthreads)
{
    orc_flow fl;
    flow_init(&fl, kind, params, grid9, cu, cv, spherical, extrap, r);
    orc_rhs_fn f;
    const void* ctx;
    flow_select(&fl, &f, &ctx);
    const double s0 = params[0] * t0, s1 = params[0] * (t0 + T);
    if (row_hi > nx) row_hi = nx;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int i = row_lo; i < row_hi; ++i) {
        for (int j = 0; j < ny; ++j) {
            long k = (long)i * ny + j;
            if (mask && mask[k]) {
                out[2 * k] = 0.0;
                out[2 * k + 1] = 0.0;
                if (status) status[k] = 0;
                if (nsteps) nsteps[k] = 0;
                continue;
            }
            double yy[2] = {x[i], y[j]};
            long ns = 0, na = 0;
            int idid = orc_dop853_2d(f, ctx, s0, s1, yy, rtol, atol, nmax, &ns, &na);
            out[2 * k] = yy[0];
            out[2 * k + 1] = yy[1];
            if (status) status[k] = idid;
            if (nsteps) nsteps[k] = (int32_t)ns;
        }
    }
}

/* integration.py:7-61: flow map of a list of points pts (npts, 2). */
void orc_flowmap_pts(int kind, const double* params, const double* grid9, const double* cu,
                     const double* cv, int spherical, int extrap, double r, double t0, double T,
                     const double* pts, long npts, double rtol, double atol, long nmax,
                     double* out, int32_t* status, int nthreads)
{
    orc_flow fl;
    flow_init(&fl, kind, params, grid9, cu, cv, spherical, extrap, r);
    orc_rhs_fn f;
    const void* ctx;
    flow_select(&fl, &f, &ctx);
    const double s0 = params[0] * t0, s1 = params[0] * (t0 + T);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long k = 0; k < npts; ++k) {
        double yy[2] = {pts[2 * k], pts[2 * k + 1]};
        int idid = orc_dop853_2d(f, ctx, s0, s1, yy, rtol, atol, nmax, NULL, NULL);
        out[2 * k] = yy[0];
        out[2 * k + 1] = yy[1];
        if (status) status[k] = idid;
    }
}

/* flowmap_aux_grid_2D (integration.py:249-464): n_aux = 5 particles per
 * grid point (offsets (h,0), (-h,0), (0,h), (0,-h), (0,0)) when eig_main,
 * else the first 4.  compute_edge: edge points integrate only the centre
 * particle (eig_main) or all 4 (not eig_main); without compute_edge only
 * interior points are integrated.  Skipped and masked particles stay 0.
 * out: (nx, ny, n_aux, 2).                                                 */
void orc_flowmap_aux_grid_2d(int kind, const double* params, const double* grid9,
                             const double* cu, const double* cv, int spherical, int extrap,
                             double r, double t0, double T, const double* x, int nx,
                             const double* y, int ny, double h, int eig_main, int compute_edge,
                             double rtol, double atol, const uint8_t* mask, long nmax,
                             double* out, int nthreads)
{
    orc_flow fl;
    flow_init(&fl, kind, params, grid9, cu, cv, spherical, extrap, r);
    orc_rhs_fn f;
    const void* ctx;
    flow_select(&fl, &f, &ctx);
    const double s0 = params[0] * t0, s1 = params[0] * (t0 + T);
    const int naux = eig_main ? 5 : 4;
    const double off[5][2] = {{h, 0.0}, {-h, 0.0}, {0.0, h}, {0.0, -h}, {0.0, 0.0}};
    memset(out, 0, sizeof(double) * (size_t)nx * ny * naux * 2);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            long c = (long)i * ny + j;
            if (mask && mask[c]) continue;
            int edge = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
            if (edge && !compute_edge) continue;
            for (int k = 0; k < naux; ++k) {
                if (edge && eig_main && k != 4) continue;
                double yy[2] = {x[i] + off[k][0], y[j] + off[k][1]};
                orc_dop853_2d(f, ctx, s0, s1, yy, rtol, atol, nmax, NULL, NULL);
                out[2 * (c * naux + k)] = yy[0];
                out[2 * (c * naux + k) + 1] = yy[1];
            }
        }
    }
}

static void eig2_sym(double a, double b, double d, double ev[2], double V[2][2]);

/* gradF_aux_stencil_2D (utils.py:50-84) at cell c of F_aux (n_aux pairs) */
static void grad_aux(const double* F, long c, int naux, double h, double g[4])
{
    const double* p = F + 2 * c * naux;
    g[0] = (p[0] - p[2]) / (2 * h);  /* dFx/dx */
    g[1] = (p[4] - p[6]) / (2 * h);  /* dFx/dy */
    g[2] = (p[1] - p[3]) / (2 * h);  /* dFy/dx */
    g[3] = (p[5] - p[7]) / (2 * h);  /* dFy/dy */
}

/* C_tensor_2D (diagnostics.py:67-110): (C11, C12, C22) on i, j in [2, n-3] */
void orc_C_tensor_aux_2d(const double* F, int nx, int ny, int naux, double h,
                         const uint8_t* mask, double* C)
{
    memset(C, 0, sizeof(double) * (size_t)nx * ny * 3);
    for (int i = 2; i < nx - 2; ++i) {
        for (int j = 2; j < ny - 2; ++j) {
            long c = (long)i * ny + j;
            if (mask && mask[c]) continue;
            double g[4];
            grad_aux(F, c, naux, h, g);
            C[3 * c] = g[0] * g[0] + g[2] * g[2];
            C[3 * c + 1] = g[0] * g[1] + g[2] * g[3];
            C[3 * c + 2] = g[1] * g[1] + g[3] * g[3];
        }
    }
}

/* C_eig_aux_2D (diagnostics.py:113-197).  eig_main: on i, j in [2, n-3]
 * the eigenvectors of the aux-grid tensor and the eigenvalues of the
 * main-grid tensor (gradF_main_stencil_2D, utils.py:88-124, central
 * differences of the centre particles); otherwise on i, j in [1, n-2] both
 * from the aux-grid tensor.                                                */
void orc_C_eig_aux_2d(const double* F, int nx, int ny, int naux, double dx, double dy, double h,
                      int eig_main, const uint8_t* mask, double* evals, double* evecs)
{
    memset(evals, 0, sizeof(double) * (size_t)nx * ny * 2);
    memset(evecs, 0, sizeof(double) * (size_t)nx * ny * 4);
    const int lo = eig_main ? 2 : 1;
    for (int i = lo; i < nx - lo; ++i) {
        for (int j = lo; j < ny - lo; ++j) {
            long c = (long)i * ny + j;
            if (mask && mask[c]) continue;
            double g[4];
            grad_aux(F, c, naux, h, g);
            double ev[2], V[2][2];
            eig2_sym(g[0] * g[0] + g[2] * g[2], g[0] * g[1] + g[2] * g[3],
                     g[1] * g[1] + g[3] * g[3], ev, V);
            if (eig_main) {
                /* centre particle (index n_aux - 1) of the four neighbours */
                const int m = naux - 1;
                const double* xp = F + 2 * ((c + ny) * naux + m);
                const double* xm = F + 2 * ((c - ny) * naux + m);
                const double* yp = F + 2 * ((c + 1) * naux + m);
                const double* ym = F + 2 * ((c - 1) * naux + m);
                double dxdx = (xp[0] - xm[0]) / (2 * dx);
                double dxdy = (yp[0] - ym[0]) / (2 * dy);
                double dydx = (xp[1] - xm[1]) / (2 * dx);
                double dydy = (yp[1] - ym[1]) / (2 * dy);
                double ev2[2], V2[2][2];
                eig2_sym(dxdx * dxdx + dydx * dydx, dxdx * dxdy + dydx * dydy,
                         dxdy * dxdy + dydy * dydy, ev2, V2);
                ev[0] = ev2[0];
                ev[1] = ev2[1];
            }
            evals[2 * c] = ev[0];
            evals[2 * c + 1] = ev[1];
            evecs[4 * c + 0] = V[0][0];
            evecs[4 * c + 1] = V[0][1];
            evecs[4 * c + 2] = V[1][0];
            evecs[4 * c + 3] = V[1][1];
        }
    }
}

/* eval_linear on a uniform 2-D grid with CONSTANT extrapolation: 0 outside
 * the closed box [x0, x1] x [y0, y1] (the rule pinned by the reference's
 * composition fixture fm_ci.npy), bilinear inside with the cell index
 * clamped to [0, n-2].  vals: (nx, ny) at stride vstride doubles.          */
static double bilinear0(const double* vals, long vstride, int nx, int ny, const double* g4,
                        double qx, double qy)
{
    if (qx < g4[0] || qx > g4[1] || qy < g4[2] || qy > g4[3]) return 0.0;
    double ux = (qx - g4[0]) / ((g4[1] - g4[0]) / (nx - 1));
    double uy = (qy - g4[2]) / ((g4[3] - g4[2]) / (ny - 1));
    int i = (int)floor(ux), j = (int)floor(uy);
    if (i > nx - 2) i = nx - 2;
    if (j > ny - 2) j = ny - 2;
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    double lx = ux - i, ly = uy - j;
    double v00 = vals[vstride * ((long)i * ny + j)], v01 = vals[vstride * ((long)i * ny + j + 1)];
    double v10 = vals[vstride * ((long)(i + 1) * ny + j)];
    double v11 = vals[vstride * ((long)(i + 1) * ny + j + 1)];
    return (1 - lx) * ((1 - ly) * v00 + ly * v01) + lx * ((1 - ly) * v10 + ly * v11);
}

/* flowmap_composition (integration.py:609-644): the positions after the
 * first interval pass through the interpolated flow maps 1 .. nT-2, then
 * the last one (flowmaps[nT-1]; for nT = 1 that is flowmaps[0] again).
 * flowmaps: (nT, nx, ny, 2); g4 = (x0, x1, y0, y1); out: (nx, ny, 2).      */
void orc_flowmap_composition(const double* flowmaps, int nT, int nx, int ny, const double* g4,
                             double* out)
{
    const long n = (long)nx * ny;
    for (long p = 0; p < n; ++p) {
        double px = flowmaps[2 * p], py = flowmaps[2 * p + 1];
        for (int k = 1; k < nT - 1; ++k) {
            const double* F = flowmaps + 2 * n * k;
            double fx = bilinear0(F, 2, nx, ny, g4, px, py);
            double fy = bilinear0(F + 1, 2, nx, ny, g4, px, py);
            px = fx;
            py = fy;
        }
        const double* L = flowmaps + 2 * n * (nT - 1);
        out[2 * p] = bilinear0(L, 2, nx, ny, g4, px, py);
        out[2 * p + 1] = bilinear0(L + 1, 2, nx, ny, g4, px, py);
    }
}

/* velocity sampler (for checking the right-hand sides against literals) */
void orc_rhs_eval(int kind, const double* params, const double* grid9, const double* cu,
                  const double* cv, int spherical, int extrap, double r, const double* t,
                  const double* pts, long n, double* out)
{
    orc_flow fl;
    flow_init(&fl, kind, params, grid9, cu, cv, spherical, extrap, r);
    orc_rhs_fn f;
    const void* ctx;
    flow_select(&fl, &f, &ctx);
    for (long k = 0; k < n; ++k) f(ctx, t[k], pts + 2 * k, out + 2 * k);
}

/* ------------------------------------------------------------------------ */
/* Natural cubic B-spline prefilter along one axis: n samples f -> n+2
 * coefficients c[-1..n] with (c[i-1] + 4 c[i] + c[i+1]) / 6 = f[i] and zero
 * second derivative at both ends (c[-1] = 2c[0]-c[1], c[n] = 2c[n-1]-c[n-2]),
 * which gives c[0] = f[0], c[n-1] = f[n-1] and a tridiagonal system for the
 * interior.  p points at c[-1]; stride in elements; work has n doubles.      */
static void prefilter_line(double* p, long stride, int n, double* work)
{
#define C(k) p[((long)(k) + 1) * stride]
    if (n == 1) {
        C(-1) = C(0);
        C(1) = C(0);
        return;
    }
    /* the samples already sit at c[0..n-1]; solve in place (Thomas) */
    if (n > 2) {
        int m = n - 2; /* unknowns c[1..n-2] */
        double* cp = work;
        /* rhs_i = 6 f_i minus the known ends */
        double prev_d = 0;
        for (int k = 0; k < m; ++k) {
            int i = k + 1;
            double rhs = 6.0 * C(i);
            if (i == 1) rhs -= C(0);
            if (i == n - 2) rhs -= C(n - 1);
            double denom = 4.0 - (k > 0 ? cp[k - 1] : 0.0);
            cp[k] = 1.0 / denom;
            double d = (rhs - (k > 0 ? prev_d : 0.0)) / denom;
            C(i) = d;
            prev_d = d;
        }
        for (int k = m - 2; k >= 0; --k) C(k + 1) = C(k + 1) - cp[k] * C(k + 2);
    }
    C(-1) = 2 * C(0) - C(1);
    C(n) = 2 * C(n - 1) - C(n - 2);
#undef C
}

/* U (nt, nx, ny) -> C (nt+2, nx+2, ny+2) */
void orc_prefilter3(const double* U, int nt, int nx, int ny, double* C)
{
    const long N2 = ny + 2, N1 = (long)(nx + 2) * N2;
    int nmax = nt > nx ? nt : nx;
    if (ny > nmax) nmax = ny;
    double* work = (double*)malloc(sizeof(double) * (nmax + 2));
    memset(C, 0, sizeof(double) * (size_t)(nt + 2) * N1);
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nx; ++b)
            for (int c = 0; c < ny; ++c)
                C[(a + 1) * N1 + (b + 1) * N2 + (c + 1)] = U[((long)a * nx + b) * ny + c];
    /* axis 2 (y), for interior t and x planes */
    for (int a = 1; a <= nt; ++a)
        for (int b = 1; b <= nx; ++b) prefilter_line(C + a * N1 + b * N2, 1, ny, work);
    /* axis 1 (x), for all y columns incl. pads */
    for (int a = 1; a <= nt; ++a)
        for (int c = 0; c < ny + 2; ++c) prefilter_line(C + a * N1 + c, N2, nx, work);
    /* axis 0 (t) */
    for (int b = 0; b < nx + 2; ++b)
        for (int c = 0; c < ny + 2; ++c) prefilter_line(C + b * N2 + c, N1, nt, work);
    free(work);
}

/* U (nx, ny) -> C (nx+2, ny+2) */
void orc_prefilter2(const double* U, int nx, int ny, double* C)
{
    const long N2 = ny + 2;
    int nmax = nx > ny ? nx : ny;
    double* work = (double*)malloc(sizeof(double) * (nmax + 2));
    memset(C, 0, sizeof(double) * (size_t)(nx + 2) * N2);
    for (int b = 0; b < nx; ++b)
        for (int c = 0; c < ny; ++c) C[(b + 1) * N2 + (c + 1)] = U[(long)b * ny + c];
    for (int b = 1; b <= nx; ++b) prefilter_line(C + b * N2, 1, ny, work);
    for (int c = 0; c < ny + 2; ++c) prefilter_line(C + c, N2, nx, work);
    free(work);
}

/* tricubic evaluation at n points (t, x, y) */
void orc_spline3_eval_pts(const double* grid9, const double* C, int extrap, const double* pts,
                          long n, double* out)
{
    orc_spline g;
    memset(&g, 0, sizeof(g));
    for (int d = 0; d < 3; ++d) {
        g.lo[d] = grid9[3 * d];
        g.hi[d] = grid9[3 * d + 1];
        g.n[d] = (int)grid9[3 * d + 2];
    }
    g.extrap = extrap;
    for (long k = 0; k < n; ++k)
        out[k] = orc_spline3_eval(&g, C, pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
}

void orc_linear3_eval_pts(const double* grid9, const double* V, int extrap, const double* pts,
                          long n, double* out)
{
    orc_spline g;
    memset(&g, 0, sizeof(g));
    for (int d = 0; d < 3; ++d) {
        g.lo[d] = grid9[3 * d];
        g.hi[d] = grid9[3 * d + 1];
        g.n[d] = (int)grid9[3 * d + 2];
    }
    g.extrap = extrap;
    for (long k = 0; k < n; ++k)
        out[k] = orc_linear3_eval(&g, V, pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
}

/* ------------------------------------------------------------------------ */
/* ftle_grid_2D (diagnostics.py:50-65): interior points only; borders and
 * masked points are 0; ftle = log(lmax)/(2|T|) when lmax > 1.              */
void orc_ftle_grid_2d(const double* F, int nx, int ny, double T, double dx, double dy,
                      const uint8_t* mask, double* ftle)
{
    const double scaling = 1 / (2 * fabs(T));
    memset(ftle, 0, sizeof(double) * (size_t)nx * ny);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 1; i < nx - 1; ++i) {
        for (int j = 1; j < ny - 1; ++j) {
            long k = (long)i * ny + j;
            if (mask && mask[k]) continue;
            /* utils.py:41-44 */
            double dxdx = (F[2 * (k + ny)] - F[2 * (k - ny)]) / (2 * dx);
            double dxdy = (F[2 * (k + 1)] - F[2 * (k - 1)]) / (2 * dy);
            double dydx = (F[2 * (k + ny) + 1] - F[2 * (k - ny) + 1]) / (2 * dx);
            double dydy = (F[2 * (k + 1) + 1] - F[2 * (k - 1) + 1]) / (2 * dy);
            double off = dxdx * dxdy + dydx * dydy;
            double a = dxdx * dxdx + dydx * dydx;
            double d = dxdy * dxdy + dydy * dydy;
            /* utils.py:185-189 */
            double trace = a + d;
            double disc = sqrt((a - d) * (a - d) + 4 * (off * off));
            double lmax = 0.5 * (trace + disc);
            if (lmax > 1) ftle[k] = scaling * log(lmax);
        }
    }
}

/* Symmetric 2x2 eigen-decomposition [[a, b], [b, d]]: ascending eigenvalues
 * and orthonormal eigenvectors as columns V[row][col] (the layout of
 * numpy.linalg.eigh).  Closed form: lmax = (a+d)/2 + sqrt(((a-d)/2)^2 + b^2),
 * lmin = (a d - b^2)/lmax when lmax != 0; eigenvector of lmax from the larger
 * of the two rows of (C - lmax I).                                          */
static void eig2_sym(double a, double b, double d, double ev[2], double V[2][2])
{
    double h = 0.5 * (a - d);
    double rad = sqrt(h * h + b * b);
    double m = 0.5 * (a + d);
    double lmax = m + rad;
    double lmin = m - rad;
    if (lmax != 0.0) {
        double det = a * d - b * b;
        double alt = det / lmax;
        /* the product form is accurate when |lmin| << |lmax| */
        if (fabs(lmin) < 0.5 * fabs(lmax)) lmin = alt;
    }
    ev[0] = lmin;
    ev[1] = lmax;
    double ex, ey;
    if (b == 0.0) {
        if (a > d) { ex = 1.0; ey = 0.0; }
        else { ex = 0.0; ey = 1.0; }
    } else {
        /* rows of (C - lmax I): (a - lmax, b) and (b, d - lmax) */
        double r1x = b, r1y = lmax - a;   /* orthogonal to row 1 */
        double r2x = lmax - d, r2y = b;   /* orthogonal to row 2 */
        double n1 = r1x * r1x + r1y * r1y, n2 = r2x * r2x + r2y * r2y;
        if (n1 >= n2) { double s = 1.0 / sqrt(n1); ex = r1x * s; ey = r1y * s; }
        else { double s = 1.0 / sqrt(n2); ex = r2x * s; ey = r2y * s; }
    }
    V[0][1] = ex;
    V[1][1] = ey;
    V[0][0] = -ey;
    V[1][0] = ex;
}

/* C_eig_2D (diagnostics.py:232-244): eigvals (nx,ny,2), eigvecs (nx,ny,2,2) */
void orc_C_eig_2d(const double* F, int nx, int ny, double dx, double dy, const uint8_t* mask,
                  double* evals, double* evecs)
{
    memset(evals, 0, sizeof(double) * (size_t)nx * ny * 2);
    memset(evecs, 0, sizeof(double) * (size_t)nx * ny * 4);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 1; i < nx - 1; ++i) {
        for (int j = 1; j < ny - 1; ++j) {
            long k = (long)i * ny + j;
            if (mask && mask[k]) continue;
            double dxdx = (F[2 * (k + ny)] - F[2 * (k - ny)]) / (2 * dx);
            double dxdy = (F[2 * (k + 1)] - F[2 * (k - 1)]) / (2 * dy);
            double dydx = (F[2 * (k + ny) + 1] - F[2 * (k - ny) + 1]) / (2 * dx);
            double dydy = (F[2 * (k + 1) + 1] - F[2 * (k - 1) + 1]) / (2 * dy);
            double off = dxdx * dxdy + dydx * dydy;
            double a = dxdx * dxdx + dydx * dydx;
            double d = dxdy * dxdy + dydy * dydy;
            double ev[2], V[2][2];
            eig2_sym(a, off, d, ev, V);
            evals[2 * k] = ev[0];
            evals[2 * k + 1] = ev[1];
            evecs[4 * k + 0] = V[0][0];
            evecs[4 * k + 1] = V[0][1];
            evecs[4 * k + 2] = V[1][0];
            evecs[4 * k + 3] = V[1][1];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* numpy.percentile(f, q) with the default 'linear' method: virtual index
 * (n-1)*q/100, lerp between the two order statistics with numpy's
 * two-sided formula.                                                        */
static int cmp_double(const void* pa, const void* pb)
{
    double a = *(const double*)pa, b = *(const double*)pb;
    return (a > b) - (a < b);
}

double orc_percentile(const double* f, long n, double q)
{
    double* s = (double*)malloc(sizeof(double) * n);
    memcpy(s, f, sizeof(double) * n);
    qsort(s, n, sizeof(double), cmp_double);
    double vi = (double)(n - 1) * (q / 100.0);
    double lo = floor(vi);
    long ilo = (long)lo;
    if (ilo < 0) ilo = 0;
    if (ilo > n - 1) ilo = n - 1;
    long ihi = ilo + 1 < n ? ilo + 1 : n - 1;
    double g = vi - lo;
    double a = s[ilo], b = s[ihi];
    double diff = b - a;
    double r = (g >= 0.5) ? b - diff * (1 - g) : a + diff * g;
    free(s);
    return r;
}

/* Ridge predicate of ftle_ridge_pts / _ftle_ridges (ridges.py:52-74) with a
 * given f_min.  ridge_bool (nx*ny) and pts (nx*ny*2) are fully written
 * (zeros where not a ridge point).  Returns the number of ridge points.     */
long orc_ridge_predicate(const double* f, const double* evec, const double* x, int nx,
                         const double* y, int ny, double sdd_thresh, double f_min,
                         uint8_t* ridge_bool, double* pts)
{
    const double dx = x[1] - x[0], dy = y[1] - y[0];
    long count = 0;
    memset(ridge_bool, 0, (size_t)nx * ny);
    memset(pts, 0, sizeof(double) * (size_t)nx * ny * 2);
    for (int i = 2; i < nx - 2; ++i) {
        for (int j = 2; j < ny - 2; ++j) {
            long k = (long)i * ny + j;
            double f0 = f[k];
            if (!(f0 > f_min)) continue;
            double fx = (f[k + ny] - f[k - ny]) / (2 * dx);
            double fy = (f[k + 1] - f[k - 1]) / (2 * dy);
            double fxx = (f[k + ny] - 2 * f[k] + f[k - ny]) / (dx * dx);
            double fyy = (f[k + 1] - 2 * f[k] + f[k - 1]) / (dy * dy);
            double fxy = (f[k + ny + 1] - f[k + ny - 1] - f[k - ny + 1] + f[k - ny - 1]) /
                         (4 * dx * dy);
            double ex = evec[2 * k], ey = evec[2 * k + 1];
            double c2v = ex * (fxx * ex + fxy * ey) + ey * (fxy * ex + fyy * ey);
            if (c2v < -sdd_thresh) {
                double t = -(fx * ex + fy * ey) / c2v;
                if (fabs(t * ex) <= dx / 2 && fabs(t * ey) <= dy / 2) {
                    pts[2 * k] = x[i] + t * ex;
                    pts[2 * k + 1] = y[j] + t * ey;
                    ridge_bool[k] = 1;
                    count++;
                }
            }
        }
    }
    return count;
}

/* 8-connected component labelling of a boolean (nx, ny) grid; labels are
 * numbered 1.. in raster order of each component's first pixel (the order
 * scipy.ndimage.label produces); background is 0.  Returns the count.       */
static int32_t uf_find(int32_t* par, int32_t a)
{
    while (par[a] != a) {
        par[a] = par[par[a]];
        a = par[a];
    }
    return a;
}

int32_t orc_label8(const uint8_t* b, int nx, int ny, int32_t* labels)
{
    long n = (long)nx * ny;
    int32_t* par = (int32_t*)malloc(sizeof(int32_t) * n);
    for (long k = 0; k < n; ++k) par[k] = (int32_t)k;
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            long k = (long)i * ny + j;
            if (!b[k]) continue;
            /* neighbours earlier in raster order */
            const int di[4] = {-1, -1, -1, 0};
            const int dj[4] = {-1, 0, 1, -1};
            for (int q = 0; q < 4; ++q) {
                int ii = i + di[q], jj = j + dj[q];
                if (ii < 0 || jj < 0 || jj >= ny) continue;
                long kk = (long)ii * ny + jj;
                if (!b[kk]) continue;
                int32_t ra = uf_find(par, (int32_t)k), rb = uf_find(par, (int32_t)kk);
                if (ra < rb) par[rb] = ra;
                else if (rb < ra) par[ra] = rb;
            }
        }
    }
    int32_t next = 0;
    int32_t* lab_of_root = (int32_t*)calloc(n, sizeof(int32_t));
    for (long k = 0; k < n; ++k) {
        if (!b[k]) { labels[k] = 0; continue; }
        int32_t r = uf_find(par, (int32_t)k);
        if (lab_of_root[r] == 0) lab_of_root[r] = ++next;
        labels[k] = lab_of_root[r];
    }
    free(lab_of_root);
    free(par);
    return next;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
