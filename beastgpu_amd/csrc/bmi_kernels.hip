// gfx950 kernels of the BEM assembly: pair classification, special-pair
// integrals (Sauter-Schwab, near, FP64-kept far pairs) and the far-pair panel
// kernel with its owner-computes write-back.
//
// Reference path replaced (file:line in the reference):
//   quadrule_determine_type          quadrule_determine.jl:54-77      -> k_classify
//   sort_data (compaction)           nonMainCaseQuadratures.jl:13-32  -> k_classify (atomic block ranges)
//   test_toll! + SS kernel           pulledback_integrand_gpu.jl:22-228,
//                                    SauterSchwab.jl:212-298,
//                                    Integrand_gpu.jl:58-132          -> k_special
//   doubleQuadRuleMainCaseKernel!    doubleQuadRuleMainCase.jl:60-105 -> k_panels (phase 1)
//   store_with_kernel_register!      store_with_kernel.jl:88-97       -> k_panels (phase 3, write-back)
//   make_complex                     assamble_gpu.jl:141-144          -> not needed: complex is written directly
#include "bmi_internal.h"

#include <algorithm>
#include <type_traits>

#define DTOL (2.220446049250313e-16 * 1.0e3)

// ------------------------------------------------------------------ math

// 1/sqrt(x) in FP64: the hardware estimate y (relative error below 1e-7)
// and one third-order correction: with e = 1 - x y^2,
// 1/sqrt(x) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + O(e^3)).
__device__ __forceinline__ double rsqrt64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// sin and cos of 2*pi*t (t in turns), FP64.  Reduction to a quarter turn
// (exact: t - q/4 by Sterbenz), then Taylor polynomials on |x| <= pi/4:
// sine to x^15, cosine to x^16 (truncation below 5e-17).
__device__ __forceinline__ void sincos_turns64(double t, double &s, double &c) {
    const double q = rint(4.0 * t);
    const double f = fma(q, -0.25, t);
    const double x = f * 6.283185307179586476925;
    const double x2 = x * x;
    double ps = 1.0 / 1307674368000.0;
    ps = fma(ps, -x2, 1.0 / 6227020800.0);
    ps = fma(ps, -x2, 1.0 / 39916800.0);
    ps = fma(ps, -x2, 1.0 / 362880.0);
    ps = fma(ps, -x2, 1.0 / 5040.0);
    ps = fma(ps, -x2, 1.0 / 120.0);
    ps = fma(ps, -x2, 1.0 / 6.0);
    const double sp = fma(-x * x2, ps, x);
    double pc = 1.0 / 20922789888000.0;
    pc = fma(pc, -x2, 1.0 / 87178291200.0);
    pc = fma(pc, -x2, 1.0 / 479001600.0);
    pc = fma(pc, -x2, 1.0 / 3628800.0);
    pc = fma(pc, -x2, 1.0 / 40320.0);
    pc = fma(pc, -x2, 1.0 / 720.0);
    pc = fma(pc, -x2, 1.0 / 24.0);
    pc = fma(pc, -x2, 0.5);
    const double cp = fma(-x2, pc, 1.0);
    const int qi = ((int)q) & 3;
    const double a = (qi & 1) ? cp : sp;
    const double b = (qi & 1) ? sp : cp;
    s = (qi & 2) ? -a : a;
    c = ((qi + 1) & 2) ? -b : b;
}

// sin and cos of 2*pi*u/64 (u in 64ths of a turn), FP64, with a table of the
// 64 angles 2*pi*m/64: u = m + g exactly (m = rint(u), |g| <= 1/2), the angle
// of g is x = (pi/32) g, |x| <= pi/64; sine: Taylor polynomial to x^7
// (truncation below 5e-18); cosine: 1 + g^2 p(g^2) with p the degree-2
// Chebyshev fit of (cos x - 1) / g^2 on g^2 in [0, 1/4] (computed with 60
// digits; |error| of the cosine below 2.7e-17, where Taylor to x^6 leaves
// 8.6e-16), the powers of pi/32 folded into the coefficients; then the angle
// addition formulas with tab[m mod 64] = (sin, cos).
__device__ __forceinline__ void sincos_64ths_tab(double u, const double2 *tab, double &s, double &c) {
    constexpr double W = 0.098174770424681038701957605727484465;  // pi/32
    constexpr double W2 = W * W;
    constexpr double S1 = W, S3 = -W * W2 / 6.0, S5 = W * W2 * W2 / 120.0, S7 = -W * W2 * W2 * W2 / 5040.0;
    constexpr double C2 = -0.004819142773969309, C4 = 3.870689505125998e-06, C6 = -1.2434801001097126e-09;
    const double m = rint(u);
    const double g = u - m;
    const double g2 = g * g;
    double ps = fma(g2, S7, S5);
    ps = fma(g2, ps, S3);
    ps = fma(g2, ps, S1);
    const double sx = g * ps;
    double pc = fma(g2, C6, C4);
    pc = fma(g2, pc, C2);
    const double cx = fma(g2, pc, 1.0);
    // m + 1.5 * 2^48 is exact (|m| < 2^47) with 1/16 as the unit in the last
    // place, so its low word is 16 m mod 2^32: masked, the byte offset of
    // tab[m mod 64]
    const int off = (int)__builtin_bit_cast(unsigned long long, m + 0x1.8p48) & (63 * (int)sizeof(double2));
    const double2 sc = *(const double2 *)((const char *)tab + off);
    s = fma(sc.x, cx, sc.y * sx);
    c = fma(sc.y, cx, -(sc.x * sx));
}

// Fills tab[0..63] with (sin, cos)(2 pi m / 64); threads 0..63 of the block.
__device__ __forceinline__ void fill_sincos_table(double2 *tab) {
    if (threadIdx.x < 64) {
        double s, c;
        sincospi((double)threadIdx.x / 32.0, &s, &c);
        tab[threadIdx.x] = make_double2(s, c);
    }
}

// The same table copied into LDS from the plan's copy in global memory (the
// kernels that start many work-groups: a load instead of a sine and cosine
// per work-group).
__device__ __forceinline__ void load_sincos_table(double2 *tab, const double2 *src) {
    if (threadIdx.x < 64) tab[threadIdx.x] = src[threadIdx.x];
}

// G * w for one point pair in FP64: h = w4pi * exp(-gamma R) / R, returned as
// (re, im).  TAB: sine and cosine from the 64-entry table in LDS, else from the
// quarter-turn polynomials; lossy = false drops the attenuation factor.
template <bool TAB>
__device__ __forceinline__ void green64(const OpConst &op, bool lossy, double w4pi, double dx, double dy, double dz,
                                        double &hr, double &hi, const double2 *tab) {
    const double R2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double rinv = rsqrt64(R2);
    const double R = R2 * rinv;
    double s, c;
    if (TAB) sincos_64ths_tab(R * (64.0 * op.k2pi), tab, s, c);
    else sincos_turns64(R * op.k2pi, s, c);
    double h = w4pi * rinv;
    if (lossy) h *= exp(-op.gamma_re * R);
    hr = h * c;
    hi = -h * s;
}

// Smooth remainder of G for the near-pair subtraction, times w:
// h = w4pi (exp(-gamma R) - 1 - gamma^2 R^2 / 2) / R = w4pi (-gamma + O(R^2)),
// exp(-gamma R) - 1 taken as exp(-gr R)(-2 sin^2(kR/2) - i sin kR) + expm1(-gr R)
// so that no term loses its leading digits for small R.
template <bool TAB>
__device__ __forceinline__ void green_smooth64(const OpConst &op, bool lossy, double w4pi, double dx, double dy,
                                               double dz, double &hr, double &hi, const double2 *tab) {
    const double R2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double k = op.k2pi * 6.283185307179586476925;
    if (R2 == 0.0) {
        hr = -w4pi * op.gamma_re;
        hi = -w4pi * k;
        return;
    }
    const double rinv = rsqrt64(R2);
    const double R = R2 * rinv;
    double sh, ch;
    if (TAB) sincos_64ths_tab(R * (32.0 * op.k2pi), tab, sh, ch);
    else sincos_turns64(0.5 * R * op.k2pi, sh, ch);
    const double s = 2.0 * sh * ch, c1 = -2.0 * sh * sh;  // sin kR, cos kR - 1
    double re, im;
    if (lossy) {
        const double gr = op.gamma_re, E = exp(-gr * R);
        re = fma(E, c1, expm1(-gr * R)) - 0.5 * (gr * gr - k * k) * R2;
        im = -fma(E, s, gr * k * R2);
    } else {
        re = fma(0.5 * k * k, R2, c1);
        im = -s;
    }
    const double h = w4pi * rinv;
    hr = h * re;
    hi = h * im;
}

// grad_x G factor for one point pair in FP64: grad_x G(r) = q(R) r with
// q = -(1 + gamma R) exp(-gamma R) / (4 pi R^3); returns h = 4 pi w4pi q as (re, im).
template <bool TAB>
__device__ __forceinline__ void green_grad64(const OpConst &op, bool lossy, double w4pi, double dx, double dy,
                                             double dz, double &hr, double &hi, const double2 *tab) {
    const double R2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double rinv = rsqrt64(R2);
    const double R = R2 * rinv;
    double s, c, kr;  // kr = kappa R
    if (TAB) {
        const double u = R * (64.0 * op.k2pi);
        sincos_64ths_tab(u, tab, s, c);
        kr = u * (6.283185307179586476925 / 64.0);
    } else {
        const double tt = R * op.k2pi;
        sincos_turns64(tt, s, c);
        kr = tt * 6.283185307179586476925;
    }
    double a = 1.0, h = -w4pi * rinv * rinv * rinv;
    if (lossy) {
        a = fma(op.gamma_re, R, 1.0);
        h *= exp(-op.gamma_re * R);
    }
    // (a + i kr)(c - i s)
    hr = h * fma(a, c, kr * s);
    hi = h * fma(kr, c, -(a * s));
}

// FP32 variant: phase reduced modulo one turn in FP64 first (as green32).
__device__ __forceinline__ void green_grad32(const OpConst &op, bool lossy, float w4pi, float dx, float dy, float dz,
                                             float &hr, float &hi) {
    const float R2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float rinv = __builtin_amdgcn_rsqf(R2);
    const float R = R2 * rinv;
    const float f = (float)__builtin_amdgcn_fract((double)R * op.k2pi);  // turns, 0 <= f < 1
    const float s = __builtin_amdgcn_sinf(f);
    const float c = __builtin_amdgcn_cosf(f);
    const float kr = R * (float)(op.k2pi * 6.283185307179586476925);
    float a = 1.0f, h = -w4pi * rinv * rinv * rinv;
    if (lossy) {
        a = fmaf((float)op.gamma_re, R, 1.0f);
        h *= __expf(-(float)op.gamma_re * R);
    }
    hr = h * fmaf(a, c, kr * s);
    hi = h * fmaf(kr, c, -(a * s));
}

// The far triangle rules of order 1..3 as compile-time values, equal to the
// ones the host writes into RuleConst (weights on the reference simplex sum to
// 1/2; w4pi = w_I w_J / (4 pi) in the same operation order).
__host__ __device__ constexpr int frule_n(int o) { return o == 1 ? 1 : (o == 2 ? 3 : 4); }
__host__ __device__ constexpr double frule_w(int o, int i) {
    return o == 1 ? 0.5 : (o == 2 ? 1.0 / 6.0 : (i == 0 ? -27.0 / 96.0 : 25.0 / 96.0));
}
__host__ __device__ constexpr double frule_l(int o, int i, int d) {
    return o == 1 ? 1.0 / 3.0
                  : (o == 2 ? (i == d ? 2.0 / 3.0 : 1.0 / 6.0)
                            : (i == 0 ? 1.0 / 3.0 : (i - 1 == d ? 3.0 / 5.0 : 1.0 / 5.0)));
}
#define BMI_FAST_TEST_ORDER 2
#define BMI_FAST_TRIAL_ORDER 3

// Far kernels with the compile-time rule accumulate the entries divided by
// alpha * fast_scale and multiply by it at the write-back.  The mixed EFIE
// pair (pair_mw_rule in FP32) leaves out the weight factor that all of the
// rule's point pairs share, w_I w_J / (4 pi) = (1/6)(25/96)/(4 pi) times 1
// for the trial vertices' points and -27/25 for the centroid, so that an
// evaluation does not multiply by its weight (2-3 % of the mixed EFIE far
// phase; the mixed double layer measured slower so, and keeps its weights).
template <int OP, int MIXED, int FAST>
__host__ __device__ constexpr double fast_scale() {
    return (OP == OPK_MW && MIXED && FAST)
               ? frule_w(BMI_FAST_TEST_ORDER, 0) * frule_w(BMI_FAST_TRIAL_ORDER, 1) / (4.0 * 3.14159265358979323846)
               : 1.0;
}
// the weight of point pair (I, J) of the compile-time rule: w_I w_J / (4 pi),
// or (FOLD) relative to fast_scale's factor
template <bool FOLD>
__host__ __device__ constexpr double fast_weight(int I, int J) {
    return !FOLD ? frule_w(BMI_FAST_TEST_ORDER, I) * frule_w(BMI_FAST_TRIAL_ORDER, J) /
                                (4.0 * 3.14159265358979323846)
                          : (J == 0 ? frule_w(BMI_FAST_TRIAL_ORDER, 0) / frule_w(BMI_FAST_TRIAL_ORDER, 1) : 1.0);
}

// Far rule seen by the pair integrators: FAST = 1 is the default 3 x 4 point
// rule of a lossless operator, fixed at compile time (straight-line code, no
// rule loads); FAST = 0 reads the rule and the attenuation flag at run time.
template <int FAST>
struct FarRule {
    const RuleConst &r;
    const OpConst &op;
    __device__ __forceinline__ int nt() const { return FAST ? frule_n(BMI_FAST_TEST_ORDER) : r.nt; }
    __device__ __forceinline__ int nb() const { return FAST ? frule_n(BMI_FAST_TRIAL_ORDER) : r.nb; }
    __device__ __forceinline__ bool lossy() const { return FAST ? false : op.lossy != 0; }
    __device__ __forceinline__ double w4pi(int I, int J) const {
        return FAST ? frule_w(BMI_FAST_TEST_ORDER, I) * frule_w(BMI_FAST_TRIAL_ORDER, J) / (4.0 * 3.14159265358979323846)
                    : r.w4pi[I][J];
    }
    __device__ __forceinline__ double lt(int I, int d) const {
        return FAST ? frule_l(BMI_FAST_TEST_ORDER, I, d) : r.lt[I][d];
    }
    __device__ __forceinline__ double lb(int J, int d) const {
        return FAST ? frule_l(BMI_FAST_TRIAL_ORDER, J, d) : r.lb[J][d];
    }
};

// ------------------------------------------------- moment-form epilogues

// EFIE on RWG.  With x' = x - c_K, a_i = v_i - c_K (and y', b_j on the trial
// cell), the reference-weighted sums give
//   sum w w G (x - v_i).(y - v_j) = Mxy - a_i.My - b_j.Mx + (a_i.b_j) M0
// and the divergence term 4 M0; the area factors of f_i = (x - v_i)/(2A)
// cancel against the Jacobians 2A of the point weights.
struct MomMW {
    double m0r, m0i, mxr[3], mxi[3], myr[3], myi[3], mxyr, mxyi;
};

__device__ __forceinline__ void mw_epilogue(const OpConst &op, const MomMW &m, const double *xa, const double *ya,
                                            double2 loc[9]) {
    // divergence term 4 beta M0
    const double dr = 4.0 * (op.beta_re * m.m0r - op.beta_im * m.m0i);
    const double di = 4.0 * (op.beta_re * m.m0i + op.beta_im * m.m0r);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double a0 = xa[3 * i], a1 = xa[3 * i + 1], a2 = xa[3 * i + 2];
        const double ur0 = m.mxr[0] - a0 * m.m0r, ur1 = m.mxr[1] - a1 * m.m0r, ur2 = m.mxr[2] - a2 * m.m0r;
        const double ui0 = m.mxi[0] - a0 * m.m0i, ui1 = m.mxi[1] - a1 * m.m0i, ui2 = m.mxi[2] - a2 * m.m0i;
        // chains of fused multiply-adds, one instruction per term
        const double ar = fma(-a2, m.myr[2], fma(-a1, m.myr[1], fma(-a0, m.myr[0], m.mxyr)));
        const double ai = fma(-a2, m.myi[2], fma(-a1, m.myi[1], fma(-a0, m.myi[0], m.mxyi)));
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double b0 = ya[3 * j], b1 = ya[3 * j + 1], b2 = ya[3 * j + 2];
            const double vr = fma(-b2, ur2, fma(-b1, ur1, fma(-b0, ur0, ar)));
            const double vi = fma(-b2, ui2, fma(-b1, ui1, fma(-b0, ui0, ai)));
            double2 z;
            z.x = fma(op.alpha_re, vr, fma(-op.alpha_im, vi, dr));
            z.y = fma(op.alpha_re, vi, fma(op.alpha_im, vr, di));
            loc[3 * i + j] = z;
        }
    }
}

__device__ __forceinline__ void hh_epilogue(const OpConst &op, const double *pr, const double *pi, double scale,
                                            double2 loc[9]) {
    const double ar = op.alpha_re * scale, ai = op.alpha_im * scale;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        loc[v].x = ar * pr[v] - ai * pi[v];
        loc[v].y = ar * pi[v] + ai * pr[v];
    }
}

// ------------------------------------- far pair, double layer (n x K)

// Double layer tested with rotated RWG functions, one far pair.  With
// r = d + x' - y' (d = c_K - c_L, primed points relative to the centroids),
// a_i = x' - alpha_i, b_j = y' - beta_j, t = n.y', p = n.r = n.d - t:
//   (n x a_i).(grad G x b_j) = q [p (a_i.b_j) - (n.b_j)(a_i.r)]
// = q [p x'.y' - beta_j.(p x') - alpha_i.(p y') + (alpha_i.beta_j) p
//      - t x'.r + alpha_i.(t r) + (n.beta_j) x'.r - (n.beta_j) alpha_i.r],
// so moments of h = w w' q carry the pair (eight, of which two pairs enter
// only as differences and are accumulated so); the 1/(2A) of each RWG
// shape cancels the Jacobian 2A of the point weights.  S is the arithmetic
// type (double, or float for the mixed path).  FAST: entries divided by alpha.
// Double-layer far pair with the default 3 x 4 point rules, entries divided
// by alpha.  The rules' points are fixed fractions of the vertices relative
// to the centroids (test point I: a_I / 2; trial point 0: the centroid,
// trial point J: (2/5) b_(J-1)), so the moment sums of pair_dl collapse onto
// the evaluations h_IJ themselves: with R_Ij = h_I,j+1, H0_I = h_I0,
// C_ij = a_i.b_j, A_iI = a_i.a_I, ad_i = a_i.d, nb_j = n.b_j, nd = n.d and
//   shr_I = H0_I + sum_j R_Ij,  str_I = (2/5) sum_j R_Ij nb_j,
//   pr_I = nd shr_I - str_I,  Q_iI = sum_j R_Ij C_ij,  xu_I = ad_I/2 + A_II/4
// the entries are ir_i - bB_j + C_ij D + nb_j g_i (pair_dl's notation) with
//   D = sum_I pr_I,  bB_j = sum_I (C_Ij / 2) pr_I,
//   ir_i = sum_I [nd Q_II / 5 - xu_I str_I + (ad_i + A_iI / 2) str_I - (2/5) nd Q_iI],
//   g_i = sum_I [xu_I shr_I - Q_II / 5 - (ad_i + A_iI / 2) shr_I + (2/5) Q_iI]:
// no moment accumulation per evaluation.  R is taken between the records'
// points, as computed.  S: arithmetic type; G: type of the points and vertices.
template <typename S, bool TAB, typename G>
__device__ __forceinline__ void pair_dl_rule(const OpConst &op, const G *xp, const G *xa, const double *xc,
                                             const G *yp, const G *ya, const double *yc, double2 loc[9],
                                             const double2 *tab) {
    const S dcx = (S)(xc[0] - yc[0]), dcy = (S)(xc[1] - yc[1]), dcz = (S)(xc[2] - yc[2]);
    double nx, ny, nz;  // unit normal of the test cell, (v1 - v3) x (v2 - v3) direction
    {
        const double e1x = (double)xa[0] - xa[6], e1y = (double)xa[1] - xa[7], e1z = (double)xa[2] - xa[8];
        const double e2x = (double)xa[3] - xa[6], e2y = (double)xa[4] - xa[7], e2z = (double)xa[5] - xa[8];
        nx = e1y * e2z - e1z * e2y;
        ny = e1z * e2x - e1x * e2z;
        nz = e1x * e2y - e1y * e2x;
        const double ninv = rsqrt64(nx * nx + ny * ny + nz * nz);
        nx *= ninv;
        ny *= ninv;
        nz *= ninv;
    }
    const S n0 = (S)nx, n1 = (S)ny, n2 = (S)nz;
    const S nd = fma(n2, dcz, fma(n1, dcy, n0 * dcx));
    auto dot = [](const G *u, const G *v) -> S {
        return fma((S)u[2], (S)v[2], fma((S)u[1], (S)v[1], (S)u[0] * (S)v[0]));
    };
    S C[3][3], A[3][3], ad[3], nb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i][j] = dot(xa + 3 * i, ya + 3 * j);
#pragma unroll
        for (int I = 0; I <= i; ++I) A[i][I] = A[I][i] = dot(xa + 3 * i, xa + 3 * I);
        ad[i] = fma((S)xa[3 * i + 2], dcz, fma((S)xa[3 * i + 1], dcy, (S)xa[3 * i] * dcx));
        nb[i] = fma(n2, (S)ya[3 * i + 2], fma(n1, (S)ya[3 * i + 1], n0 * (S)ya[3 * i]));
    }
    S Dr = 0, Di = 0, irr[3] = {}, iri[3] = {}, gr[3] = {}, gi[3] = {}, bBr[3] = {}, bBi[3] = {};
#pragma unroll
    for (int I = 0; I < 3; ++I) {
        const S ux = dcx + (S)xp[3 * I], uy = dcy + (S)xp[3 * I + 1], uz = dcz + (S)xp[3 * I + 2];
        S h0r, h0i, rr[3], ri[3];
#pragma unroll
        for (int J = 0; J < 4; ++J) {
            const S w = (S)fast_weight<false>(I, J);
            S hr, hi;
            if constexpr (sizeof(S) == 8) {
                green_grad64<TAB>(op, false, w, ux - (S)yp[3 * J], uy - (S)yp[3 * J + 1], uz - (S)yp[3 * J + 2], hr,
                                  hi, tab);
            } else {
                green_grad32(op, false, w, ux - (S)yp[3 * J], uy - (S)yp[3 * J + 1], uz - (S)yp[3 * J + 2], hr,
                             hi);
            }
            if (J == 0) {
                h0r = hr;
                h0i = hi;
            } else {
                rr[J - 1] = hr;
                ri[J - 1] = hi;
            }
        }
        const S shr = h0r + (rr[0] + rr[1] + rr[2]), shi = h0i + (ri[0] + ri[1] + ri[2]);
        const S str = (S)0.4 * fma(rr[2], nb[2], fma(rr[1], nb[1], rr[0] * nb[0]));
        const S sti = (S)0.4 * fma(ri[2], nb[2], fma(ri[1], nb[1], ri[0] * nb[0]));
        const S pr = fma(nd, shr, -str), pi = fma(nd, shi, -sti);
        S qr[3], qi[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            qr[i] = fma(rr[2], C[i][2], fma(rr[1], C[i][1], rr[0] * C[i][0]));
            qi[i] = fma(ri[2], C[i][2], fma(ri[1], C[i][1], ri[0] * C[i][0]));
        }
        const S xu = fma((S)0.25, A[I][I], (S)0.5 * ad[I]);
        // the parts that do not depend on i: nd Q_II / 5 - xu str and xu shr - Q_II / 5
        const S er = fma((S)0.2 * nd, qr[I], -xu * str), ei = fma((S)0.2 * nd, qi[I], -xu * sti);
        const S fr = fma(xu, shr, (S)-0.2 * qr[I]), fi = fma(xu, shi, (S)-0.2 * qi[I]);
        Dr += pr;
        Di += pi;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const S cf = fma((S)0.5, A[i][I], ad[i]);
            irr[i] += fma(cf, str, fma((S)-0.4 * nd, qr[i], er));
            iri[i] += fma(cf, sti, fma((S)-0.4 * nd, qi[i], ei));
            gr[i] += fma(-cf, shr, fma((S)0.4, qr[i], fr));
            gi[i] += fma(-cf, shi, fma((S)0.4, qi[i], fi));
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            bBr[j] = fma((S)0.5 * C[I][j], pr, bBr[j]);
            bBi[j] = fma((S)0.5 * C[I][j], pi, bBi[j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            loc[3 * i + j].x = (double)fma(nb[j], gr[i], fma(C[i][j], Dr, irr[i] - bBr[j]));
            loc[3 * i + j].y = (double)fma(nb[j], gi[i], fma(C[i][j], Di, iri[i] - bBi[j]));
        }
}

template <typename S, int FAST, bool TAB, typename G = double>
__device__ __forceinline__ void pair_dl(const OpConst &op, const RuleConst &rule_, const G *xp, const G *xa,
                                        const double *xc, const G *yp, const G *ya, const double *yc,
                                        double2 loc[9], const double2 *tab) {
    if constexpr (FAST) {
        pair_dl_rule<S, TAB, G>(op, xp, xa, xc, yp, ya, yc, loc, tab);
        return;
    }
    const FarRule<FAST> rule{rule_, op};
    const bool lossy = rule.lossy();
    const S dcx = (S)(xc[0] - yc[0]), dcy = (S)(xc[1] - yc[1]), dcz = (S)(xc[2] - yc[2]);
    // unit normal of the test cell, (v1 - v3) x (v2 - v3) direction
    double nx, ny, nz;
    {
        const double e1x = (double)xa[0] - xa[6], e1y = (double)xa[1] - xa[7], e1z = (double)xa[2] - xa[8];
        const double e2x = (double)xa[3] - xa[6], e2y = (double)xa[4] - xa[7], e2z = (double)xa[5] - xa[8];
        nx = e1y * e2z - e1z * e2y;
        ny = e1z * e2x - e1x * e2z;
        nz = e1x * e2y - e1y * e2x;
        const double ninv = rsqrt64(nx * nx + ny * ny + nz * nz);
        nx *= ninv;
        ny *= ninv;
        nz *= ninv;
    }
    const S n0 = (S)nx, n1 = (S)ny, n2 = (S)nz;
    const S nd = n0 * dcx + n1 * dcy + n2 * dcz;
    // Of the eight moments only A - E and F - C enter the entries, and in both
    // the sums of h t y' cancel: AE = sum nd (x'.h y') - (x'.u) h t and
    // FC = sum u h t - nd h y' (u = d + x')
    S AEr = 0, AEi = 0, Dr = 0, Di = 0, Gr = 0, Gi = 0;
    S Br[3] = {}, Bi[3] = {}, FCr[3] = {}, FCi[3] = {}, Hr[3] = {}, Hi[3] = {};
#pragma unroll
    for (int I = 0; I < BMI_MAX_RULE; ++I) {
        if (I >= rule.nt()) break;
        const S x0 = (S)xp[3 * I], x1 = (S)xp[3 * I + 1], x2 = (S)xp[3 * I + 2];
        const S u0 = dcx + x0, u1 = dcy + x1, u2 = dcz + x2;
        S shr = 0, shi = 0, str = 0, sti = 0;
        S syr[3] = {}, syi[3] = {};
#pragma unroll
        for (int J = 0; J < BMI_MAX_RULE; ++J) {
            if (J >= rule.nb()) break;
            const S y0 = (S)yp[3 * J], y1 = (S)yp[3 * J + 1], y2 = (S)yp[3 * J + 2];
            S hr, hi;
            if constexpr (sizeof(S) == 8) {
                green_grad64<TAB>(op, lossy, rule.w4pi(I, J), u0 - y0, u1 - y1, u2 - y2, hr, hi, tab);
            } else {
                green_grad32(op, lossy, (float)rule.w4pi(I, J), u0 - y0, u1 - y1, u2 - y2, hr, hi);
            }
            const S t = fma(n2, y2, fma(n1, y1, n0 * y0));
            shr += hr;
            shi += hi;
            str = fma(hr, t, str);
            sti = fma(hi, t, sti);
            syr[0] = fma(hr, y0, syr[0]); syr[1] = fma(hr, y1, syr[1]); syr[2] = fma(hr, y2, syr[2]);
            syi[0] = fma(hi, y0, syi[0]); syi[1] = fma(hi, y1, syi[1]); syi[2] = fma(hi, y2, syi[2]);
        }
        // sums over J of h p (p = n.d - t)
        const S pr = fma(nd, shr, -str), pi = fma(nd, shi, -sti);
        const S xu = fma(x2, u2, fma(x1, u1, x0 * u0));  // x'.(d + x')
        const S xsr = fma(x2, syr[2], fma(x1, syr[1], x0 * syr[0]));
        const S xsi = fma(x2, syi[2], fma(x1, syi[1], x0 * syi[0]));
        const S x[3] = {x0, x1, x2}, u[3] = {u0, u1, u2};
        Dr += pr;
        Di += pi;
        AEr = fma(nd, xsr, fma(-xu, str, AEr));
        AEi = fma(nd, xsi, fma(-xu, sti, AEi));
        Gr = fma(xu, shr, Gr - xsr);
        Gi = fma(xu, shi, Gi - xsi);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            Br[k] = fma(x[k], pr, Br[k]);
            Bi[k] = fma(x[k], pi, Bi[k]);
            FCr[k] = fma(u[k], str, fma(-nd, syr[k], FCr[k]));
            FCi[k] = fma(u[k], sti, fma(-nd, syi[k], FCi[k]));
            Hr[k] = fma(u[k], shr, Hr[k] - syr[k]);
            Hi[k] = fma(u[k], shi, Hi[k] - syi[k]);
        }
    }
    const S alr = (S)op.alpha_re, ali = (S)op.alpha_im;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const S a0 = (S)xa[3 * i], a1 = (S)xa[3 * i + 1], a2 = (S)xa[3 * i + 2];
        // terms that depend on i only (chains of fused multiply-adds, one
        // instruction per term)
        const S ir = fma(a2, FCr[2], fma(a1, FCr[1], fma(a0, FCr[0], AEr)));
        const S ii = fma(a2, FCi[2], fma(a1, FCi[1], fma(a0, FCi[0], AEi)));
        const S gr = Gr - fma(a2, Hr[2], fma(a1, Hr[1], a0 * Hr[0]));
        const S gi = Gi - fma(a2, Hi[2], fma(a1, Hi[1], a0 * Hi[0]));
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const S b0 = (S)ya[3 * j], b1 = (S)ya[3 * j + 1], b2 = (S)ya[3 * j + 2];
            const S ab = fma(a2, b2, fma(a1, b1, a0 * b0)), nb = fma(n2, b2, fma(n1, b1, n0 * b0));
            const S vr = fma(nb, gr, fma(ab, Dr, fma(-b2, Br[2], fma(-b1, Br[1], fma(-b0, Br[0], ir)))));
            const S vi = fma(nb, gi, fma(ab, Di, fma(-b2, Bi[2], fma(-b1, Bi[1], fma(-b0, Bi[0], ii)))));
            if constexpr (FAST) {  // divided by alpha (applied at the write-back)
                loc[3 * i + j].x = (double)vr;
                loc[3 * i + j].y = (double)vi;
            } else {
                loc[3 * i + j].x = (double)fma(alr, vr, -ali * vi);
                loc[3 * i + j].y = (double)fma(alr, vi, ali * vr);
            }
        }
    }
}

// EFIE far pair with the default 3 x 4 point rules (below).
template <typename S, bool TAB, typename G>
__device__ __forceinline__ void pair_mw_rule(const OpConst &op, const G *xp, const G *xa, const double *xc,
                                             const G *yp, const G *ya, const double *yc, double2 loc[9],
                                             const double2 *tab);

// -------------------------------------------------- far pair, FP64 rule

// Local 3x3 matrix of one far pair with the product rule
// (doubleQuadRuleMainCase.jl:83-101), FP64 throughout.
//   xp, yp: rule points relative to the cell centroids (BMI_MAX_RULE x 3)
//   xa, ya: vertices relative to the centroids (3 x 3)
// FAST (the default rule at compile time): the entries divided by alpha.
template <int OP, int FAST, bool TAB>
__device__ __forceinline__ void pair_fp64(const OpConst &op, const RuleConst &rule_, const double *xp,
                                          const double *xa, const double *xc, double xs, const double *yp,
                                          const double *ya, const double *yc, double ys, double2 loc[9],
                                          const double2 *tab) {
    if constexpr (OP == OPK_DL) {
        pair_dl<double, FAST, TAB>(op, rule_, xp, xa, xc, yp, ya, yc, loc, tab);
        return;
    }
    if constexpr (OP == OPK_MW && FAST) {
        pair_mw_rule<double, TAB, double>(op, xp, xa, xc, yp, ya, yc, loc, tab);
        return;
    }
    const FarRule<FAST> rule{rule_, op};
    const bool lossy = rule.lossy();
    const double dcx = xc[0] - yc[0], dcy = xc[1] - yc[1], dcz = xc[2] - yc[2];
    if (OP == OPK_MW) {
        MomMW m = {};
#pragma unroll
        for (int I = 0; I < BMI_MAX_RULE; ++I) {
            if (I >= rule.nt()) break;
            const double x0 = xp[3 * I], x1 = xp[3 * I + 1], x2 = xp[3 * I + 2];
            const double ux = dcx + x0, uy = dcy + x1, uz = dcz + x2;
            double sr = 0, si = 0, tr0 = 0, tr1 = 0, tr2 = 0, ti0 = 0, ti1 = 0, ti2 = 0;
#pragma unroll
            for (int J = 0; J < BMI_MAX_RULE; ++J) {
                if (J >= rule.nb()) break;
                const double y0 = yp[3 * J], y1 = yp[3 * J + 1], y2 = yp[3 * J + 2];
                double hr, hi;
                green64<TAB>(op, lossy, rule.w4pi(I, J), ux - y0, uy - y1, uz - y2, hr, hi, tab);
                sr += hr;
                si += hi;
                tr0 = fma(hr, y0, tr0); tr1 = fma(hr, y1, tr1); tr2 = fma(hr, y2, tr2);
                ti0 = fma(hi, y0, ti0); ti1 = fma(hi, y1, ti1); ti2 = fma(hi, y2, ti2);
            }
            m.m0r += sr; m.m0i += si;
            m.mxr[0] = fma(x0, sr, m.mxr[0]); m.mxr[1] = fma(x1, sr, m.mxr[1]); m.mxr[2] = fma(x2, sr, m.mxr[2]);
            m.mxi[0] = fma(x0, si, m.mxi[0]); m.mxi[1] = fma(x1, si, m.mxi[1]); m.mxi[2] = fma(x2, si, m.mxi[2]);
            m.myr[0] += tr0; m.myr[1] += tr1; m.myr[2] += tr2;
            m.myi[0] += ti0; m.myi[1] += ti1; m.myi[2] += ti2;
            m.mxyr = fma(x2, tr2, fma(x1, tr1, fma(x0, tr0, m.mxyr)));
            m.mxyi = fma(x2, ti2, fma(x1, ti1, fma(x0, ti0, m.mxyi)));
        }
        mw_epilogue(op, m, xa, ya, loc);
    } else if constexpr (FAST) {
        // default rules: trial point J > 0 has shape value 3/5 at vertex J - 1
        // and 1/5 at the others, J = 0 is the centroid; test point I has 2/3
        // at vertex I and 1/6 at the others.  So per test point
        //   v_j = h_0 / 3 + (h_1 + h_2 + h_3) / 5 + (2/5) h_(j+1)
        // and 6 sum_I lt(I, i) v_Ij = 3 v_ij + sum_I v_Ij: no product per shape
        double vr[3][3], vi[3][3], tr[3] = {}, ti[3] = {};
#pragma unroll
        for (int I = 0; I < 3; ++I) {
            const double ux = dcx + xp[3 * I], uy = dcy + xp[3 * I + 1], uz = dcz + xp[3 * I + 2];
            double hr[4], hi[4];
#pragma unroll
            for (int J = 0; J < 4; ++J)
                green64<TAB>(op, false, rule.w4pi(I, J), ux - yp[3 * J], uy - yp[3 * J + 1], uz - yp[3 * J + 2],
                             hr[J], hi[J], tab);
            const double br = fma(0.2, hr[1] + hr[2] + hr[3], hr[0] * (1.0 / 3.0));
            const double bi = fma(0.2, hi[1] + hi[2] + hi[3], hi[0] * (1.0 / 3.0));
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                vr[I][j] = fma(0.4, hr[j + 1], br);
                vi[I][j] = fma(0.4, hi[j + 1], bi);
                tr[j] += vr[I][j];
                ti[j] += vi[I][j];
            }
        }
        // divided by alpha (applied at the write-back)
        const double sc = 4.0 * xs * ys * (1.0 / 6.0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                loc[3 * i + j].x = sc * fma(3.0, vr[i][j], tr[j]);
                loc[3 * i + j].y = sc * fma(3.0, vi[i][j], ti[j]);
            }
    } else {
        double pr[9] = {}, pi[9] = {};
#pragma unroll
        for (int I = 0; I < BMI_MAX_RULE; ++I) {
            if (I >= rule.nt()) break;
            const double ux = dcx + xp[3 * I], uy = dcy + xp[3 * I + 1], uz = dcz + xp[3 * I + 2];
            double vr[3] = {}, vi[3] = {};
#pragma unroll
            for (int J = 0; J < BMI_MAX_RULE; ++J) {
                if (J >= rule.nb()) break;
                double hr, hi;
                green64<TAB>(op, lossy, rule.w4pi(I, J), ux - yp[3 * J], uy - yp[3 * J + 1], uz - yp[3 * J + 2], hr, hi, tab);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    vr[j] = fma(hr, rule.lb(J, j), vr[j]);
                    vi[j] = fma(hi, rule.lb(J, j), vi[j]);
                }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    pr[3 * i + j] = fma(rule.lt(I, i), vr[j], pr[3 * i + j]);
                    pi[3 * i + j] = fma(rule.lt(I, i), vi[j], pi[3 * i + j]);
                }
        }
        hh_epilogue(op, pr, pi, 4.0 * xs * ys, loc);
    }
}

// -------------------------------------------- far pair, mixed precision

// FP32 evaluation with the phase reduced modulo one turn (2 pi) in FP64
// before the FP32 sin/cos; per-pair sums in FP32, result widened to FP64.
__device__ __forceinline__ void green32(const OpConst &op, bool lossy, float w4pi, float dx, float dy, float dz, float &hr,
                                        float &hi) {
    const float R2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float rinv = __builtin_amdgcn_rsqf(R2);
    const float R = R2 * rinv;
    // v_sin_f32 / v_cos_f32 take the angle in turns; 0 <= f < 1
    const float f = (float)__builtin_amdgcn_fract((double)R * op.k2pi);
    const float s = __builtin_amdgcn_sinf(f);
    const float c = __builtin_amdgcn_cosf(f);
    float h = w4pi * rinv;
    if (lossy) h *= __expf(-(float)op.gamma_re * R);
    hr = h * c;
    hi = -h * s;
}

// EFIE far pair with the default 3 x 4 point rules, entries divided by alpha
// (FP32: by alpha * fast_scale).
// Relative to the centroids the rules' points are fixed fractions of the
// vertices a_i = v_i - c_K, b_j = v_j - c_L: test point I (barycentric 2/3 at
// vertex I) is a_I / 2, trial point 0 is the centroid and trial point J
// (3/5 at vertex J - 1) is (2/5) b_(J-1).  With h_IJ = w_I w_J G / (4 pi),
// C_Ij = a_I . b_j, T_I = sum_J h_IJ, R_j = sum_I h_I,j+1 and
// S = sum_I,j h_I,j+1 C_Ij, the moment form of mw_epilogue becomes
//   sum_IJ h_IJ (x_I - v_i).(y_J - v_j)
//     = S / 5 - (2/5) sum_j' R_j' C_ij' - (1/2) sum_I T_I C_Ij + C_ij M0
// (M0 = sum_I T_I): three sums per evaluation instead of eight moments, and
// an epilogue on the 3 x 3 matrix C.  The divergence term 4 beta M0 enters
// as (4 beta / alpha) M0.  R is taken between the records' points, as
// computed.  S: arithmetic type (double, or float for the mixed path, whose
// phase is reduced in FP64); G: type of the records' points and vertices.
template <typename S, bool TAB, typename G>
__device__ __forceinline__ void pair_mw_rule(const OpConst &op, const G *xp, const G *xa, const double *xc,
                                             const G *yp, const G *ya, const double *yc, double2 loc[9],
                                             const double2 *tab) {
    const S dcx = (S)(xc[0] - yc[0]), dcy = (S)(xc[1] - yc[1]), dcz = (S)(xc[2] - yc[2]);
    // FP64: the vertices relative to a centroid sum to zero, so row and
    // column 2 of C follow from the 2 x 2 block (to rounding), which keeps
    // five registers pairs free across the evaluations; FP32 takes all nine
    // products (their rounding is that of the FP32 evaluation itself)
    constexpr bool DERIVE = sizeof(S) == 8;
    S C[3][3];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (!DERIVE || (I < 2 && j < 2))
                C[I][j] = fma((S)xa[3 * I + 2], (S)ya[3 * j + 2], fma((S)xa[3 * I + 1], (S)ya[3 * j + 1],
                                                                      (S)xa[3 * I] * (S)ya[3 * j]));
    auto Cx = [&](int I, int j) -> S {
        if (!DERIVE || (I < 2 && j < 2)) return C[I][j];
        if (I < 2) return -(C[I][0] + C[I][1]);
        if (j < 2) return -(C[0][j] + C[1][j]);
        return (C[0][0] + C[0][1]) + (C[1][0] + C[1][1]);
    };
    S Tr[3], Ti[3], Rr[3] = {}, Ri[3] = {}, Sr = 0, Si = 0;
#pragma unroll
    for (int I = 0; I < 3; ++I) {
        const S ux = dcx + (S)xp[3 * I], uy = dcy + (S)xp[3 * I + 1], uz = dcz + (S)xp[3 * I + 2];
        S tr = 0, ti = 0;
#pragma unroll
        for (int J = 0; J < 4; ++J) {
            const S w = (S)fast_weight<sizeof(S) == 4>(I, J);
            S hr, hi;
            if constexpr (sizeof(S) == 8) {
                green64<TAB>(op, false, w, ux - (S)yp[3 * J], uy - (S)yp[3 * J + 1], uz - (S)yp[3 * J + 2], hr, hi,
                             tab);
            } else {
                green32(op, false, w, ux - (S)yp[3 * J], uy - (S)yp[3 * J + 1], uz - (S)yp[3 * J + 2], hr, hi);
            }
            tr += hr;
            ti += hi;
            if (J > 0) {
                const S c = Cx(I, J - 1);
                Rr[J - 1] += hr;
                Ri[J - 1] += hi;
                Sr = fma(c, hr, Sr);
                Si = fma(c, hi, Si);
            }
        }
        Tr[I] = tr;
        Ti[I] = ti;
    }
    const S M0r = Tr[0] + Tr[1] + Tr[2], M0i = Ti[0] + Ti[1] + Ti[2];
    const S bpr = (S)op.bp_re, bpi = (S)op.bp_im;
    const S dr = fma(bpr, M0r, -bpi * M0i), di = fma(bpr, M0i, bpi * M0r);
    S Cf[3][3];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int j = 0; j < 3; ++j) Cf[I][j] = Cx(I, j);
    S wr[3], wi[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        wr[j] = fma(Tr[2], Cf[2][j], fma(Tr[1], Cf[1][j], Tr[0] * Cf[0][j]));
        wi[j] = fma(Ti[2], Cf[2][j], fma(Ti[1], Cf[1][j], Ti[0] * Cf[0][j]));
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const S ur = fma(Rr[2], Cf[i][2], fma(Rr[1], Cf[i][1], Rr[0] * Cf[i][0]));
        const S ui = fma(Ri[2], Cf[i][2], fma(Ri[1], Cf[i][1], Ri[0] * Cf[i][0]));
        const S br = fma((S)0.2, Sr, fma((S)-0.4, ur, dr));
        const S bi = fma((S)0.2, Si, fma((S)-0.4, ui, di));
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            loc[3 * i + j].x = (double)fma(Cf[i][j], M0r, fma((S)-0.5, wr[j], br));
            loc[3 * i + j].y = (double)fma(Cf[i][j], M0i, fma((S)-0.5, wi[j], bi));
        }
    }
}

// G: type of the rule points and vertices (float when the caller keeps an
// FP32 copy of the geometry, double otherwise).  FAST: the entries divided by
// alpha.
template <int OP, int FAST, typename G = double>
__device__ __forceinline__ void pair_mixed(const OpConst &op, const RuleConst &rule_, const G *xp,
                                           const G *xa, const double *xc, double xs, const G *yp,
                                           const G *ya, const double *yc, double ys, double2 loc[9]) {
    if constexpr (OP == OPK_DL) {
        pair_dl<float, FAST, false, G>(op, rule_, xp, xa, xc, yp, ya, yc, loc, nullptr);
        return;
    }
    if constexpr (OP == OPK_MW && FAST) {
        pair_mw_rule<float, false, G>(op, xp, xa, xc, yp, ya, yc, loc, nullptr);
        return;
    }
    const FarRule<FAST> rule{rule_, op};
    const bool lossy = rule.lossy();
    const float dcx = (float)(xc[0] - yc[0]), dcy = (float)(xc[1] - yc[1]), dcz = (float)(xc[2] - yc[2]);
    float ypf[3 * BMI_MAX_RULE];
#pragma unroll
    for (int v = 0; v < 3 * BMI_MAX_RULE; ++v) ypf[v] = (float)yp[v];
    if (OP == OPK_MW) {
        float m0r = 0, m0i = 0, mxr[3] = {}, mxi[3] = {}, myr[3] = {}, myi[3] = {}, mxyr = 0, mxyi = 0;
#pragma unroll
        for (int I = 0; I < BMI_MAX_RULE; ++I) {
            if (I >= rule.nt()) break;
            const float x0 = (float)xp[3 * I], x1 = (float)xp[3 * I + 1], x2 = (float)xp[3 * I + 2];
            const float ux = dcx + x0, uy = dcy + x1, uz = dcz + x2;
            float sr = 0, si = 0, tr0 = 0, tr1 = 0, tr2 = 0, ti0 = 0, ti1 = 0, ti2 = 0;
#pragma unroll
            for (int J = 0; J < BMI_MAX_RULE; ++J) {
                if (J >= rule.nb()) break;
                const float y0 = ypf[3 * J], y1 = ypf[3 * J + 1], y2 = ypf[3 * J + 2];
                float hr, hi;
                green32(op, lossy, (float)rule.w4pi(I, J), ux - y0, uy - y1, uz - y2, hr, hi);
                sr += hr;
                si += hi;
                tr0 = fmaf(hr, y0, tr0); tr1 = fmaf(hr, y1, tr1); tr2 = fmaf(hr, y2, tr2);
                ti0 = fmaf(hi, y0, ti0); ti1 = fmaf(hi, y1, ti1); ti2 = fmaf(hi, y2, ti2);
            }
            m0r += sr; m0i += si;
            mxr[0] = fmaf(x0, sr, mxr[0]); mxr[1] = fmaf(x1, sr, mxr[1]); mxr[2] = fmaf(x2, sr, mxr[2]);
            mxi[0] = fmaf(x0, si, mxi[0]); mxi[1] = fmaf(x1, si, mxi[1]); mxi[2] = fmaf(x2, si, mxi[2]);
            myr[0] += tr0; myr[1] += tr1; myr[2] += tr2;
            myi[0] += ti0; myi[1] += ti1; myi[2] += ti2;
            mxyr = fmaf(x2, tr2, fmaf(x1, tr1, fmaf(x0, tr0, mxyr)));
            mxyi = fmaf(x2, ti2, fmaf(x1, ti1, fmaf(x0, ti0, mxyi)));
        }
        const float alr = (float)op.alpha_re, ali = (float)op.alpha_im;
        const float b4r = 4.0f * (float)op.beta_re, b4i = 4.0f * (float)op.beta_im;
        const float dr = b4r * m0r - b4i * m0i, di = b4r * m0i + b4i * m0r;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float a0 = (float)xa[3 * i], a1 = (float)xa[3 * i + 1], a2 = (float)xa[3 * i + 2];
            const float ur0 = mxr[0] - a0 * m0r, ur1 = mxr[1] - a1 * m0r, ur2 = mxr[2] - a2 * m0r;
            const float ui0 = mxi[0] - a0 * m0i, ui1 = mxi[1] - a1 * m0i, ui2 = mxi[2] - a2 * m0i;
            // written as chains of fused multiply-adds (one instruction per
            // term; the compiler may not reassociate the sums itself)
            const float ar = fmaf(-a2, myr[2], fmaf(-a1, myr[1], fmaf(-a0, myr[0], mxyr)));
            const float ai = fmaf(-a2, myi[2], fmaf(-a1, myi[1], fmaf(-a0, myi[0], mxyi)));
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float b0 = (float)ya[3 * j], b1 = (float)ya[3 * j + 1], b2 = (float)ya[3 * j + 2];
                const float vr = fmaf(-b2, ur2, fmaf(-b1, ur1, fmaf(-b0, ur0, ar)));
                const float vi = fmaf(-b2, ui2, fmaf(-b1, ui1, fmaf(-b0, ui0, ai)));
                loc[3 * i + j].x = (double)fmaf(alr, vr, fmaf(-ali, vi, dr));
                loc[3 * i + j].y = (double)fmaf(alr, vi, fmaf(ali, vr, di));
            }
        }
    } else if constexpr (FAST) {
        // default rules, as in pair_fp64
        float vr[3][3], vi[3][3], tr[3] = {}, ti[3] = {};
#pragma unroll
        for (int I = 0; I < 3; ++I) {
            const float ux = dcx + (float)xp[3 * I], uy = dcy + (float)xp[3 * I + 1], uz = dcz + (float)xp[3 * I + 2];
            float hr[4], hi[4];
#pragma unroll
            for (int J = 0; J < 4; ++J)
                green32(op, false, (float)rule.w4pi(I, J), ux - ypf[3 * J], uy - ypf[3 * J + 1], uz - ypf[3 * J + 2],
                        hr[J], hi[J]);
            const float br = fmaf(0.2f, hr[1] + hr[2] + hr[3], hr[0] * (1.0f / 3.0f));
            const float bi = fmaf(0.2f, hi[1] + hi[2] + hi[3], hi[0] * (1.0f / 3.0f));
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                vr[I][j] = fmaf(0.4f, hr[j + 1], br);
                vi[I][j] = fmaf(0.4f, hi[j + 1], bi);
                tr[j] += vr[I][j];
                ti[j] += vi[I][j];
            }
        }
        // divided by alpha (applied at the write-back)
        const float sc = 4.0f * (float)(xs * ys * (1.0 / 6.0));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                loc[3 * i + j].x = (double)(sc * fmaf(3.0f, vr[i][j], tr[j]));
                loc[3 * i + j].y = (double)(sc * fmaf(3.0f, vi[i][j], ti[j]));
            }
    } else {
        float pr[9] = {}, pi[9] = {};
#pragma unroll
        for (int I = 0; I < BMI_MAX_RULE; ++I) {
            if (I >= rule.nt()) break;
            const float ux = dcx + (float)xp[3 * I], uy = dcy + (float)xp[3 * I + 1], uz = dcz + (float)xp[3 * I + 2];
            float vr[3] = {}, vi[3] = {};
#pragma unroll
            for (int J = 0; J < BMI_MAX_RULE; ++J) {
                if (J >= rule.nb()) break;
                float hr, hi;
                green32(op, lossy, (float)rule.w4pi(I, J), ux - ypf[3 * J], uy - ypf[3 * J + 1], uz - ypf[3 * J + 2], hr, hi);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    vr[j] = fmaf(hr, (float)rule.lb(J, j), vr[j]);
                    vi[j] = fmaf(hi, (float)rule.lb(J, j), vi[j]);
                }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    pr[3 * i + j] = fmaf((float)rule.lt(I, i), vr[j], pr[3 * i + j]);
                    pi[3 * i + j] = fmaf((float)rule.lt(I, i), vi[j], pi[3 * i + j]);
                }
        }
        const float sc = 4.0f * (float)(xs * ys);
        const float ar = (float)op.alpha_re * sc, ai = (float)op.alpha_im * sc;
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            loc[v].x = (double)(ar * pr[v] - ai * pi[v]);
            loc[v].y = (double)(ar * pi[v] + ai * pr[v]);
        }
    }
}

// ------------------------------------------------- batched work mapping

// Problem of work unit u (a group, chunk or entry) under a prefix table
// base[0..n] of the problems' unit counts (base[n] = total).
__device__ __forceinline__ int find_problem(const int *base, int n, int u) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------- classification

// One work-group per test sub-batch p.  For each trial patch q the bounding
// spheres give a lower bound D on the distance between any cell of p and any
// cell of q; a block can hold a special pair only if D is below the largest
// distance at which a pair can touch (0), be near (dmin < 0.8 sqrt(A_trial),
// from the near test below) or, in the mixed path, be kept in FP64 (centroid
// distance < near_factor * max edge).  Blocks that pass are classified pair
// by pair (256 threads = 16 x 16 pairs), as quadrule_determine.jl:54-77:
// hits = vertex pairs closer than 1e3 eps (compared as squared distances);
// otherwise near iff dmin^2 max(16|gamma|^2, 1/A_trial) < 0.64; the mixed
// path marks far pairs with close centroids as class 5.
// A block's special pairs are written, in thread order, to a range of entries
// reserved with one atomic add (its first entry and count are recorded per
// block); culled blocks record a count of zero.
// Work-group g takes sub-batch g - gbase[i] of problem i (one problem, or
// every problem of a batch).
// split: work-groups per sub-batch, each taking a contiguous share of the
// trial patches (small problems still fill the device).
__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_classify(const ClassifyArgs *args, const int *gbase, int nprob,
                                                                int split) {
    // candidate trial patches are taken CG at a time: their cells' data is
    // loaded together, so a sub-batch waits for one round of loads per group
    constexpr int CG = 4;
    __shared__ double tv[BMI_PATCH][9], bv[CG][BMI_PATCH][9];
    __shared__ double tc[BMI_PATCH][3], bc[CG][BMI_PATCH][3];
    __shared__ double barea[CG][BMI_PATCH];
    __shared__ int tt[BMI_PATCH], bt[CG][BMI_PATCH];
    __shared__ double trad[BMI_PATCH];  // test cell: largest vertex distance from its centroid
    __shared__ int nearq[BMI_BLOCK_THREADS];
    __shared__ int wsum[CG][BMI_BLOCK_THREADS / 64];
    __shared__ int s_off;

    const int grp = (int)blockIdx.x / split, part = (int)blockIdx.x - grp * split;
    const int pi = find_problem(gbase, nprob, grp);
    const ClassifyArgs a = args[pi];  // a copy: the kernel's stores cannot alias it
    const int p = grp - gbase[pi];
    const int qshare = (a.npb + split - 1) / split;
    const int qlo = part * qshare, qhi = min(a.npb, qlo + qshare);
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    if (t < BMI_PATCH) tt[t] = a.X.patch_tri[p * BMI_PATCH + t];
    __syncthreads();
    for (int i = t; i < BMI_PATCH * 9; i += BMI_BLOCK_THREADS) {
        const int c = i / 9, e = i - 9 * (i / 9);
        tv[c][e] = tt[c] >= 0 ? a.X.vtx[9 * tt[c] + e] : 0.0;
    }
    if (t < BMI_PATCH * 3) {
        const int c = t / 3, e = t - 3 * (t / 3);
        tc[c][e] = tt[c] >= 0 ? a.X.cen[3 * tt[c] + e] : 0.0;
    }
    const double px = a.X.patch_sph[4 * p], py = a.X.patch_sph[4 * p + 1], pz = a.X.patch_sph[4 * p + 2];
    const double pr = a.X.patch_sph[4 * p + 3];
    const double dtol2 = DTOL * DTOL;
    __syncthreads();
    if (t < BMI_PATCH) {
        double r2 = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double dx = tv[t][3 * i] - tc[t][0], dy = tv[t][3 * i + 1] - tc[t][1], dz = tv[t][3 * i + 2] - tc[t][2];
            r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
        }
        trad[t] = sqrt(r2);
    }
    __syncthreads();

    for (int q0 = qlo; q0 < qhi; q0 += BMI_BLOCK_THREADS) {
        // cull this chunk of trial patches
        const int q = q0 + t;
        int cand = 0;
        if (q < qhi) {
            const double dx = px - a.Y.patch_sph[4 * q], dy = py - a.Y.patch_sph[4 * q + 1],
                         dz = pz - a.Y.patch_sph[4 * q + 2];
            const double qr = a.Y.patch_sph[4 * q + 3];
            const double D = sqrt(dx * dx + dy * dy + dz * dz) - pr - qr;
            double reach = fmax(1e-9, 0.8 * sqrt(a.Y.patch_amax[q]));
            if (a.near_d >= 0.0) reach = fmax(reach, a.near_d);
            const double lim = 1.0001 * reach + 1e-12;
            cand = D < lim;
            if (cand) {
                // then cell by cell: a pair can only be special if its test
                // cell's bounding sphere comes within reach of the patch's
                const double qx = a.Y.patch_sph[4 * q], qy = a.Y.patch_sph[4 * q + 1], qz = a.Y.patch_sph[4 * q + 2];
                cand = 0;
                for (int k = 0; k < BMI_PATCH; ++k) {
                    if (tt[k] < 0) continue;
                    const double ex = tc[k][0] - qx, ey = tc[k][1] - qy, ez = tc[k][2] - qz;
                    if (sqrt(ex * ex + ey * ey + ez * ez) - trad[k] - qr < lim) {
                        cand = 1;
                        break;
                    }
                }
            }
            if (!cand) a.blk_cnt[(long long)p * a.npb + q] = 0;
        }
        const unsigned long long m = __ballot(cand);
        if (lane == 0) wsum[0][wv] = __popcll(m);
        __syncthreads();
        int base = 0, ncand = 0;
#pragma unroll
        for (int i = 0; i < BMI_BLOCK_THREADS / 64; ++i) {
            base += i < wv ? wsum[0][i] : 0;
            ncand += wsum[0][i];
        }
        if (cand) nearq[base + __popcll(m & ((1ull << lane) - 1ull))] = q;
        __syncthreads();

        for (int u0 = 0; u0 < ncand; u0 += CG) {
          const int ng = min(CG, ncand - u0);
          if (t < CG * BMI_PATCH) {
              const int g = t / BMI_PATCH, c = t - BMI_PATCH * g;
              bt[g][c] = g < ng ? a.Y.patch_tri[nearq[u0 + g] * BMI_PATCH + c] : -1;
          }
          __syncthreads();
          for (int i = t; i < CG * BMI_PATCH * 9; i += BMI_BLOCK_THREADS) {
              const int g = i / (BMI_PATCH * 9), r = i - g * BMI_PATCH * 9;
              const int c = r / 9, e = r - 9 * (r / 9);
              const int L = bt[g][c];
              bv[g][c][e] = L >= 0 ? a.Y.vtx[9 * L + e] : 0.0;
          }
          for (int i = t; i < CG * BMI_PATCH * 3; i += BMI_BLOCK_THREADS) {
              const int g = i / (BMI_PATCH * 3), r = i - g * BMI_PATCH * 3;
              const int c = r / 3, e = r - 3 * (r / 3);
              const int L = bt[g][c];
              bc[g][c][e] = L >= 0 ? a.Y.cen[3 * L + e] : 0.0;
          }
          if (t < CG * BMI_PATCH) {
              const int g = t / BMI_PATCH, c = t - BMI_PATCH * g;
              barea[g][c] = bt[g][c] >= 0 ? a.Y.area[bt[g][c]] : 1.0;
          }
          __syncthreads();
          // the group's blocks are classified first, then their entries get
          // one range reserved with one atomic for the group (in block order)
          const int kl = t >> 4, ll = t & 15;
          const int K = tt[kl];
          int gcls[CG];
#pragma unroll
          for (int g = 0; g < CG; ++g) {
            const int L = bt[g][ll];
            int cls = -1;
            if (g < ng && K >= 0 && L >= 0) {
                int hits = 0;
                double dmin2 = 1.79769313486231570815e308;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const double dx = tv[kl][3 * i] - bv[g][ll][3 * j];
                        const double dy = tv[kl][3 * i + 1] - bv[g][ll][3 * j + 1];
                        const double dz = tv[kl][3 * i + 2] - bv[g][ll][3 * j + 2];
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        dmin2 = fmin(dmin2, d2);
                        hits += d2 < dtol2;
                    }
                if (hits > 0) {
                    cls = hits > 3 ? 3 : hits;
                } else if (dmin2 * fmax(a.k2x16, 1.0 / barea[g][ll]) < 0.2 * 0.2 * 16) {
                    cls = CLS_NEAR;
                } else {
                    cls = CLS_FAR;
                    if (a.near_d2 >= 0.0) {
                        const double cx = tc[kl][0] - bc[g][ll][0], cy = tc[kl][1] - bc[g][ll][1],
                                     cz = tc[kl][2] - bc[g][ll][2];
                        if (cx * cx + cy * cy + cz * cz < a.near_d2) cls = CLS_FAR64;
                    }
                }
            }
            gcls[g] = cls;
            const unsigned long long sm = __ballot(cls > 0);
            if (lane == 0) wsum[g][wv] = __popcll(sm);
          }
          __syncthreads();
          // block g's entries start at gstart[g] of the group's range; a
          // wave's at sbase[g] within its block
          int gstart[CG], sbase[CG], gtotal = 0;
#pragma unroll
          for (int g = 0; g < CG; ++g) {
            int tot = 0, sb = 0;
#pragma unroll
            for (int i = 0; i < BMI_BLOCK_THREADS / 64; ++i) {
                sb += i < wv ? wsum[g][i] : 0;
                tot += wsum[g][i];
            }
            gstart[g] = gtotal;
            sbase[g] = sb;
            gtotal += g < ng ? tot : 0;
          }
          // the range's place varies from run to run, its contents do not
          if (t == 0) s_off = gtotal > 0 ? atomicAdd(a.counter, gtotal) : 0;
          __syncthreads();
          const int off = s_off;
          if (t < ng) {
            const long long b = (long long)p * a.npb + nearq[u0 + t];
            int tot = 0;
#pragma unroll
            for (int i = 0; i < BMI_BLOCK_THREADS / 64; ++i) tot += wsum[t][i];
            int st = 0;
            for (int g = 0; g < t; ++g)
#pragma unroll
                for (int i = 0; i < BMI_BLOCK_THREADS / 64; ++i) st += wsum[g][i];
            a.blk_off[b] = tot > 0 ? off + st : 0;
            a.blk_cnt[b] = tot;
          }
#pragma unroll
          for (int g = 0; g < CG; ++g) {
            const int cls = gcls[g];
            const unsigned long long sm = __ballot(cls > 0);
            if (cls > 0) {
                const int e = off + gstart[g] + sbase[g] + __popcll(sm & ((1ull << lane) - 1ull));
                if (e < a.capacity) {
                    a.spec_K[e] = K;
                    a.spec_L[e] = bt[g][ll];
                    // bit 7: the pair is integrated in the test cell's home group, not here
                    a.spec_cls[e] = (unsigned char)(cls | (a.X.home[K] == p ? 0 : SPEC_NOT_OWNER));
                    a.spec_loc[e] = (unsigned char)t;
                }
            }
          }
          __syncthreads();  // wsum, s_off and the group's cells are reused by the next group
        }
    }
}

// ------------------------------------------------- plan geometry on device

// Far-rule points of every cell relative to its centroid,
// x = v3 + b1 (v1 - v3) + b2 (v2 - v3) (assamble_chunk_gpu.jl:75 chart);
// thread = (cell, point, coordinate).  Products are not fused, so the points
// equal a host evaluation of the same expression.
__global__ void k_rule_points(const double *vtx, const double *cen, long long nt, int n, double l0, double l1,
                              double l2, double l3, double l4, double l5, double l6, double l7, double *pts) {
#pragma clang fp contract(off)
    const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nt * 3 * BMI_MAX_RULE) return;
    const long long k = u / (3 * BMI_MAX_RULE);
    const int I = (int)(u % (3 * BMI_MAX_RULE)) / 3, d = (int)(u % 3);
    if (I >= n) {
        pts[u] = 0.0;
        return;
    }
    const double b1[BMI_MAX_RULE] = {l0, l2, l4, l6}, b2[BMI_MAX_RULE] = {l1, l3, l5, l7};
    const double *v = vtx + 9 * k;
    const double x = v[6 + d] + b1[I] * (v[d] - v[6 + d]) + b2[I] * (v[3 + d] - v[6 + d]);
    pts[u] = x - cen[3 * k + d];
}

// Packed geometry records (BMI_GEO doubles) of groups of BMI_PATCH cells:
// rule points, centroids, areas, vertices relative to the centroid; zero for
// empty slots.  One work-group per group.
__global__ void k_geo_records(const double *pts, const double *vtx, const double *cen, const double *area,
                              const int *cells, int ngroups, double *g) {
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    if (b >= ngroups) return;
    for (int i = threadIdx.x; i < BMI_GEO; i += blockDim.x) {
        double v = 0.0;
        int c, e;
        if (i < BMI_GEO_CEN) { c = i / (3 * BMI_MAX_RULE); e = i % (3 * BMI_MAX_RULE); }
        else if (i < BMI_GEO_AREA) { c = (i - BMI_GEO_CEN) / 3; e = (i - BMI_GEO_CEN) % 3; }
        else if (i < BMI_GEO_VREL) { c = i - BMI_GEO_AREA; e = 0; }
        else { c = (i - BMI_GEO_VREL) / 9; e = (i - BMI_GEO_VREL) % 9; }
        const long long K = cells[(long long)b * BMI_PATCH + c];
        if (K >= 0) {
            if (i < BMI_GEO_CEN) v = pts[K * 3 * BMI_MAX_RULE + e];
            else if (i < BMI_GEO_AREA) v = cen[3 * K + e];
            else if (i < BMI_GEO_VREL) v = area[K];
            else v = vtx[9 * K + e] - cen[3 * K + e % 3];
        }
        g[(long long)b * BMI_GEO + i] = v;
    }
}

hipError_t launch_rule_points(const double *vtx, const double *cen, long long nt, int n, const double l[][3],
                              double *pts, hipStream_t s) {
    const long long tot = nt * 3 * BMI_MAX_RULE;
    if (tot <= 0) return hipSuccess;
    double b[8] = {};
    for (int I = 0; I < n && I < BMI_MAX_RULE; ++I) {
        b[2 * I] = l[I][0];
        b[2 * I + 1] = l[I][1];
    }
    hipLaunchKernelGGL(k_rule_points, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, vtx, cen, nt, n, b[0], b[1],
                       b[2], b[3], b[4], b[5], b[6], b[7], pts);
    return hipGetLastError();
}

hipError_t launch_geo_records(const double *pts, const double *vtx, const double *cen, const double *area,
                              const int *cells, int ngroups, double *g, hipStream_t s) {
    if (ngroups <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_geo_records, dim3((unsigned)ngroups), dim3(256), 0, s, pts, vtx, cen, area, cells, ngroups, g);
    return hipGetLastError();
}

hipError_t launch_classify(const ClassifyArgs *args, const int *gbase, int nprob, int total_groups, hipStream_t s) {
    if (total_groups <= 0) return hipSuccess;
    // up to 8 work-groups per sub-batch, for at least 8192 work-groups
    const int split = std::max(1, std::min(8, 8192 / total_groups));
    hipLaunchKernelGGL(k_classify, dim3((unsigned)total_groups * (unsigned)split), dim3(BMI_BLOCK_THREADS), 0, s, args,
                       gbase, nprob, split);
    return hipGetLastError();
}

// ------------------------------------------------------- special pairs

// Accumulator of one lane for a special pair: moment form for RWG (16 for
// the single layer, 24 for the double layer), 3x3 shape products for P1.
constexpr int spec_nacc(int op) { return op == OPK_DL ? 24 : (op == OPK_MW ? 16 : 18); }
template <int OP>
struct SpecAcc {
    double v[spec_nacc(OP)];
};


// Sums of each of the N accumulators over each segment of SEG lanes (64: the
// whole wave; 32: its halves), returned to every lane of the segment.
// Groups of eight accumulators are reduced and scattered at once: at the
// exchanges with lanes SEG/2, SEG/4 and SEG/8 apart each lane keeps half of
// its remaining values and sends the other half (4 + 2 + 1 exchanges), so
// that lane l ends with value (l / (SEG/8)) mod 8 summed over SEG/8 lanes;
// butterfly steps over the lowest bits finish that sum, and one exchange per
// accumulator hands it to every lane of the segment: 17-18 exchanges per
// group instead of 5-6 per accumulator.  The order of the additions is
// fixed, and both lanes of an exchange add the same two numbers, so every
// lane that holds a sum holds the same one.
template <int N, int SEG = 64>
__device__ __forceinline__ void wave_sum_all(double (&v)[N]) {
    static_assert(SEG == 64 || SEG == 32, "segments of 64 or 32 lanes");
    constexpr int G = (N + 7) / 8, Q = SEG / 8;
    const int lane = threadIdx.x & 63;
    double t[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        double w[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = 8 * g + j < N ? v[8 * g + j] : 0.0;
        // (three rounds written out: a loop with two counters kept the
        // halves' indices at run time, and w in scratch memory)
        auto halve = [&](auto o_c) {
            constexpr int o = decltype(o_c)::value, h = o / Q;
            const bool up = (lane & o) != 0;
#pragma unroll
            for (int j = 0; j < h; ++j) {
                const double keep = up ? w[h + j] : w[j];
                const double give = up ? w[j] : w[h + j];
                w[j] = keep + __shfl_xor(give, o);
            }
        };
        halve(std::integral_constant<int, 4 * Q>{});
        halve(std::integral_constant<int, 2 * Q>{});
        halve(std::integral_constant<int, Q>{});
#pragma unroll
        for (int o = Q / 2; o >= 1; o >>= 1) w[0] += __shfl_xor(w[0], o);
        t[g] = w[0];
    }
    const int base = lane & ~(SEG - 1);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = __shfl(t[i / 8], base + Q * (i % 8));
}

// Entry (i, j) of the EFIE local matrix from the moments m[0..16) (M0, Mx,
// My, Mxy as real/imaginary pairs in the order of the special-pair
// accumulators) and a = v_i - c_K, b = v_j - c_L, the vertices opposite the
// two shapes' edges relative to the centroids; as mw_epilogue.
__device__ __forceinline__ double2 mw_entry(const OpConst &op, const double *m, const double *a, const double *b) {
    const double ur0 = m[2] - a[0] * m[0], ur1 = m[3] - a[1] * m[0], ur2 = m[4] - a[2] * m[0];
    const double ui0 = m[5] - a[0] * m[1], ui1 = m[6] - a[1] * m[1], ui2 = m[7] - a[2] * m[1];
    const double ar = fma(-a[2], m[10], fma(-a[1], m[9], fma(-a[0], m[8], m[14])));
    const double ai = fma(-a[2], m[13], fma(-a[1], m[12], fma(-a[0], m[11], m[15])));
    const double vr = fma(-b[2], ur2, fma(-b[1], ur1, fma(-b[0], ur0, ar)));
    const double vi = fma(-b[2], ui2, fma(-b[1], ui1, fma(-b[0], ui0, ai)));
    const double dr = 4.0 * (op.beta_re * m[0] - op.beta_im * m[1]);
    const double di = 4.0 * (op.beta_re * m[1] + op.beta_im * m[0]);
    double2 z;
    z.x = fma(op.alpha_re, vr, fma(-op.alpha_im, vi, dr));
    z.y = fma(op.alpha_re, vi, fma(op.alpha_im, vr, di));
    return z;
}

struct CellG {
    double o[3];   // v3 - centroid
    double t1[3];  // v1 - v3
    double t2[3];  // v2 - v3
};

// SMOOTH: the smooth remainder of G (near pairs of the single layers) in
// place of G.  PAIR (single layers on one cell, the charts equal): the node
// stands for itself and for its x <-> y swap, whose kernel value is the same
// (R is symmetric), so h is counted for both orders: the moments of x and of
// y each gather x + y and x.y is counted twice (Maxwell), the shape products
// f_i g_j + g_i f_j (Helmholtz).
template <int OP, bool SMOOTH = false, bool PAIR = false>
__device__ __forceinline__ void spec_point(const OpConst &op, bool lossy, const double2 *tab, SpecAcc<OP> &acc, const CellG &X, const CellG &Y,
                                           const double *dc, double bx1, double bx2, double by1, double by2,
                                           double w, const double *nrm) {
    double x[3], y[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = X.o[d] + X.t1[d] * bx1 + X.t2[d] * bx2;
        y[d] = Y.o[d] + Y.t1[d] * by1 + Y.t2[d] * by2;
    }
    if constexpr (OP == OPK_DL) {
        // (n x a_c).(grad G x b_c') = q [(n.r)(a_c.b_c') - (n.b_c')(a_c.r)] for the
        // chart vertices c, c' (a_c = x - X_c, b_c' = y - Y_c', relative to the
        // centroids); summed over the nodes it is a combination of moments of
        // h = w q, twelve complex accumulators:
        //   M1 = h n.r, M25 = h [(n.r)(x.y) - (n.y)(x.r)], M7 = h x.r,
        //   M63 = h [(n.y) r - (n.r) y], M4 = h (n.r) x, M8 = h r
        // entry (c, c') = M25 + X.M63 - Y.(M4 - X M1) + (n.Y)(M7 - X.M8)
        const double r0 = dc[0] + x[0] - y[0], r1 = dc[1] + x[1] - y[1], r2 = dc[2] + x[2] - y[2];
        double hr, hi;
        green_grad64<true>(op, lossy, w * (1.0 / (4.0 * M_PI)), r0, r1, r2, hr, hi, tab);
        const double nr = fma(nrm[2], r2, fma(nrm[1], r1, nrm[0] * r0));
        const double ny = fma(nrm[2], y[2], fma(nrm[1], y[1], nrm[0] * y[0]));
        const double xy = fma(x[2], y[2], fma(x[1], y[1], x[0] * y[0]));
        const double xr = fma(x[2], r2, fma(x[1], r1, x[0] * r0));
        const double r[3] = {r0, r1, r2};
        const double hnr = hr * nr, hni = hi * nr;
        const double t25 = fma(nr, xy, -ny * xr);
        acc.v[0] += hnr;
        acc.v[1] += hni;
        acc.v[2] = fma(hr, t25, acc.v[2]);
        acc.v[3] = fma(hi, t25, acc.v[3]);
        acc.v[4] = fma(hr, xr, acc.v[4]);
        acc.v[5] = fma(hi, xr, acc.v[5]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double t63 = fma(ny, r[d], -nr * y[d]);
            acc.v[6 + d] = fma(hr, t63, acc.v[6 + d]);
            acc.v[9 + d] = fma(hi, t63, acc.v[9 + d]);
            acc.v[12 + d] = fma(hnr, x[d], acc.v[12 + d]);
            acc.v[15 + d] = fma(hni, x[d], acc.v[15 + d]);
            acc.v[18 + d] = fma(hr, r[d], acc.v[18 + d]);
            acc.v[21 + d] = fma(hi, r[d], acc.v[21 + d]);
        }
        return;
    }
    double hr, hi;
    if (SMOOTH)
        green_smooth64<true>(op, lossy, w * (1.0 / (4.0 * M_PI)), dc[0] + x[0] - y[0], dc[1] + x[1] - y[1],
                             dc[2] + x[2] - y[2], hr, hi, tab);
    else
        green64<true>(op, lossy, w * (1.0 / (4.0 * M_PI)), dc[0] + x[0] - y[0], dc[1] + x[1] - y[1],
                      dc[2] + x[2] - y[2], hr, hi, tab);
    if (OP == OPK_MW) {
        if constexpr (PAIR) {
            const double xy = x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
            acc.v[0] = fma(2.0, hr, acc.v[0]);
            acc.v[1] = fma(2.0, hi, acc.v[1]);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double sd = x[d] + y[d];
                acc.v[2 + d] = fma(hr, sd, acc.v[2 + d]);
                acc.v[5 + d] = fma(hi, sd, acc.v[5 + d]);
            }
            acc.v[14] = fma(2.0 * hr, xy, acc.v[14]);
            acc.v[15] = fma(2.0 * hi, xy, acc.v[15]);
            return;  // the moments of y equal those of x: copied after the reduction
        }
        acc.v[0] += hr;
        acc.v[1] += hi;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            acc.v[2 + d] = fma(hr, x[d], acc.v[2 + d]);
            acc.v[5 + d] = fma(hi, x[d], acc.v[5 + d]);
            acc.v[8 + d] = fma(hr, y[d], acc.v[8 + d]);
            acc.v[11 + d] = fma(hi, y[d], acc.v[11 + d]);
        }
        const double xy = x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
        acc.v[14] = fma(hr, xy, acc.v[14]);
        acc.v[15] = fma(hi, xy, acc.v[15]);
    } else {
        const double f[3] = {bx1, bx2, 1.0 - bx1 - bx2};
        const double g[3] = {by1, by2, 1.0 - by1 - by2};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double fr = hr * f[i], fi = hi * f[i];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if constexpr (PAIR) {
                    const double sij = fma(f[i], g[j], g[i] * f[j]);
                    acc.v[2 * (3 * i + j)] = fma(hr, sij, acc.v[2 * (3 * i + j)]);
                    acc.v[2 * (3 * i + j) + 1] = fma(hi, sij, acc.v[2 * (3 * i + j) + 1]);
                } else {
                    acc.v[2 * (3 * i + j)] = fma(fr, g[j], acc.v[2 * (3 * i + j)]);
                    acc.v[2 * (3 * i + j) + 1] = fma(fi, g[j], acc.v[2 * (3 * i + j) + 1]);
                }
            }
        }
    }
}



// Hypercube maps of the Sauter-Schwab regions (SauterSchwab.jl:265-298):
// region r of class cls at the node (e1, e2, e3, xi) with product weight w
// gives the reference coordinates (u1, u2) of the test and (v1, v2) of the
// trial chart and the weight times the region's Jacobian.
__device__ __forceinline__ void ss_map(int cls, int r, double e1, double e2, double e3, double xi, double w,
                                       double &u1, double &u2, double &v1, double &v2, double &m) {
    const double xi3 = xi * xi * xi;
    if (cls == CLS_CV) {
        const double xe1 = xi * e1, xe2 = xi * e2;
        m = w * xi3 * e2;
        if (r == 0) { u1 = 1 - xi; u2 = xe1; v1 = 1 - xe2; v2 = xe2 * e3; }
        else { u1 = 1 - xe2; u2 = xe2 * e3; v1 = 1 - xi; v2 = xe1; }
    } else if (cls == CLS_CE) {
        const double xe1 = xi * e1, e12 = e1 * e2, e23 = e2 * e3, e123 = e12 * e3;
        m = w * xi3 * e1 * e1 * (r == 4 ? 1.0 : e2);
        switch (r) {
        case 0: u1 = 1 - xi; u2 = xe1; v1 = 1 - xi * (1 - e123); v2 = xe1 * e2 * (1 - e3); break;
        case 1: u1 = 1 - xi * (1 - e12); u2 = xe1 * (1 - e2); v1 = 1 - xi; v2 = xe1 * e23; break;
        case 2: u1 = 1 - xi * (1 - e123); u2 = xe1 * e2 * (1 - e3); v1 = 1 - xi; v2 = xe1; break;
        case 3: u1 = 1 - xi * (1 - e123); u2 = xe1 * (1 - e23); v1 = 1 - xi; v2 = xe1 * e2; break;
        default: u1 = 1 - xi; u2 = xe1 * e3; v1 = 1 - xi * (1 - e12); v2 = xe1 * (1 - e2); break;
        }
    } else {
        const double p = xi * e1;
        m = w * xi3 * e1 * e1 * e2;
        switch (r) {
        case 0: u1 = 1 - xi; u2 = xi - p + p * e2; v1 = 1 - (xi - p * e2 * e3); v2 = xi - p; break;
        case 1: u1 = 1 - (xi - p * e2 * e3); u2 = xi - p; v1 = 1 - xi; v2 = xi - p + p * e2; break;
        case 2: u1 = 1 - xi; u2 = p * (1 - e2 + e2 * e3); v1 = 1 - (xi - p * e2); v2 = p * (1 - e2); break;
        case 3: u1 = 1 - (xi - p * e2); u2 = p * (1 - e2); v1 = 1 - xi; v2 = p * (1 - e2 + e2 * e3); break;
        case 4: u1 = 1 - (xi - p * e2 * e3); u2 = p * (1 - e2 * e3); v1 = 1 - xi; v2 = p * (1 - e2); break;
        default: u1 = 1 - xi; u2 = p * (1 - e2); v1 = 1 - (xi - p * e2 * e3); v2 = p * (1 - e2 * e3); break;
        }
    }
}

// The region maps of ss_map are linear in the radial variable xi and put the
// singular point at xi = 0, chart point (1, 0) on both cells:
//   u1 = 1 + xi b1,  u2 = xi b2,  v1 = 1 + xi d1,  v2 = xi d2,
// with the Jacobian xi^3 f (times the product weight).  ss_dir gives the
// directions (b1, b2, d1, d2) and f of region r of class cls at (e1, e2, e3),
// the same maps written as such.
__host__ __device__ __forceinline__ void ss_dir(int cls, int r, double e1, double e2, double e3, double &b1,
                                                double &b2, double &d1, double &d2, double &f) {
    if (cls == CLS_CV) {
        f = e2;
        const double e23 = e2 * e3;
        if (r == 0) { b1 = -1.0; b2 = e1; d1 = -e2; d2 = e23; }
        else { b1 = -e2; b2 = e23; d1 = -1.0; d2 = e1; }
    } else if (cls == CLS_CE) {
        const double e12 = e1 * e2, e23 = e2 * e3, e123 = e12 * e3;
        f = e1 * e1 * (r == 4 ? 1.0 : e2);
        switch (r) {
        case 0: b1 = -1.0; b2 = e1; d1 = e123 - 1.0; d2 = e12 * (1.0 - e3); break;
        case 1: b1 = e12 - 1.0; b2 = e1 * (1.0 - e2); d1 = -1.0; d2 = e1 * e23; break;
        case 2: b1 = e123 - 1.0; b2 = e12 * (1.0 - e3); d1 = -1.0; d2 = e1; break;
        case 3: b1 = e123 - 1.0; b2 = e1 * (1.0 - e23); d1 = -1.0; d2 = e12; break;
        default: b1 = -1.0; b2 = e1 * e3; d1 = e12 - 1.0; d2 = e1 * (1.0 - e2); break;
        }
    } else {
        const double e12 = e1 * e2, e123 = e12 * e3;
        f = e1 * e1 * e2;
        switch (r) {
        case 0: b1 = -1.0; b2 = 1.0 - e1 + e12; d1 = e123 - 1.0; d2 = 1.0 - e1; break;
        case 1: b1 = e123 - 1.0; b2 = 1.0 - e1; d1 = -1.0; d2 = 1.0 - e1 + e12; break;
        case 2: b1 = -1.0; b2 = e1 * (1.0 - e2 + e2 * e3); d1 = e12 - 1.0; d2 = e1 * (1.0 - e2); break;
        case 3: b1 = e12 - 1.0; b2 = e1 * (1.0 - e2); d1 = -1.0; d2 = e1 * (1.0 - e2 + e2 * e3); break;
        case 4: b1 = e123 - 1.0; b2 = e1 * (1.0 - e2 * e3); d1 = -1.0; d2 = e1 * (1.0 - e2); break;
        default: b1 = -1.0; b2 = e1 * (1.0 - e2); d1 = e123 - 1.0; d2 = e1 * (1.0 - e2 * e3); break;
        }
    }
}

// Closed-form integrals over the flat trial triangle w[0..2] (unit normal n)
// at the point x, R = |x - y|, xp the projection of x onto its plane:
//   I0 = int 1/R,  I1 = int (y - xp)/R,  K1 = int R,  K1v = int (y - xp) R.
// Per edge a -> b (counter-clockwise about n): l the unit edge vector, m = l x n
// its outward in-plane normal, t+- = (b, a - xp).l, p = (a - xp).m, h = |x - xp|,
// r+- = |x - b|, |x - a|, r0^2 = p^2 + h^2, L = ln((r+ + t+)/(r- + t-)); the
// edge integrals E1 = int_e R = [t r + r0^2 L]/2 and
// E3 = int_e R^3 = [t r^3]/4 + 3 r0^2 [t r]/8 + 3 r0^4 L/8 ([f] = f(+) - f(-))
// give I1 = sum m E1, K1v = sum m E3/3, K1 = (sum p E1 + h^2 I0)/3 and
// I0 = sum p L - h [atan(p t+ / (r0^2 + h r+)) - atan(p t- / (r0^2 + h r-))]
// (Wilton et al., IEEE TAP 32, 1984; Graglia, IEEE TAP 41, 1993).  L is taken
// in the one of three equal forms that subtracts no nearly equal numbers.
__device__ __forceinline__ void tri_potentials64(const double (&w)[3][3], const double (&n)[3], const double (&x)[3],
                                                 double &I0, double (&I1)[3], double &K1, double (&K1v)[3],
                                                 double (&xp)[3]) {
    const double h0 = (x[0] - w[0][0]) * n[0] + (x[1] - w[0][1]) * n[1] + (x[2] - w[0][2]) * n[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) xp[d] = x[d] - h0 * n[d];
    const double h = fabs(h0);
    double i0 = 0.0, k1 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) I1[d] = K1v[d] = 0.0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int e1 = e == 2 ? 0 : e + 1;
        double l[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) l[d] = w[e1][d] - w[e][d];
        const double il = rsqrt64(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) l[d] *= il;
        const double m[3] = {l[1] * n[2] - l[2] * n[1], l[2] * n[0] - l[0] * n[2], l[0] * n[1] - l[1] * n[0]};
        double tp = 0.0, tm = 0.0, p = 0.0, rp2 = 0.0, rm2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            tp = fma(w[e1][d] - xp[d], l[d], tp);
            tm = fma(w[e][d] - xp[d], l[d], tm);
            p = fma(w[e][d] - xp[d], m[d], p);
            rp2 = fma(x[d] - w[e1][d], x[d] - w[e1][d], rp2);
            rm2 = fma(x[d] - w[e][d], x[d] - w[e][d], rm2);
        }
        const double rp = sqrt(rp2), rm = sqrt(rm2);
        const double r0sq = fma(p, p, h * h);
        double lg = 0.0;
        if (tm >= 0.0) lg = log((rp + tp) / (rm + tm));
        else if (tp <= 0.0) lg = log((rm - tm) / (rp - tp));
        else if (r0sq > 0.0) lg = log((rp + tp) * (rm - tm) / r0sq);
        if (fabs(p) * il > 1e-14)
            i0 += p * lg - h * (atan(p * tp / fma(h, rp, r0sq)) - atan(p * tm / fma(h, rm, r0sq)));
        const double tr = tp * rp - tm * rm;
        const double E1 = 0.5 * fma(r0sq, lg, tr);
        const double E3 = 0.25 * (tp * rp * rp2 - tm * rm * rm2) + 0.375 * r0sq * fma(r0sq, lg, tr);
        k1 = fma(p, E1, k1);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            I1[d] = fma(E1, m[d], I1[d]);
            K1v[d] = fma(E3 * (1.0 / 3.0), m[d], K1v[d]);
        }
    }
    I0 = i0;
    K1 = fma(h * h, i0, k1) * (1.0 / 3.0);
}

// Near pair of a single layer, outer point (b1, b2) of weight nw on the test
// cell: adds the trial-cell integrals of the singular terms
// [1/R + (gamma^2/2) R]/(4 pi) of G, in closed form, to the accumulators in the
// form spec_point gives them (reference weights: sum_J nw_J f(y_J) stands for
// int_L f / (2 A_L)).  The remainder is left to spec_point<OP, true>.
template <int OP>
__device__ __forceinline__ void near_static(const OpConst &op, SpecAcc<OP> &acc, const CellG &X, const CellG &Y,
                                            const double *dc, double b1, double b2, double nw) {
    double x[3], xl[3], w[3][3], n[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = X.o[d] + X.t1[d] * b1 + X.t2[d] * b2;  // relative to the test centroid
        xl[d] = x[d] + dc[d];                        // relative to the trial centroid
        w[0][d] = Y.o[d] + Y.t1[d];
        w[1][d] = Y.o[d] + Y.t2[d];
        w[2][d] = Y.o[d];
    }
    n[0] = Y.t1[1] * Y.t2[2] - Y.t1[2] * Y.t2[1];
    n[1] = Y.t1[2] * Y.t2[0] - Y.t1[0] * Y.t2[2];
    n[2] = Y.t1[0] * Y.t2[1] - Y.t1[1] * Y.t2[0];
    const double nn2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    const double inv = rsqrt64(nn2);
#pragma unroll
    for (int d = 0; d < 3; ++d) n[d] *= inv;
    const double two_area = nn2 * inv;
    double I0, I1[3], K1, K1v[3], xp[3];
    tri_potentials64(w, n, xl, I0, I1, K1, K1v, xp);
    // gamma^2 / 2
    const double k = op.k2pi * 6.283185307179586476925, gr = op.gamma_re;
    const double g2r = 0.5 * (gr * gr - k * k), g2i = gr * k;
    const double sc = nw / (two_area * 4.0 * M_PI);
    const double s0r = sc * fma(g2r, K1, I0), s0i = sc * g2i * K1;
    double s1r[3], s1i[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = fma(xp[d], K1, K1v[d]);
        s1r[d] = sc * (fma(xp[d], I0, I1[d]) + g2r * v);
        s1i[d] = sc * g2i * v;
    }
    if (OP == OPK_MW) {
        acc.v[0] += s0r;
        acc.v[1] += s0i;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            acc.v[2 + d] = fma(s0r, x[d], acc.v[2 + d]);
            acc.v[5 + d] = fma(s0i, x[d], acc.v[5 + d]);
            acc.v[8 + d] += s1r[d];
            acc.v[11 + d] += s1i[d];
        }
        acc.v[14] += x[0] * s1r[0] + x[1] * s1r[1] + x[2] * s1r[2];
        acc.v[15] += x[0] * s1i[0] + x[1] * s1i[1] + x[2] * s1i[2];
    } else {
        // barycentric coordinates of the trial chart as affine functions
        // lambda_j(y) = c0_j + c_j . y (Gram-matrix inverse of t1, t2)
        const double g11 = Y.t1[0] * Y.t1[0] + Y.t1[1] * Y.t1[1] + Y.t1[2] * Y.t1[2];
        const double g12 = Y.t1[0] * Y.t2[0] + Y.t1[1] * Y.t2[1] + Y.t1[2] * Y.t2[2];
        const double g22 = Y.t2[0] * Y.t2[0] + Y.t2[1] * Y.t2[1] + Y.t2[2] * Y.t2[2];
        const double idet = 1.0 / (g11 * g22 - g12 * g12);
        double c[3][3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            c[0][d] = (g22 * Y.t1[d] - g12 * Y.t2[d]) * idet;
            c[1][d] = (g11 * Y.t2[d] - g12 * Y.t1[d]) * idet;
            c[2][d] = -c[0][d] - c[1][d];
        }
        const double f[3] = {b1, b2, 1.0 - b1 - b2};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double c0 = (j == 2 ? 1.0 : 0.0) - (c[j][0] * Y.o[0] + c[j][1] * Y.o[1] + c[j][2] * Y.o[2]);
            const double glr = c0 * s0r + c[j][0] * s1r[0] + c[j][1] * s1r[1] + c[j][2] * s1r[2];
            const double gli = c0 * s0i + c[j][0] * s1i[0] + c[j][1] * s1i[1] + c[j][2] * s1i[2];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                acc.v[2 * (3 * i + j)] = fma(f[i], glr, acc.v[2 * (3 * i + j)]);
                acc.v[2 * (3 * i + j) + 1] = fma(f[i], gli, acc.v[2 * (3 * i + j) + 1]);
            }
        }
    }
}

// One wave per touching pair (NEARK false; the Maxwell operators with the
// compile-time order: half a wave per pair) or per near pair (NEARK true),
// owner entries only: a pair whose test cell belongs to several sub-batches
// is integrated once, in the sub-batch the cell calls home.  Touching pairs:
// regions in turn, the pair's lanes striding over the hypercube nodes, one
// Green evaluation (or, in the radial form, one 1/|B| and the xi loop) each.
// Near pairs: closed-form singular terms, one outer point per lane, then
// lanes striding over the product rule.  wave_sum_all reduces the
// accumulators over the pair's lanes.  The two kinds are separate instantiations
// so that the near-pair code does not add to the registers of the
// Sauter-Schwab loop.
// NG: Gauss-Legendre points per hypercube dimension fixed at compile time
// (the default 5; constant-divisor node decoding), 0: read per class.
template <int OP, int NG, bool NEARK>
__global__ __launch_bounds__(BMI_BLOCK_THREADS, (OP == OPK_MW && !NEARK) ? 3 : 2) void k_special(
    const SpecialArgs *args, const int *cbase, int nprob, int cs) {
    const int pi = find_problem(cbase, nprob, (int)blockIdx.x);
    const SpecialArgs &a = args[pi];
    __shared__ double s_glx[3][BMI_MAX_GL], s_glw[3][BMI_MAX_GL];
    __shared__ double s_nb1[BMI_MAX_NEAR], s_nb2[BMI_MAX_NEAR], s_nw[BMI_MAX_NEAR];
    __shared__ double2 s_sc[64];
    // compile-time order (touching pairs) of the Maxwell operators: two
    // pairs per wave, 32 lanes each, so that the per-pair set-up and the
    // reduction serve two pairs (EFIE special phase -14 %; four pairs per
    // wave, or two in the Helmholtz kernel, gained nothing)
    constexpr int PPW = (NG != 0 && !NEARK && OP != OPK_HH3D) ? 2 : 1, LPP = 64 / PPW;
    constexpr int N3 = NG * NG * NG, NH3 = NG ? (N3 + LPP - 1) / LPP : 1;
    __shared__ double4 s_node[NG ? LPP * NH3 : 1];
    __shared__ int s_items[BMI_BLOCK_THREADS];
    __shared__ int s_wsum[BMI_BLOCK_THREADS / 64];
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    const int ns = min(*a.n_special, a.capacity);
    const bool lossy = a.op.lossy != 0;
    // one chunk of cs (<= 256) entries: the threads pick out the touching (or
    // near) entries the problem owns, listed by decreasing cost (common face:
    // 6 regions, common edge: 5, common vertex: 2), then each wave integrates
    // every fourth, so the waves of the work-group finish close together
    {
        const int c0 = ((int)blockIdx.x - cbase[pi]) * cs;
        const int ec = c0 + t;
        int take = 0, cl = 0;
        if (t < cs && ec < ns) {
            cl = a.spec_cls[ec];
            take = !(cl & SPEC_NOT_OWNER) && (NEARK ? cl == CLS_NEAR : (cl >= CLS_CV && cl <= CLS_CF));
            // blocks that k_link fills: zero common faces, mirrored common vertices
            if (!NEARK && take && cl == CLS_CF && a.cf_zero) take = 0;
            if (!NEARK && take && cl == CLS_CV && a.sym && a.spec_K[ec] > a.spec_L[ec]) take = 0;
        }
        int nitem = 0;
#pragma unroll
        for (int pass = 0; pass < (NEARK ? 1 : 3); ++pass) {
            const bool mine = take && (NEARK || cl == CLS_CF - pass);
            const unsigned long long m = __ballot(mine);
            __syncthreads();  // s_items, s_wsum of the previous chunk (pass) are consumed
            if (lane == 0) s_wsum[wv] = __popcll(m);
            __syncthreads();
            int base = 0, n = 0;
#pragma unroll
            for (int i = 0; i < BMI_BLOCK_THREADS / 64; ++i) {
                base += i < wv ? s_wsum[i] : 0;
                n += s_wsum[i];
            }
            if (mine) s_items[nitem + base + __popcll(m & ((1ull << lane) - 1ull))] = ec;
            nitem += n;
        }
        // a chunk without such entries (the mixed path's chunks of class-5
        // entries, the near kernel's chunks of touching ones) ends here,
        // before the tables
        if (nitem == 0) return;
        load_sincos_table(s_sc, a.sc);
        if (t < 3 * BMI_MAX_GL) {
            s_glx[t / BMI_MAX_GL][t % BMI_MAX_GL] = a.glx[t / BMI_MAX_GL][t % BMI_MAX_GL];
            s_glw[t / BMI_MAX_GL][t % BMI_MAX_GL] = a.glw[t / BMI_MAX_GL][t % BMI_MAX_GL];
        }
        if (t < BMI_MAX_NEAR) {
            s_nb1[t] = a.nb1[t];
            s_nb2[t] = a.nb2[t];
            s_nw[t] = a.nw[t];
        }
        __syncthreads();
        // compile-time order: node table of three Gauss-Legendre indices, entry
        // u = (x[i], x[j], x[k], w[i] w[j] w[k]) for u = (i NG + j) NG + k < NG^3,
        // padded to whole rounds of a pair's lanes with the last node at weight
        // zero; (i, j, k) are (e1, e2, e3) of the radial form
        if constexpr (NG != 0) {
            for (int u = t; u < LPP * NH3; u += BMI_BLOCK_THREADS) {
                const int nd = min(u, N3 - 1);
                const int i2 = nd / (NG * NG), i3 = (nd / NG) % NG, i4 = nd % NG;
                s_node[u] = make_double4(s_glx[0][i2], s_glx[0][i3], s_glx[0][i4],
                                         (u < N3 ? 1.0 : 0.0) * s_glw[0][i2] * s_glw[0][i3] * s_glw[0][i4]);
            }
            __syncthreads();
        }
    // pair slot ps of the wave takes item w0 + ps on its LPP lanes (lane pl
    // of the pair); a slot past the last item repeats the wave's first and
    // stores nothing
    const int pl = lane % LPP, ps = lane / LPP;
    for (int w0 = wv * PPW; w0 < nitem; w0 += (BMI_BLOCK_THREADS / 64) * PPW) {
        const bool valid = w0 + ps < nitem;
        const int e = s_items[valid ? w0 + ps : w0];
        const int cls = a.spec_cls[e];
        const int K = a.spec_K[e], L = a.spec_L[e];
        // vertex permutations of the two cells (identity for near pairs):
        // chart vertex c of the test cell is its vertex PT[c]
        int PT[3] = {0, 1, 2}, PS[3] = {0, 1, 2};
        if constexpr (!NEARK) {
            // vertex correspondence (pulledback_integrand_gpu.jl:22-155)
            int mi[3] = {0, 0, 0}, mj[3] = {0, 0, 0}, ne = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double dx = a.X.vtx[9 * K + 3 * i] - a.Y.vtx[9 * L + 3 * j];
                    const double dy = a.X.vtx[9 * K + 3 * i + 1] - a.Y.vtx[9 * L + 3 * j + 1];
                    const double dz = a.X.vtx[9 * K + 3 * i + 2] - a.Y.vtx[9 * L + 3 * j + 2];
                    if (dx * dx + dy * dy + dz * dz < DTOL * DTOL && ne < 3) {
                        mi[ne] = i;
                        mj[ne] = j;
                        ++ne;
                    }
                }
            if (cls == CLS_CF) {
                PS[0] = PS[1] = PS[2] = 0;
                for (int k = 0; k < ne; ++k) PS[mi[k]] = mj[k];
            } else if (cls == CLS_CE) {
                // common edge between chart vertices 1 and 3
                PT[0] = mi[1]; PT[1] = 3 - mi[0] - mi[1]; PT[2] = mi[0];
                PS[0] = mj[1]; PS[1] = 3 - mj[0] - mj[1]; PS[2] = mj[0];
            } else {
                // common vertex at chart vertex 1
                PT[0] = mi[0]; PT[1] = mi[0] == 0 ? 1 : 0; PT[2] = 3 - PT[0] - PT[1];
                PS[0] = mj[0]; PS[1] = mj[0] == 0 ? 1 : 0; PS[2] = 3 - PS[0] - PS[1];
            }
        }
        // permuted charts relative to the centroids: a chart point (u1, u2)
        // is u1 V[P0] + u2 V[P1] + (1 - u1 - u2) V[P2], the reference chart
        // map of Integrand_gpu.jl:51-55 composed with the cell's own chart
        CellG X, Y;
        double dc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double cx = a.X.cen[3 * K + d], cy = a.Y.cen[3 * L + d];
            const double x0 = a.X.vtx[9 * K + 3 * PT[0] + d], x1 = a.X.vtx[9 * K + 3 * PT[1] + d],
                         x2 = a.X.vtx[9 * K + 3 * PT[2] + d];
            const double y0 = a.Y.vtx[9 * L + 3 * PS[0] + d], y1 = a.Y.vtx[9 * L + 3 * PS[1] + d],
                         y2 = a.Y.vtx[9 * L + 3 * PS[2] + d];
            X.o[d] = x2 - cx;
            X.t1[d] = x0 - x2;
            X.t2[d] = x1 - x2;
            Y.o[d] = y2 - cy;
            Y.t1[d] = y0 - y2;
            Y.t2[d] = y1 - y2;
            dc[d] = cx - cy;
        }
        // unit normal of the test cell ((v1 - v3) x (v2 - v3) direction; the
        // chart permutation may reverse t1 x t2)
        double nrm[3] = {0.0, 0.0, 0.0};
        if (OP == OPK_DL) {
            double e1[3], e2[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                e1[d] = a.X.vtx[9 * K + d] - a.X.vtx[9 * K + 6 + d];
                e2[d] = a.X.vtx[9 * K + 3 + d] - a.X.vtx[9 * K + 6 + d];
            }
            nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
            nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
            nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
            const double ninv = rsqrt64(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            nrm[0] *= ninv;
            nrm[1] *= ninv;
            nrm[2] *= ninv;
        }
        SpecAcc<OP> acc;
#pragma unroll
        for (int i = 0; i < spec_nacc(OP); ++i) acc.v[i] = 0.0;

        if constexpr (NEARK) {
            // single layers: singular terms in closed form, one outer point
            // per lane, then the smooth remainder by the product rule; the
            // double layer keeps the plain product rule
            const int np = a.near_np, tot = np * np;
            constexpr bool SUB = OP != OPK_DL;
            if (SUB && lane < np) near_static<OP>(a.op, acc, X, Y, dc, s_nb1[lane], s_nb2[lane], s_nw[lane]);
#pragma unroll 1
            for (int idx = lane; idx < tot; idx += 64) {
                const int I = idx / np, J = idx - np * (idx / np);
                spec_point<OP, SUB>(a.op, lossy, s_sc, acc, X, Y, dc, s_nb1[I], s_nb2[I], s_nb1[J], s_nb2[J],
                                    s_nw[I] * s_nw[J], nrm);
            }
        } else if constexpr (NG != 0) {
            // compile-time order: radial form.  Every region
            // map is linear in xi with x - y = xi B at the node (ss_dir: the
            // singular point is the same physical point on both charts), so
            // R = xi |B|, and x = P + xi Q, y = P' + xi Q' with P, P' that
            // point relative to the centroids.  Lanes over the (e1, e2, e3)
            // node table; per lane one 1/|B| and the xi nodes (uniform over
            // the wave) in a loop that evaluates only the phase: with
            // S_n = sum over xi of h xi^n the lane adds M0 += S_0 and
            // (Maxwell) Q S_1, Q' S_1, (Q.Q') S_2, or (Helmholtz) the shape
            // products of f = F0 + xi F1, g = G0 + xi G1; the P, P' terms
            // are added after the reduction.  The Jacobian xi^3 and 1/R give
            // the node weight w xi^2 / |B|.  Double layer: grad G = q r with
            // q ~ 1/R^3, so the node weight is w / |B|^3 and every moment has
            // a factor r = xi B: the lane adds the (n.r)(x.y), (n.y)(x.r),
            // x.r, (n.y) r - (n.r) y, (n.r) x, r moments from S_1, S_2, S_3
            // (their P, P' parts through M1 = sum h n.r after the reduction).
            double Px[3], Py[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Px[d] = X.o[d] + X.t1[d];
                Py[d] = Y.o[d] + Y.t1[d];
            }
            const double ny0 = nrm[0] * Py[0] + nrm[1] * Py[1] + nrm[2] * Py[2];
            const double c0 = Px[0] * Py[0] + Px[1] * Py[1] + Px[2] * Py[2];
            auto region = [&](auto cls_c, auto r_c) {
                constexpr int C = decltype(cls_c)::value, RR = decltype(r_c)::value;
#pragma unroll 1
                for (int h0 = 0; h0 < NH3; ++h0) {
                    // (e1, e2, e3, w1 w2 w3), weight zero past the table
                    const double4 nd = s_node[pl + LPP * h0];
                    double b1, b2, d1, d2, f;
                    ss_dir(C, RR, nd.x, nd.y, nd.z, b1, b2, d1, d2, f);
                    double qx[3], qy[3], bb = 0.0;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        qx[d] = fma(X.t1[d], b1, X.t2[d] * b2);
                        qy[d] = fma(Y.t1[d], d1, Y.t2[d] * d2);
                        const double bd = qx[d] - qy[d];
                        bb = fma(bd, bd, bb);
                    }
                    const double rb = rsqrt64(bb);
                    const double nb = bb * rb;
                    const double u64 = nb * (64.0 * a.op.k2pi);
                    if constexpr (OP == OPK_DL) {
                        // h = -w (a + i kR) exp(-gamma R) / (4 pi |B|^3)
                        const double mr = -(nd.w * f * rb * rb * rb * (1.0 / (4.0 * M_PI)));
                        const double kb = nb * (2.0 * M_PI * a.op.k2pi);
                        double s1r = 0.0, s1i = 0.0, s2r = 0.0, s2i = 0.0, s3r = 0.0, s3i = 0.0;
#pragma unroll
                        for (int i4 = 0; i4 < NG; ++i4) {
                            const double xi = a.glx[0][i4];
                            double sn, cs;
                            sincos_64ths_tab(u64 * xi, s_sc, sn, cs);
                            double hm = mr * a.glw[0][i4], ga = 1.0;
                            if (lossy) {
                                ga = fma(a.op.gamma_re, nb * xi, 1.0);
                                hm *= exp(-a.op.gamma_re * (nb * xi));
                            }
                            const double kr = kb * xi;
                            const double hr = hm * fma(ga, cs, kr * sn), hi = hm * fma(kr, cs, -(ga * sn));
                            s1r = fma(hr, xi, s1r);
                            s1i = fma(hi, xi, s1i);
                            s2r = fma(hr, xi * xi, s2r);
                            s2i = fma(hi, xi * xi, s2i);
                            s3r = fma(hr, xi * xi * xi, s3r);
                            s3i = fma(hi, xi * xi * xi, s3i);
                        }
                        double B[3], nB = 0.0, ny1 = 0.0, c1 = 0.0, c2 = 0.0, xr1 = 0.0, xr2 = 0.0;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            B[d] = qx[d] - qy[d];
                            nB = fma(nrm[d], B[d], nB);
                            ny1 = fma(nrm[d], qy[d], ny1);
                            c1 = fma(Px[d], qy[d], fma(qx[d], Py[d], c1));
                            c2 = fma(qx[d], qy[d], c2);
                            xr1 = fma(Px[d], B[d], xr1);
                            xr2 = fma(qx[d], B[d], xr2);
                        }
                        const double k1 = fma(nB, c0, -ny0 * xr1);
                        const double k2 = fma(nB, c1, -fma(ny0, xr2, ny1 * xr1));
                        const double k3 = fma(nB, c2, -ny1 * xr2);
                        acc.v[0] = fma(nB, s1r, acc.v[0]);
                        acc.v[1] = fma(nB, s1i, acc.v[1]);
                        acc.v[2] = fma(k1, s1r, fma(k2, s2r, fma(k3, s3r, acc.v[2])));
                        acc.v[3] = fma(k1, s1i, fma(k2, s2i, fma(k3, s3i, acc.v[3])));
                        acc.v[4] = fma(xr1, s1r, fma(xr2, s2r, acc.v[4]));
                        acc.v[5] = fma(xr1, s1i, fma(xr2, s2i, acc.v[5]));
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const double g2 = fma(ny1, B[d], -nB * qy[d]), g1 = ny0 * B[d];
                            acc.v[6 + d] = fma(g1, s1r, fma(g2, s2r, acc.v[6 + d]));
                            acc.v[9 + d] = fma(g1, s1i, fma(g2, s2i, acc.v[9 + d]));
                            const double q4 = nB * qx[d];
                            acc.v[12 + d] = fma(q4, s2r, acc.v[12 + d]);
                            acc.v[15 + d] = fma(q4, s2i, acc.v[15 + d]);
                            acc.v[18 + d] = fma(B[d], s1r, acc.v[18 + d]);
                            acc.v[21 + d] = fma(B[d], s1i, acc.v[21 + d]);
                        }
                        continue;
                    }
                    const double mr = nd.w * f * rb * (1.0 / (4.0 * M_PI));
                    double s0r = 0.0, s0i = 0.0, s1r = 0.0, s1i = 0.0, s2r = 0.0, s2i = 0.0;
#pragma unroll
                    for (int i4 = 0; i4 < NG; ++i4) {
                        const double xi = a.glx[0][i4];
                        double sn, cs;
                        sincos_64ths_tab(u64 * xi, s_sc, sn, cs);
                        double hm = mr * (a.glw[0][i4] * xi * xi);
                        if (lossy) hm *= exp(-a.op.gamma_re * (nb * xi));
                        const double hr = hm * cs, hi = -(hm * sn);
                        s0r += hr;
                        s0i += hi;
                        s1r = fma(hr, xi, s1r);
                        s1i = fma(hi, xi, s1i);
                        s2r = fma(hr, xi * xi, s2r);
                        s2i = fma(hi, xi * xi, s2i);
                    }
                    if constexpr (OP == OPK_MW) {
                        acc.v[0] += s0r;
                        acc.v[1] += s0i;
                        double qq = 0.0;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            acc.v[2 + d] = fma(qx[d], s1r, acc.v[2 + d]);
                            acc.v[5 + d] = fma(qx[d], s1i, acc.v[5 + d]);
                            acc.v[8 + d] = fma(qy[d], s1r, acc.v[8 + d]);
                            acc.v[11 + d] = fma(qy[d], s1i, acc.v[11 + d]);
                            qq = fma(qx[d], qy[d], qq);
                        }
                        acc.v[14] = fma(qq, s2r, acc.v[14]);
                        acc.v[15] = fma(qq, s2i, acc.v[15]);
                    } else {
                        // chart barycentrics f = (u1, u2, 1 - u1 - u2) = F0 + xi F1
                        // with F0 = (1, 0, 0), F1 = (b1, b2, -b1 - b2); g alike
                        const double F1[3] = {b1, b2, -(b1 + b2)}, G1[3] = {d1, d2, -(d1 + d2)};
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int j = 0; j < 3; ++j) {
                                const double c2 = F1[i] * G1[j];
                                double &ar = acc.v[2 * (3 * i + j)], &ai = acc.v[2 * (3 * i + j) + 1];
                                ar = fma(c2, s2r, ar);
                                ai = fma(c2, s2i, ai);
                                if (i == 0 || j == 0) {
                                    const double c1 = (i == 0 ? G1[j] : 0.0) + (j == 0 ? F1[i] : 0.0);
                                    ar = fma(c1, s1r, ar);
                                    ai = fma(c1, s1i, ai);
                                }
                                if (i == 0 && j == 0) {
                                    ar += s0r;
                                    ai += s0i;
                                }
                            }
                    }
                }
            };
            using I0 = std::integral_constant<int, 0>;
            using I1 = std::integral_constant<int, 1>;
            using I2 = std::integral_constant<int, 2>;
            using I3 = std::integral_constant<int, 3>;
            using I4 = std::integral_constant<int, 4>;
            using I5 = std::integral_constant<int, 5>;
            using CCV = std::integral_constant<int, CLS_CV>;
            using CCE = std::integral_constant<int, CLS_CE>;
            using CCF = std::integral_constant<int, CLS_CF>;
            if (cls == CLS_CV) {
                region(CCV{}, I0{});
                region(CCV{}, I1{});
            } else if (cls == CLS_CE) {
                region(CCE{}, I0{});
                region(CCE{}, I1{});
                region(CCE{}, I2{});
                region(CCE{}, I3{});
                region(CCE{}, I4{});
            } else if (OP != OPK_DL && a.sym) {
                // regions 2m + 1 are the x <-> y swaps of regions 2m: added
                // after the reduction
                region(CCF{}, I0{});
                region(CCF{}, I2{});
                region(CCF{}, I4{});
            } else {
                region(CCF{}, I0{});
                region(CCF{}, I1{});
                region(CCF{}, I2{});
                region(CCF{}, I3{});
                region(CCF{}, I4{});
                region(CCF{}, I5{});
            }
        } else {
            const int g = cls - 1;
            const int n = a.ss_n[g];
            const int n2 = n * n, n4 = n2 * n2;
            const int nreg = cls == CLS_CV ? 2 : (cls == CLS_CE ? 5 : 6);
            // common face of a symmetric kernel: regions 2m + 1 are the swaps
            // of regions 2m, evaluated with them
            const bool pair = OP != OPK_DL && cls == CLS_CF && a.sym;
            // region loop outermost: the region map is uniform over the wave
            for (int r = 0; r < nreg; r += pair ? 2 : 1) {
                // U nodes per lane and iteration, so that independent
                // evaluations overlap their latencies (the double layer's
                // registers allow one); past the last node the extra ones
                // repeat it with weight zero
                constexpr int U = OP == OPK_DL ? 1 : 2;
#pragma unroll 1
                for (int node = lane; node < n4; node += 64 * U) {
#pragma unroll
                    for (int h2 = 0; h2 < U; ++h2) {
                        const int nd = min(node + 64 * h2, n4 - 1);
                        const int i1 = nd / (n2 * n), r1 = nd - i1 * n2 * n;
                        const int i2 = r1 / n2, r2 = r1 - i2 * n2;
                        const int i3 = r2 / n, i4 = r2 - i3 * n;
                        const double wgt = (node + 64 * h2 < n4 ? 1.0 : 0.0) * s_glw[g][i1] * s_glw[g][i2] *
                                           s_glw[g][i3] * s_glw[g][i4];
                        double u1, u2, v1, v2, m;
                        ss_map(cls, r, s_glx[g][i1], s_glx[g][i2], s_glx[g][i3], s_glx[g][i4], wgt, u1, u2, v1, v2, m);
                        if (pair) spec_point<OP, false, true>(a.op, lossy, s_sc, acc, X, Y, dc, u1, u2, v1, v2, m, nrm);
                        else spec_point<OP>(a.op, lossy, s_sc, acc, X, Y, dc, u1, u2, v1, v2, m, nrm);
                    }
                }
            }
        }
        // the sums over the pair's lanes, to each of them (wave_sum_all);
        // lane l < 9 of the pair then forms entry l of the chart-order 3 x 3
        // matrix
        wave_sum_all<spec_nacc(OP), LPP>(acc.v);
        if constexpr (NG != 0 && !NEARK && OP == OPK_MW) {
            // radial form: the moments of x = P + xi Q and y = P' + xi Q'
            // from the lane sums (S_0, Q S_1, Q' S_1, (Q.Q') S_2), P and P'
            // the singular point relative to the centroids; paired common
            // faces add the x <-> y swaps: M0 and Mxy twice, Mx = My = Mx + My
            const bool pr = cls == CLS_CF && a.sym;
            const double a0r = acc.v[0], a0i = acc.v[1];
            double xyr = acc.v[14], xyi = acc.v[15];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double px = X.o[d] + X.t1[d], py = Y.o[d] + Y.t1[d];
                const double axr = acc.v[2 + d], axi = acc.v[5 + d], ayr = acc.v[8 + d], ayi = acc.v[11 + d];
                xyr = fma(px, fma(py, a0r, ayr), fma(py, axr, xyr));
                xyi = fma(px, fma(py, a0i, ayi), fma(py, axi, xyi));
                const double mxr = fma(px, a0r, axr), mxi = fma(px, a0i, axi);
                const double myr = fma(py, a0r, ayr), myi = fma(py, a0i, ayi);
                acc.v[2 + d] = pr ? mxr + myr : mxr;
                acc.v[5 + d] = pr ? mxi + myi : mxi;
                acc.v[8 + d] = pr ? mxr + myr : myr;
                acc.v[11 + d] = pr ? mxi + myi : myi;
            }
            const double sc = pr ? 2.0 : 1.0;
            acc.v[0] = sc * a0r;
            acc.v[1] = sc * a0i;
            acc.v[14] = sc * xyr;
            acc.v[15] = sc * xyi;
        } else if constexpr (NG != 0 && !NEARK && OP == OPK_DL) {
            // radial form: the P, P' parts of (n.y) r - (n.r) y and (n.r) x
            // are -P' M1 and P M1
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double px = X.o[d] + X.t1[d], py = Y.o[d] + Y.t1[d];
                acc.v[6 + d] = fma(-py, acc.v[0], acc.v[6 + d]);
                acc.v[9 + d] = fma(-py, acc.v[1], acc.v[9 + d]);
                acc.v[12 + d] = fma(px, acc.v[0], acc.v[12 + d]);
                acc.v[15 + d] = fma(px, acc.v[1], acc.v[15 + d]);
            }
        } else if constexpr (NG != 0 && !NEARK && OP == OPK_HH3D) {
            // radial form, paired common faces: the swaps add the transposed
            // shape products, f_i g_j + g_i f_j
            if (cls == CLS_CF && a.sym) {
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = i; j < 3; ++j)
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            const double v = acc.v[2 * (3 * i + j) + c] + acc.v[2 * (3 * j + i) + c];
                            acc.v[2 * (3 * i + j) + c] = v;
                            acc.v[2 * (3 * j + i) + c] = v;
                        }
            }
        } else if (OP == OPK_MW && !NEARK && cls == CLS_CF && a.sym) {
#pragma unroll
            for (int i = 0; i < 6; ++i) acc.v[8 + i] = acc.v[2 + i];  // paired regions: moments of y = those of x
        }
        const double *red = acc.v;
        if (valid && pl < 9) {
            // chart vertices c (test) and c2 (trial) relative to the
            // centroids (o + t1, o + t2, o); entry (c, c2) belongs to the
            // local shapes (PT[c], PS[c2])
            const int c = pl / 3, c2 = pl - 3 * (pl / 3);
            double Xc[3], Yc[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Xc[d] = X.o[d] + (c == 0 ? X.t1[d] : (c == 1 ? X.t2[d] : 0.0));
                Yc[d] = Y.o[d] + (c2 == 0 ? Y.t1[d] : (c2 == 1 ? Y.t2[d] : 0.0));
            }
            const int dst = 3 * (c == 0 ? PT[0] : (c == 1 ? PT[1] : PT[2])) + (c2 == 0 ? PS[0] : (c2 == 1 ? PS[1] : PS[2]));
            double2 z;
            if constexpr (OP == OPK_MW) {
                // moments M0, Mx, My, Mxy (real, imaginary)
                z = mw_entry(a.op, red, Xc, Yc);
            } else {
                double pr, pi;
                if constexpr (OP == OPK_DL) {
                    // M25 + X.M63 - Y.(M4 - X M1) + (n.Y)(M7 - X.M8)
                    double tr = red[2], ti = red[3];
                    double qr = red[4], qi = red[5];
                    double ur[3], ui[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        tr = fma(Xc[d], red[6 + d], tr);
                        ti = fma(Xc[d], red[9 + d], ti);
                        ur[d] = fma(-Xc[d], red[0], red[12 + d]);
                        ui[d] = fma(-Xc[d], red[1], red[15 + d]);
                        qr = fma(-Xc[d], red[18 + d], qr);
                        qi = fma(-Xc[d], red[21 + d], qi);
                    }
                    const double nY = nrm[0] * Yc[0] + nrm[1] * Yc[1] + nrm[2] * Yc[2];
                    pr = tr - (Yc[0] * ur[0] + Yc[1] * ur[1] + Yc[2] * ur[2]) + nY * qr;
                    pi = ti - (Yc[0] * ui[0] + Yc[1] * ui[1] + Yc[2] * ui[2]) + nY * qi;
                } else {
                    // red[2 lane], red[2 lane + 1] by selects (a register
                    // array indexed by the lane would go to scratch memory)
                    pr = red[0];
                    pi = red[1];
#pragma unroll
                    for (int k = 1; k < 9; ++k)
                        if (pl == k) {
                            pr = red[2 * k];
                            pi = red[2 * k + 1];
                        }
                }
                // Helmholtz: Jacobian 4 A_K A_L; double layer: it cancels the
                // 1/(2A) of the two RWG shapes
                const double scale = OP == OPK_DL ? 1.0 : 4.0 * a.X.area[K] * a.Y.area[L];
                const double sr = a.op.alpha_re * scale, si = a.op.alpha_im * scale;
                z.x = sr * pr - si * pi;
                z.y = sr * pi + si * pr;
            }
            a.spec_val[9 * (long long)e + dst] = z;
        }
    }
    }
}

// One thread per entry of a far pair that the mixed path keeps in FP64
// (class 5), owners and the other entries of the pair alike (computing those
// again costs less than k_link's search for their owner); thread u takes
// entry u - ebase[i] of problem i.
// FAST: the default 3 x 4 point rule of a lossless operator at compile time.
template <int OP, int FAST>
__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_far64(const SpecialArgs *args, const int *ebase, int nprob) {
    __shared__ double2 s_sc[64];
    __shared__ int s_items[BMI_BLOCK_THREADS];
    __shared__ int s_wsum[BMI_BLOCK_THREADS / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // the work-group's entries of class 5 packed to its first threads, so
    // that the waves past them exit at once (all of them, before the table,
    // where there are none)
    const int u0 = blockIdx.x * BMI_BLOCK_THREADS + t;
    bool mine = false;
    if (u0 < ebase[nprob]) {
        const int p0 = find_problem(ebase, nprob, u0);
        const int e0 = u0 - ebase[p0];
        mine = e0 < min(*args[p0].n_special, args[p0].capacity) &&
               (args[p0].spec_cls[e0] & ~SPEC_NOT_OWNER) == CLS_FAR64;
    }
    const unsigned long long m = __ballot(mine);
    if (lane == 0) s_wsum[wv] = __popcll(m);
    __syncthreads();
    int base = 0, n = 0;
#pragma unroll
    for (int i = 0; i < BMI_BLOCK_THREADS / 64; ++i) {
        base += i < wv ? s_wsum[i] : 0;
        n += s_wsum[i];
    }
    if (n == 0) return;
    if (mine) s_items[base + __popcll(m & ((1ull << lane) - 1ull))] = u0;
    load_sincos_table(s_sc, args[0].sc);
    __syncthreads();
    if (t >= n) return;
    const int u = s_items[t];
    const int pi = find_problem(ebase, nprob, u);
    const SpecialArgs &a = args[pi];
    {
        const int e = u - ebase[pi];
        const int K = a.spec_K[e], L = a.spec_L[e];
        double xp[3 * BMI_MAX_RULE], yp[3 * BMI_MAX_RULE], xa[9], ya[9], xc[3], yc[3];
#pragma unroll
        for (int v = 0; v < 3 * BMI_MAX_RULE; ++v) {
            xp[v] = a.X.pts[3 * BMI_MAX_RULE * (long long)K + v];
            yp[v] = a.Y.pts[3 * BMI_MAX_RULE * (long long)L + v];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xc[d] = a.X.cen[3 * K + d];
            yc[d] = a.Y.cen[3 * L + d];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            xa[v] = a.X.vtx[9 * K + v] - xc[v % 3];
            ya[v] = a.Y.vtx[9 * L + v] - yc[v % 3];
        }
        double2 loc[9];
        pair_fp64<OP, FAST, true>(a.op, a.rule, xp, xa, xc, a.X.area[K], yp, ya, yc, a.Y.area[L], loc, s_sc);
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            double2 z = loc[v];
            if constexpr (FAST) {  // the compile-time-rule pair gives the entries divided by alpha
                z.x = fma(a.op.alpha_re, loc[v].x, -a.op.alpha_im * loc[v].y);
                z.y = fma(a.op.alpha_re, loc[v].y, a.op.alpha_im * loc[v].x);
            }
            a.spec_val[9 * (long long)e + v] = z;
        }
    }
}

// One thread per entry that k_special and k_far64 leave: zero common faces
// (cf_zero), common-vertex entries (K, L) with K > L of a symmetric problem
// (the transpose of the block of (L, K)), and non-owner entries, whose pair
// sits in the block (home sub-batch of the test cell, patch of the trial
// cell).  The source entry is found in its block by binary search: a block's
// entries are in the order of their pair's thread kl * 16 + ll, kl the test
// cell's place in its sub-batch and ll the trial cell's place in its patch.
// Class-5 entries are all k_far64's.  A missing source (which the classifier rules out: both blocks classify the
// same pair with the same arithmetic, and the vertex test is symmetric)
// leaves NaN.
__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_link(const SpecialArgs *args, const int *ebase, int nprob) {
    const int u = blockIdx.x * BMI_BLOCK_THREADS + threadIdx.x;
    if (u >= ebase[nprob]) return;
    const int pi = find_problem(ebase, nprob, u);
    const SpecialArgs &a = args[pi];
    const int ns = min(*a.n_special, a.capacity);
    const int e = u - ebase[pi];
    if (e >= ns) return;
    const int raw = a.spec_cls[e], cls = raw & ~SPEC_NOT_OWNER;
    if (cls == CLS_FAR64) return;  // k_far64 integrates every entry of the pair, owner or not
    if (cls == CLS_CF && a.cf_zero) {
#pragma unroll
        for (int v = 0; v < 9; ++v) a.spec_val[9 * (long long)e + v] = make_double2(0.0, 0.0);
        return;
    }
    const int K = a.spec_K[e], L = a.spec_L[e];
    const bool mirror = a.sym && cls == CLS_CV && K > L;
    if (!(raw & SPEC_NOT_OWNER) && !mirror) return;
    // the source pair: (K, L) in its owner entry, or (L, K)
    const int S = mirror ? L : K, T = mirror ? K : L;
    const int hs = a.X.home[S], hp = a.Y.home[T];
    const long long b = (long long)hs * a.npb + hp;
    const int o0 = a.blk_off[b], o1 = min(o0 + a.blk_cnt[b], a.capacity);
    int kl = 0, ll = a.spec_loc[e] & (BMI_PATCH - 1);  // a non-owner's trial cell sits in the same patch
    for (int k = 0; k < BMI_PATCH; ++k)
        if (a.X.patch_tri[(long long)hs * BMI_PATCH + k] == S) kl = k;
    if (mirror)
        for (int k = 0; k < BMI_PATCH; ++k)
            if (a.Y.patch_tri[(long long)hp * BMI_PATCH + k] == T) ll = k;
    const int want = kl * BMI_PATCH + ll;
    int lo = o0, hi = o1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a.spec_loc[mid] < want) lo = mid + 1;
        else hi = mid;
    }
    const int o = lo < o1 && a.spec_K[lo] == S && a.spec_L[lo] == T ? lo : -1;
    const double2 nan2 = make_double2(__longlong_as_double(0x7ff8000000000000LL),
                                      __longlong_as_double(0x7ff8000000000000LL));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            a.spec_val[9 * (long long)e + 3 * i + j] =
                o >= 0 ? a.spec_val[9 * (long long)o + (mirror ? 3 * j + i : 3 * i + j)] : nan2;
}

__global__ void k_sincos_table(double2 *tab) { fill_sincos_table(tab); }

hipError_t launch_sincos_table(double2 *tab, hipStream_t s) {
    hipLaunchKernelGGL(k_sincos_table, dim3(1), dim3(64), 0, s, tab);
    return hipGetLastError();
}

hipError_t launch_special(const SpecialArgs *args, const int *cbase, const int *ebase, int nprob, int total_chunks,
                          int total_entries, int cs, int mixed, int op_kind, int ng5, int fast, hipStream_t s) {
    if (total_entries <= 0) return hipSuccess;
    const dim3 b(BMI_BLOCK_THREADS);
    const dim3 g(total_chunks);
    const int g2 = (total_entries + BMI_BLOCK_THREADS - 1) / BMI_BLOCK_THREADS;
    // the short kernels first (in a batch they then share the device with the
    // other operators' Sauter-Schwab kernels rather than their far kernels)
    if (mixed && fast) {
        if (op_kind == OPK_MW) hipLaunchKernelGGL((k_far64<OPK_MW, 1>), dim3(g2), b, 0, s, args, ebase, nprob);
        else if (op_kind == OPK_DL) hipLaunchKernelGGL((k_far64<OPK_DL, 1>), dim3(g2), b, 0, s, args, ebase, nprob);
        else hipLaunchKernelGGL((k_far64<OPK_HH3D, 1>), dim3(g2), b, 0, s, args, ebase, nprob);
    } else if (mixed) {
        if (op_kind == OPK_MW) hipLaunchKernelGGL((k_far64<OPK_MW, 0>), dim3(g2), b, 0, s, args, ebase, nprob);
        else if (op_kind == OPK_DL) hipLaunchKernelGGL((k_far64<OPK_DL, 0>), dim3(g2), b, 0, s, args, ebase, nprob);
        else hipLaunchKernelGGL((k_far64<OPK_HH3D, 0>), dim3(g2), b, 0, s, args, ebase, nprob);
    }
    if (op_kind == OPK_MW) hipLaunchKernelGGL((k_special<OPK_MW, 0, true>), g, b, 0, s, args, cbase, nprob, cs);
    else if (op_kind == OPK_DL) hipLaunchKernelGGL((k_special<OPK_DL, 0, true>), g, b, 0, s, args, cbase, nprob, cs);
    else hipLaunchKernelGGL((k_special<OPK_HH3D, 0, true>), g, b, 0, s, args, cbase, nprob, cs);
    if (ng5) {
        if (op_kind == OPK_MW) hipLaunchKernelGGL((k_special<OPK_MW, 5, false>), g, b, 0, s, args, cbase, nprob, cs);
        else if (op_kind == OPK_DL) hipLaunchKernelGGL((k_special<OPK_DL, 5, false>), g, b, 0, s, args, cbase, nprob, cs);
        else hipLaunchKernelGGL((k_special<OPK_HH3D, 5, false>), g, b, 0, s, args, cbase, nprob, cs);
    } else {
        if (op_kind == OPK_MW) hipLaunchKernelGGL((k_special<OPK_MW, 0, false>), g, b, 0, s, args, cbase, nprob, cs);
        else if (op_kind == OPK_DL) hipLaunchKernelGGL((k_special<OPK_DL, 0, false>), g, b, 0, s, args, cbase, nprob, cs);
        else hipLaunchKernelGGL((k_special<OPK_HH3D, 0, false>), g, b, 0, s, args, cbase, nprob, cs);
    }
    hipLaunchKernelGGL(k_link, dim3(g2), b, 0, s, args, ebase, nprob);
    return hipGetLastError();
}

// Entries per chunk of k_special: enough chunks to fill the device
// (>= 8192 work-groups, so the last round of work-groups is a small part)
// down to 16 entries per chunk.
int special_chunk(int capacity) {
    int cs = 256;
    while (cs > 16 && (capacity + cs - 1) / cs < 8192) cs >>= 1;
    return cs;
}

// ------------------------------------------------------ far-pair blocks

// The write-back's matrix accesses, non-temporal: each block's entries are
// stored once (or read back once, by the launch of a later colour), so they
// stream past the caches instead of displacing the records every block reads
// (far phase about 2 % faster).
typedef double bmi_d2v __attribute__((ext_vector_type(2)));
typedef float bmi_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(double2 *p, double2 v) {
    const bmi_d2v x = {v.x, v.y};
    __builtin_nontemporal_store(x, (bmi_d2v *)p);
}
__device__ __forceinline__ void store_stream(float2 *p, float2 v) {
    const bmi_f2v x = {v.x, v.y};
    __builtin_nontemporal_store(x, (bmi_f2v *)p);
}
__device__ __forceinline__ double2 load_stream(const double2 *p) {
    const bmi_d2v x = __builtin_nontemporal_load((const bmi_d2v *)p);
    return make_double2(x.x, x.y);
}
__device__ __forceinline__ float2 load_stream(const float2 *p) {
    const bmi_f2v x = __builtin_nontemporal_load((const bmi_f2v *)p);
    return make_float2(x.x, x.y);
}

__device__ __forceinline__ void out_write(double2 *Z, long long off, double re, double im, bool first) {
    double2 v;
    if (first) {
        v.x = re;
        v.y = im;
    } else {
        v = Z[off];
        v.x += re;
        v.y += im;
    }
    Z[off] = v;
}

__device__ __forceinline__ void out_write(float2 *Z, long long off, double re, double im, bool first) {
    float2 v;
    if (first) {
        v.x = (float)re;
        v.y = (float)im;
    } else {
        v = Z[off];
        v.x = (float)((double)v.x + re);
        v.y = (float)((double)v.y + im);
    }
    Z[off] = v;
}

// One work-group per (test panel P, trial patch q).  The panel owns output
// rows [64 P, 64 P + 64) completely: its cells (all cells carrying one of
// those functions) are integrated against the 16 cells of q in sub-batches of
// 16, and every output entry (row of P, column of q) is accumulated in a
// register of the thread that owns it.  Launches go trial colour by trial
// colour: inside one launch no two work-groups share an output entry (rows
// belong to one panel, columns of one colour to one patch), so there are no
// atomics; a column is stored by the launch of its first trial patch and
// read-modify-written by the launch of its second.
//   per sub-batch:  phase 1  thread (k, l) integrates cell pair (sub-batch[k], q[l]) -> LDS
//                   phase 2  special pairs of the block overwrite their slots
//                   phase 3  each owned entry adds its slot contributions, in a fixed order
//   at the end:     each wave writes whole 64-row column segments (1 KB, contiguous)
// The records of sub-batch s + 1 are read into registers while s is being
// integrated.  Thread t owns row t % 64 and the columns t / 64 + 4 j, j < 12.
template <int OP, int MIXED, int FAST, typename OutT>
__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_panels(PanelArgs a) {
    __shared__ double2 V[9 * BMI_BLOCK_THREADS];  // V[v * 256 + kl * 16 + ll]
    __shared__ double sg[BMI_GEO];                // test sub-batch geometry
    __shared__ double bg[BMI_GEO];                // trial patch geometry
    __shared__ int s_cfn[BMI_MAX_ROWS];
    __shared__ unsigned char s_cfirst[BMI_MAX_ROWS];
    __shared__ int s_cs[BMI_MAX_ROWS + 1];
    __shared__ unsigned char s_ckl[BMI_MAX_ROWS];
    __shared__ double s_ccf[BMI_MAX_ROWS];
    __shared__ int s_rs[BMI_PANEL_ROWS + 1];
    __shared__ unsigned char s_rkl[BMI_MAX_ROWS];
    __shared__ double s_rcf[BMI_MAX_ROWS];
    __shared__ double2 s_sc[64];

    const int t = threadIdx.x;
    if (!MIXED) load_sincos_table(s_sc, a.sc);
    const int w = blockIdx.x;
    const int pl = w / a.nbq;  // panel of the launch's range
    const int P = a.panel0 + pl;
    const int q = a.bq_list[w - pl * a.nbq];
    const int r0 = a.prow[P];
    const int nr = a.prow[P + 1] - r0;
    const int nc = a.bnc[q];
    const int myr = t & (BMI_PANEL_ROWS - 1);
    const int myc = t >> 6;
    const int kl = t >> 4, ll = t & 15;
    const int sb0 = a.sb_off[P], sb1 = a.sb_off[P + 1];

    // trial patch records (once per work-group)
    {
        const double *g = a.bgeo + (long long)q * BMI_GEO;
        bg[t] = g[t];
        if (t < BMI_GEO - BMI_BLOCK_THREADS) bg[BMI_BLOCK_THREADS + t] = g[BMI_BLOCK_THREADS + t];
        if (t < BMI_MAX_ROWS) {
            s_cfn[t] = a.bfn[q * BMI_MAX_ROWS + t];
            s_cfirst[t] = a.bfirst[q * BMI_MAX_ROWS + t];
            s_ckl[t] = a.bkl[q * BMI_MAX_ROWS + t];
            s_ccf[t] = a.bcf[q * BMI_MAX_ROWS + t];
        }
        if (t <= BMI_MAX_ROWS) s_cs[t] = a.bcs[q * (BMI_MAX_ROWS + 1) + t];
    }

    // registers holding the next sub-batch's records
    double pg0 = 0.0, pg1 = 0.0, pcf = 0.0;
    int prs = 0, pkl = 0, pe0 = 0, pe1 = 0;
    auto fetch = [&](int s) {
        const double *g = a.tgeo + (long long)s * BMI_GEO;
        pg0 = g[t];
        if (t < BMI_GEO - BMI_BLOCK_THREADS) pg1 = g[BMI_BLOCK_THREADS + t];
        if (t <= BMI_PANEL_ROWS) prs = a.trs[(long long)s * (BMI_PANEL_ROWS + 1) + t];
        if (t < BMI_MAX_ROWS) {
            pkl = a.tkl[(long long)s * BMI_MAX_ROWS + t];
            pcf = a.tcf[(long long)s * BMI_MAX_ROWS + t];
        }
        const long long b = (long long)s * a.npb + q;
        pe0 = a.blk_off[b];
        pe1 = pe0 + a.blk_cnt[b];
    };
    if (sb0 < sb1) fetch(sb0);

    double accr[BMI_PANEL_ACC], acci[BMI_PANEL_ACC];
#pragma unroll
    for (int j = 0; j < BMI_PANEL_ACC; ++j) accr[j] = acci[j] = 0.0;

    for (int s = sb0; s < sb1; ++s) {
        __syncthreads();  // the previous sub-batch is done with sg, V and the row tables
        sg[t] = pg0;
        if (t < BMI_GEO - BMI_BLOCK_THREADS) sg[BMI_BLOCK_THREADS + t] = pg1;
        if (t <= BMI_PANEL_ROWS) s_rs[t] = prs;
        if (t < BMI_MAX_ROWS) {
            s_rkl[t] = (unsigned char)pkl;
            s_rcf[t] = pcf;
        }
        const int e0 = pe0, e1 = min(pe1, a.capacity);
        if (s + 1 < sb1) fetch(s + 1);
        __syncthreads();

        // phase 1
        {
            double2 loc[9];
            const double xs = sg[BMI_GEO_AREA + kl], ys = bg[BMI_GEO_AREA + ll];
            if (xs > 0.0 && ys > 0.0) {
                const double *xp = &sg[BMI_GEO_PTS + 3 * BMI_MAX_RULE * kl];
                const double *xc = &sg[BMI_GEO_CEN + 3 * kl];
                const double *xa = &sg[BMI_GEO_VREL + 9 * kl];
                const double *yp = &bg[BMI_GEO_PTS + 3 * BMI_MAX_RULE * ll];
                const double *yc = &bg[BMI_GEO_CEN + 3 * ll];
                const double *ya = &bg[BMI_GEO_VREL + 9 * ll];
                if (MIXED) pair_mixed<OP, FAST>(a.op, a.rule, xp, xa, xc, xs, yp, ya, yc, ys, loc);
                else pair_fp64<OP, FAST, true>(a.op, a.rule, xp, xa, xc, xs, yp, ya, yc, ys, loc, s_sc);
            } else {
#pragma unroll
                for (int v = 0; v < 9; ++v) loc[v] = make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int v = 0; v < 9; ++v) V[v * BMI_BLOCK_THREADS + t] = loc[v];
        }

        // phase 2 (blocks next to the diagonal only); FAST: the pair values are
        // divided by alpha * fast_scale, so are the special ones
        if (e1 > e0) {
            __syncthreads();
            for (int e = e0 + t; e < e1; e += BMI_BLOCK_THREADS) {
                const int lo = a.spec_loc[e];
#pragma unroll
                for (int v = 0; v < 9; ++v) {
                    double2 z = a.spec_val[9 * (long long)e + v];
                    if constexpr (FAST) {
                        constexpr double is = 1.0 / fast_scale<OP, MIXED, FAST>();
                        z = make_double2(is * fma(a.op.ia_re, z.x, -a.op.ia_im * z.y),
                                         is * fma(a.op.ia_re, z.y, a.op.ia_im * z.x));
                    }
                    V[v * BMI_BLOCK_THREADS + lo] = z;
                }
            }
        }
        __syncthreads();

        // phase 3
        const int u0 = s_rs[myr], u1 = s_rs[myr + 1];
        if (u0 < u1) {
#pragma unroll
            for (int j = 0; j < BMI_PANEL_ACC; ++j) {
                const int c = myc + 4 * j;
                if (c < nc) {
                    double sr = accr[j], si = acci[j];
                    const int v0 = s_cs[c], v1 = s_cs[c + 1];
                    for (int u = u0; u < u1; ++u) {
                        const int ki = s_rkl[u];
                        const int k = ki / 3, i = ki - 3 * k;
                        const double cf = s_rcf[u];
                        for (int v = v0; v < v1; ++v) {
                            const int lj = s_ckl[v];
                            const int l = lj / 3, jj = lj - 3 * l;
                            const double cc = cf * s_ccf[v];
                            const double2 z = V[(3 * i + jj) * BMI_BLOCK_THREADS + k * BMI_PATCH + l];
                            sr = fma(cc, z.x, sr);
                            si = fma(cc, z.y, si);
                        }
                    }
                    accr[j] = sr;
                    acci[j] = si;
                }
            }
        }
    }

    // write-back: a wave stores rows r0 .. r0 + 63 of one column
    __syncthreads();
    if (myr < nr) {
        OutT *Z = (OutT *)a.Z;
#pragma unroll
        for (int j = 0; j < BMI_PANEL_ACC; ++j) {
            const int c = myc + 4 * j;
            if (c < nc) {
                const long long off = (long long)(r0 + myr) + (long long)s_cfn[c] * a.ld;
                double vr = accr[j], vi = acci[j];
                if constexpr (FAST) {
                    constexpr double ks = fast_scale<OP, MIXED, FAST>();
                    vr = fma(ks * a.op.alpha_re, accr[j], -(ks * a.op.alpha_im) * acci[j]);
                    vi = fma(ks * a.op.alpha_re, acci[j], (ks * a.op.alpha_im) * accr[j]);
                }
                out_write(Z, off, vr, vi, s_cfirst[c]);
            }
        }
    }
}

// LDS of the rank-round kernels besides the tile.
struct PanelShared {
    double sg[2][BMI_GEO];             // test sub-batch geometry, double-buffered
    double bg[BMI_GEO];                // trial patch geometry
    int s_tslot[2][BMI_MAX_ROWS];
    double s_tscf[2][BMI_MAX_ROWS];
    int s_cfn[BMI_MAX_ROWS];
    unsigned char s_cfirst[BMI_MAX_ROWS];
    double2 s_sc[64];                  // sincos table (FP64 path)
    short s_ent[BMI_BLOCK_THREADS];    // special entry of each thread's pair (offset in the block), -1 none
};

// Tile entry (row r, column c) of a panel block: column-major, 64 rows per
// column, the row index XOR-ed with c mod 16, so that the 16 lanes of a rank
// round (one row, up to 16 columns) fall into different LDS banks while each
// column stays one contiguous segment of 64 entries.
template <int PR>
__device__ __forceinline__ int tile_ix(int r, int c) { return c * PR + (r ^ (c & 15)); }
// Tile entry type: FP64 for the FP64 path.  The mixed path keeps its tile in
// FP32: a tile entry sums the cell pairs of one (test, trial) function pair
// within one patch (at most 4 on RWG, the two vertex degrees' product on
// P1), each added by an FP64 multiply-add rounded once to FP32, which adds
// an error of the order of the FP32 evaluation's own; the 20 KB tile (not
// 40 KB) lets four work-groups share a CU.  Sums across patches are in the
// output type.
template <int MIXED> using tile_t = typename std::conditional<MIXED != 0, float2, double2>::type;

// Rows of a test panel and the widest trial patch of the operator's spaces:
// the Helmholtz operator works on Lagrange spaces (wide panels, narrow
// patches), the Maxwell operators on RWG.
__host__ __device__ constexpr int panel_rows(int op) { return op == OPK_HH3D ? BMI_PANEL_ROWS_WIDE : BMI_PANEL_ROWS; }
__host__ __device__ constexpr int panel_cols(int op) { return op == OPK_HH3D ? BMI_MAX_COLS_WIDE : BMI_MAX_COLS; }

// One (test panel P, trial patch q) block of the far-pair assembly with its
// write-back; the caller has filled sh.s_sc (FP64 path) and passes the tile
// of panel_rows(OP) x (widest patch) entries.  Ends after the write-back stores are
// issued; a caller that reuses the LDS for another block synchronises first.
//
// Rank rounds: each thread adds its pair's 9 values straight into the LDS
// tile of the work-group's 64 rows x the patch's columns.  Two pairs add to the same
// entry only if their test cells share the entry's row and their trial cells
// share its column; giving every slot its rank among the slots of its group
// with the same row (column), the additions go in rounds (rank_t, rank_b)
// with a barrier between rounds, so no entry is touched twice inside a round
// and the summation order is fixed.  RWG functions have at most two cells, so
// there are at most 2 x 2 rounds.  A trial patch runs only the rounds it
// needs; the Helmholtz kernels and the FP64 EFIE kernel with the fixed rule
// round by level instead of rank (below).
// REFETCH: recompute the lane offsets of the per-sub-batch record loads
// (the single-problem kernels; measured faster there, slower in the batched
// ones).
template <int OP, int MIXED, int FAST, typename OutT, bool REFETCH = false>
__device__ __forceinline__ void panel_block(const PanelArgs &a, int P, int q, tile_t<MIXED> *tile, PanelShared &sh) {
    auto &sg = sh.sg;
    auto &bg = sh.bg;
    auto &s_tslot = sh.s_tslot;
    auto &s_tscf = sh.s_tscf;
    auto &s_cfn = sh.s_cfn;
    auto &s_cfirst = sh.s_cfirst;
    const double2 *s_sc = sh.s_sc;
    const int t = threadIdx.x;
    constexpr int PR = panel_rows(OP);
    const int r0 = a.prow[P];
    const int nr = a.prow[P + 1] - r0;  // at most PR
    const int nc = a.bnc[q];
    const int kl = t >> 4, ll = t & 15;
    const int sb0 = a.sb_off[P], sb1 = a.sb_off[P + 1];

    // trial rounds by level (bslot bits 16-23) or by rank (bits 8-15): by level
    // for the Helmholtz kernels and the compile-time-rule kernels, whose
    // rounds read the three entries of one trial shape together
    constexpr bool BY_LEVEL = OP == OPK_HH3D || FAST;
    constexpr int RSH = BY_LEVEL ? 16 : 8;
    const int lane = t & 63, wv = t >> 6;
    for (int c = wv; c < nc; c += BMI_BLOCK_THREADS / 64)
#pragma unroll
        for (int h = 0; h < PR / 64; ++h) tile[c * PR + 64 * h + lane] = tile_t<MIXED>{};
    short *s_ent = sh.s_ent;
    s_ent[t] = -1;
    // MIXED: the rule points and vertices of a geometry record are kept in
    // FP32 in the front of the record's LDS (geo_put), centroids and areas in
    // FP64 in place
    auto geo_put = [&](double *rec, int i, double v) {
        if (!MIXED || (i >= BMI_GEO_CEN && i < BMI_GEO_VREL)) rec[i] = v;
        else ((float *)rec)[i < BMI_GEO_CEN ? i : BMI_GEOF_VREL + (i - BMI_GEO_VREL)] = (float)v;
    };
    {
        const double *g = a.bgeo + (long long)q * BMI_GEO;
        geo_put(bg, t, g[t]);
        if (t < BMI_GEO - BMI_BLOCK_THREADS) geo_put(bg, BMI_BLOCK_THREADS + t, g[BMI_BLOCK_THREADS + t]);
        if (t < BMI_MAX_ROWS) {
            s_cfn[t] = a.bfn[q * BMI_MAX_ROWS + t];
            s_cfirst[t] = a.bfirst[q * BMI_MAX_ROWS + t];
        }
    }
    // the three column slots of this thread's trial cell, for the whole work-group
    int bsl[3];
    double bcf[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        bsl[j] = a.bslot[q * BMI_MAX_ROWS + 3 * ll + j];
        bcf[j] = a.bscf[q * BMI_MAX_ROWS + 3 * ll + j];
    }

    // registers with the records of the next sub-batch
    double pg0 = 0.0, pg1 = 0.0, pcf = 0.0;
    int psl = 0xff, pe0 = 0, pcnt = 0;
    auto fetch = [&](int s) {
        const double *g = a.tgeo + (long long)s * BMI_GEO;
        pg0 = g[t];
        if (t < BMI_GEO - BMI_BLOCK_THREADS) pg1 = g[BMI_BLOCK_THREADS + t];
        if (t < BMI_MAX_ROWS) {
            // REFETCH: the lane offset is recomputed here (an empty asm hides
            // that it is loop-invariant); a per-lane 64-bit pointer kept
            // across the sub-batch loop costs two registers the FP64 kernels
            // lack (the compiler spilled it)
            int tt = t;
            if constexpr (REFETCH) asm volatile("" : "+v"(tt));
            psl = a.tslot[(long long)s * BMI_MAX_ROWS + tt];
            pcf = a.tscf[(long long)s * BMI_MAX_ROWS + tt];
        }
        const long long b = (long long)s * a.npb + q;
        pe0 = a.blk_off[b];
        pcnt = a.blk_cnt[b];  // added to pe0 where it is used: no wait here
    };
    auto stash = [&](int buf) {
        geo_put(sg[buf], t, pg0);
        if (t < BMI_GEO - BMI_BLOCK_THREADS) geo_put(sg[buf], BMI_BLOCK_THREADS + t, pg1);
        if (t < BMI_MAX_ROWS) {
            s_tslot[buf][t] = psl;
            s_tscf[buf][t] = pcf;
        }
    };
    int e0 = 0, e1 = 0;
    if (sb0 < sb1) {
        fetch(sb0);
        stash(0);
        e0 = pe0;
        e1 = min(pe0 + pcnt, a.capacity);
    }
    if (sb0 + 1 < sb1) fetch(sb0 + 1);
    __syncthreads();

    for (int s = sb0; s < sb1; ++s) {
        const int cur = (s - sb0) & 1;
        double2 loc[9];
        // the block's special entries (at most one per thread, in thread
        // order): each thread's entry by its pair, in LDS
        const bool has_spec = e1 > e0;  // uniform over the block
        if (has_spec && t < e1 - e0) s_ent[a.spec_loc[e0 + t]] = (short)t;
        const double xs = sg[cur][BMI_GEO_AREA + kl], ys = bg[BMI_GEO_AREA + ll];
        if (xs > 0.0 && ys > 0.0) {
            const double *xc = &sg[cur][BMI_GEO_CEN + 3 * kl];
            const double *yc = &bg[BMI_GEO_CEN + 3 * ll];
            if (MIXED) {
                const float *xf = (const float *)sg[cur], *yf = (const float *)bg;
                pair_mixed<OP, FAST, float>(a.op, a.rule, xf + 3 * BMI_MAX_RULE * kl, xf + BMI_GEOF_VREL + 9 * kl, xc,
                                            xs, yf + 3 * BMI_MAX_RULE * ll, yf + BMI_GEOF_VREL + 9 * ll, yc, ys, loc);
            } else {
                const double *xp = &sg[cur][BMI_GEO_PTS + 3 * BMI_MAX_RULE * kl];
                const double *xa = &sg[cur][BMI_GEO_VREL + 9 * kl];
                const double *yp = &bg[BMI_GEO_PTS + 3 * BMI_MAX_RULE * ll];
                const double *ya = &bg[BMI_GEO_VREL + 9 * ll];
                pair_fp64<OP, FAST, true>(a.op, a.rule, xp, xa, xc, xs, yp, ya, yc, ys, loc, s_sc);
            }
        }
        // a lane without a pair (padding, or a cell of zero area, which
        // contributes nothing) leaves loc unset and takes part in no round
        bool live = xs > 0.0 && ys > 0.0;
        // special pairs of this block take their precomputed values; the
        // slot is cleared for the next sub-batch's entries (written after
        // the round barrier below)
        if (has_spec) {
            __syncthreads();
            const int o = s_ent[t];
            if (o >= 0) {
#pragma unroll
                for (int v = 0; v < 9; ++v) {
                    const double2 z = a.spec_val[9 * (long long)(e0 + o) + v];
                    // FAST: the tile holds the entries divided by alpha * fast_scale
                    if constexpr (FAST) {
                        constexpr double is = 1.0 / fast_scale<OP, MIXED, FAST>();
                        loc[v] = make_double2(is * fma(a.op.ia_re, z.x, -a.op.ia_im * z.y),
                                              is * fma(a.op.ia_re, z.y, a.op.ia_im * z.x));
                    } else {
                        loc[v] = z;
                    }
                }
                s_ent[t] = -1;
                live = true;
            }
        }
        int tsl[3];
        double tcf[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            tsl[i] = s_tslot[cur][3 * kl + i];
            tcf[i] = s_tscf[cur][3 * kl + i];
        }
        // the next sub-batch goes to the other buffer (read last in the
        // previous iteration); the round barriers below publish it
        if (s + 1 < sb1) {
            stash(cur ^ 1);
            e0 = pe0;
            e1 = min(pe0 + pcnt, a.capacity);
            if (s + 2 < sb1) fetch(s + 2);
        }
        // With one test rank (the round-robin sub-batches), different test
        // cells of a sub-batch own different rows, so two updates of one
        // entry come from lanes of the same 16-lane group, i.e. of one wave:
        // the trial rounds need no barrier, as one wave's LDS accesses
        // are performed in program order.  More test ranks put colliding
        // rows in different waves: a barrier after each test round.
        // Trial rounds by level (the Helmholtz kernels and the kernels with
        // the compile-time rule): two lanes may add to one entry in one round
        // only through different local shapes j, and the updates go shape by
        // shape in program order, each shape's reads issued together (unused
        // slots read entry 0 and write nothing) after the previous shape's
        // writes.  Trial rounds by rank (the others): within a round a
        // thread's entries are distinct and no other thread touches them;
        // the updates go one at a time.
        const int nrb = a.brounds[q];
        for (int ra = 0; ra < a.rounds_t; ++ra) {
            bool tact[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) tact[i] = live && (tsl[i] & 0xff) != 0xff && (tsl[i] >> 8) == ra;
            for (int rb = 0; rb < nrb; ++rb) {
                if constexpr (FAST && BY_LEVEL) {
                    // by level: the three reads of one shape together
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const bool bj = (bsl[j] & 0xff) != 0xff && ((bsl[j] >> 16) & 0xff) == rb;
                        int ix[3];
                        bool act[3];
                        tile_t<MIXED> z[3];
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            act[i] = tact[i] && bj;
                            ix[i] = act[i] ? tile_ix<PR>(tsl[i] & 0xff, bsl[j] & 0xff) : 0;
                            z[i] = tile[ix[i]];
                        }
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            if (!act[i]) continue;
                            const double cc = tcf[i] * bcf[j];
                            z[i].x = fma(cc, loc[3 * i + j].x, (double)z[i].x);
                            z[i].y = fma(cc, loc[3 * i + j].y, (double)z[i].y);
                            tile[ix[i]] = z[i];
                        }
                    }
                } else {
                    // trial slot outermost: one rank test per column and
                    // round (the test slots' tests do not depend on the round)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        if ((bsl[j] & 0xff) == 0xff || ((bsl[j] >> RSH) & 0xff) != rb) continue;
                        const int col = bsl[j] & 0xff;
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            if (!tact[i]) continue;
                            const double cc = tcf[i] * bcf[j];
                            auto &z = tile[tile_ix<PR>(tsl[i] & 0xff, col)];
                            z.x = fma(cc, loc[3 * i + j].x, (double)z.x);
                            z.y = fma(cc, loc[3 * i + j].y, (double)z.y);
                        }
                    }
                }
            }
            // the next test round; after the last one, the next sub-batch's
            // records (stashed above) become visible and the tile is complete
            __syncthreads();
        }
    }

    // write-back, one column per wave at a time, lane l taking position l of
    // the column's segment (row l ^ (c mod 16)).  The stored values of all of
    // the wave's read-modify-write columns are loaded first, so that the
    // loads are in flight together, then added and stored.  A column is the
    // same for the whole wave: its function, first-touch flag and address
    // are read into scalar registers, so that each lane adds only its row.
    // (wide panels: each column in PR / 64 segments of 64 rows)
    OutT *Z = (OutT *)a.Z;
    constexpr int NH = PR / 64;
    constexpr int NCW = (panel_cols(OP) + BMI_BLOCK_THREADS / 64 - 1) / (BMI_BLOCK_THREADS / 64);
    const int wvu = __builtin_amdgcn_readfirstlane(wv);
    auto column = [&](int c) -> OutT * {
        return Z + (long long)__builtin_amdgcn_readfirstlane(s_cfn[c]) * a.ld + r0;
    };
    OutT old[NCW * NH];
#pragma unroll
    for (int k = 0; k < NCW; ++k) {
        const int c = wvu + k * (BMI_BLOCK_THREADS / 64);
#pragma unroll
        for (int h = 0; h < NH; ++h) old[NH * k + h] = OutT{};
        if (c < nc && !__builtin_amdgcn_readfirstlane(s_cfirst[c])) {
            const OutT *zc = column(c);
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const int r = 64 * h + (lane ^ (c & 15));
                if (r < nr) old[NH * k + h] = load_stream(zc + r);
            }
        }
    }
    // the tile holds the entries divided by alpha * fast_scale (FAST)
    constexpr double ks = fast_scale<OP, MIXED, FAST>();
    const double ar = ks * a.op.alpha_re, ai = ks * a.op.alpha_im;
#pragma unroll
    for (int k = 0; k < NCW; ++k) {
        const int c = wvu + k * (BMI_BLOCK_THREADS / 64);
        if (c < nc) {
            OutT *zc = column(c);
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const int r = 64 * h + (lane ^ (c & 15));
                if (r < nr) {
                    const auto tv = tile[c * PR + 64 * h + lane];
                    double2 v = make_double2(tv.x, tv.y);
                    if constexpr (FAST) v = make_double2(fma(ar, v.x, -ai * v.y), fma(ar, v.y, ai * v.x));
                    OutT o;
                    o.x = (decltype(o.x))((double)old[NH * k + h].x + v.x);
                    o.y = (decltype(o.y))((double)old[NH * k + h].y + v.y);
                    store_stream(zc + r, o);
                }
            }
        }
    }
}

// One work-group per (test panel P of the launch's range, trial patch q of its colour).
// Work-groups per CU: four on the mixed path (128 VGPRs, 20 KB FP32 tile;
// measured faster than three despite a few spilled registers), three on
// the FP64 path (the 40 KB tile and 161 VGPRs), two for the FP64 double layer.
template <int OP, int MIXED, int FAST, typename OutT>
__global__ __launch_bounds__(BMI_BLOCK_THREADS, MIXED ? 4 : OP == OPK_DL ? 2 : 3) void k_panels_rr(PanelArgs a) {
    extern __shared__ double2 tile[];             // panel_rows(OP) x a.ncm entries of tile_t<MIXED>
    __shared__ PanelShared sh;
    if (!MIXED) load_sincos_table(sh.s_sc, a.sc);
    const int w = blockIdx.x;
    const int pl = w / a.nbq;  // panel of the launch's range
    const int P = a.panel0 + pl;
    const int q = a.bq_list[w - pl * a.nbq];
    panel_block<OP, MIXED, FAST, OutT, true>(a, P, q, (tile_t<MIXED> *)tile, sh);
}

// Batched, one colour at a time: work-group i takes item i = (problem, panel,
// patch, -) of the launch's colour, across the problems of a batch.
// Work-groups per CU as in k_panels_rr.
template <int OP, int MIXED, int FAST, typename OutT>
__global__ __launch_bounds__(BMI_BLOCK_THREADS, MIXED ? 4 : OP == OPK_DL ? 2 : 3) void k_panels_items(
    const PanelArgs *args, const int4 *items) {
    extern __shared__ double2 tile[];  // panel_rows(OP) x (the batch's widest patch) entries of tile_t<MIXED>
    __shared__ PanelShared sh;
    const int4 it = items[blockIdx.x];
    if (!MIXED) load_sincos_table(sh.s_sc, args[it.x].sc);
    if constexpr (MIXED) {
        // a copy: the kernel's stores to the matrix cannot alias it, so its
        // fields are loaded once instead of after each store (the FP64
        // kernels lack the registers for it)
        const PanelArgs a = args[it.x];
        panel_block<OP, MIXED, FAST, OutT>(a, it.y, it.z, (tile_t<MIXED> *)tile, sh);
    } else {
        panel_block<OP, MIXED, FAST, OutT>(args[it.x], it.y, it.z, (tile_t<MIXED> *)tile, sh);
    }
}

// Batched, persistent and panel-major: the work-groups pop items (problem,
// test panel) from one queue, largest first, and each integrates its panel
// against every trial patch of its problem in colour order -- the order of
// the per-colour launches, so the values are the same bitwise.  A column
// shared by two patches is stored by the first of the two blocks and
// read-modify-written by the later one, both of this work-group, so the only
// ordering needed is the work-group barrier (with its workgroup-scope
// fences) between blocks: no work-group waits for another and no
// device-scope fence is needed.
template <int OP, int MIXED, int FAST, typename OutT>
__global__ __launch_bounds__(BMI_BLOCK_THREADS, (OP == OPK_DL && !MIXED) ? 2 : 3) void k_panels_persist(
    const PanelArgs *args, const int2 *items, int nitems, int *head) {
    extern __shared__ double2 tile[];
    __shared__ PanelShared sh;
    __shared__ int s_item;
    const int t = threadIdx.x;
    if (!MIXED) load_sincos_table(sh.s_sc, args[0].sc);
    for (;;) {
        __syncthreads();  // the previous block is done with the LDS and s_item; its stores are visible here
        if (t == 0) s_item = atomicAdd(head, 1);
        __syncthreads();
        const int i = s_item;
        if (i >= nitems) break;
        const int2 it = items[i];
        auto run = [&](const PanelArgs &a) {
            for (int u = 0; u < a.npb; ++u) {
                if (u > 0) __syncthreads();
                panel_block<OP, MIXED, FAST, OutT>(a, it.y, a.bq_all[u], (tile_t<MIXED> *)tile, sh);
            }
        };
        if constexpr (MIXED) {
            const PanelArgs a = args[it.x];  // as in k_panels_items
            run(a);
        } else {
            run(args[it.x]);
        }
    }
}

template <int OP, int MIXED, int FAST>
static void launch_panels_t(const PanelArgs &a, int c32, hipStream_t s) {
    const unsigned grid = (unsigned)((long long)a.npanel * a.nbq);
    // wide panels (the Helmholtz operator) only in the rank-round kernel
    if (a.rr || panel_rows(OP) != BMI_PANEL_ROWS) {
        // few rounds (RWG: at most 2; Lagrange by level: 2-4): rounds into an
        // LDS tile
        const size_t lds = sizeof(tile_t<MIXED>) * panel_rows(OP) * a.ncm;
        if (c32)
            hipLaunchKernelGGL((k_panels_rr<OP, MIXED, FAST, float2>), dim3(grid), dim3(BMI_BLOCK_THREADS), lds, s, a);
        else
            hipLaunchKernelGGL((k_panels_rr<OP, MIXED, FAST, double2>), dim3(grid), dim3(BMI_BLOCK_THREADS), lds, s, a);
        return;
    }
    if (c32)
        hipLaunchKernelGGL((k_panels<OP, MIXED, FAST, float2>), dim3(grid), dim3(BMI_BLOCK_THREADS), 0, s, a);
    else
        hipLaunchKernelGGL((k_panels<OP, MIXED, FAST, double2>), dim3(grid), dim3(BMI_BLOCK_THREADS), 0, s, a);
}

template <int OP, int MIXED>
static void launch_panels_m(const PanelArgs &a, int c32, hipStream_t s) {
    if (a.fast) launch_panels_t<OP, MIXED, 1>(a, c32, s);
    else launch_panels_t<OP, MIXED, 0>(a, c32, s);
}

hipError_t launch_panels(const PanelArgs &a, int op_kind, int mixed, int c32, hipStream_t s) {
    if (a.npanel <= 0 || a.nbq <= 0) return hipSuccess;
    if (op_kind == OPK_MW) {
        if (mixed) launch_panels_m<OPK_MW, 1>(a, c32, s);
        else launch_panels_m<OPK_MW, 0>(a, c32, s);
    } else if (op_kind == OPK_DL) {
        if (mixed) launch_panels_m<OPK_DL, 1>(a, c32, s);
        else launch_panels_m<OPK_DL, 0>(a, c32, s);
    } else {
        if (mixed) launch_panels_m<OPK_HH3D, 1>(a, c32, s);
        else launch_panels_m<OPK_HH3D, 0>(a, c32, s);
    }
    return hipGetLastError();
}

// ------------------------------------------------------ batched launches

namespace {
template <int OP, int MIXED, int FAST, typename OutT>
void items_t(const PanelArgs *args, const int4 *items, int nitems, int ncm, hipStream_t s) {
    const size_t lds = sizeof(tile_t<MIXED>) * panel_rows(OP) * ncm;
    hipLaunchKernelGGL((k_panels_items<OP, MIXED, FAST, OutT>), dim3(nitems), dim3(BMI_BLOCK_THREADS), lds, s, args,
                       items);
}

template <int OP, int MIXED, int FAST, typename OutT>
void persist_t(const PanelArgs *args, const int2 *items, int nitems, int ncm, int *head, int grid, hipStream_t s) {
    const size_t lds = sizeof(tile_t<MIXED>) * panel_rows(OP) * ncm;
    hipLaunchKernelGGL((k_panels_persist<OP, MIXED, FAST, OutT>), dim3(grid), dim3(BMI_BLOCK_THREADS), lds, s, args,
                       items, nitems, head);
}

// f(op, mixed, fast, c32) -> the instantiation
template <template <int, int, int, typename> class K, typename... A>
void dispatch4(int op_kind, int mixed, int fast, int c32, A... args) {
#define BMI_D(O, M, F)                                                                   \
    if (op_kind == O && mixed == M && fast == F) {                                       \
        if (c32) K<O, M, F, float2>::run(args...);                                       \
        else K<O, M, F, double2>::run(args...);                                          \
        return;                                                                          \
    }
    BMI_D(OPK_MW, 0, 0) BMI_D(OPK_MW, 0, 1) BMI_D(OPK_MW, 1, 0) BMI_D(OPK_MW, 1, 1)
    BMI_D(OPK_DL, 0, 0) BMI_D(OPK_DL, 0, 1) BMI_D(OPK_DL, 1, 0) BMI_D(OPK_DL, 1, 1)
    BMI_D(OPK_HH3D, 0, 0) BMI_D(OPK_HH3D, 0, 1) BMI_D(OPK_HH3D, 1, 0) BMI_D(OPK_HH3D, 1, 1)
#undef BMI_D
}

template <int OP, int MIXED, int FAST, typename OutT>
struct ItemsK {
    static void run(const PanelArgs *args, const int4 *items, int nitems, int ncm, hipStream_t s) {
        items_t<OP, MIXED, FAST, OutT>(args, items, nitems, ncm, s);
    }
};

template <int OP, int MIXED, int FAST, typename OutT>
struct PersistK {
    static void run(const PanelArgs *args, const int2 *items, int nitems, int ncm, int *head, int grid,
                    hipStream_t s) {
        persist_t<OP, MIXED, FAST, OutT>(args, items, nitems, ncm, head, grid, s);
    }
};

__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_fill_zero(uint4 *p, long long n16) {
    for (long long i = (long long)blockIdx.x * BMI_BLOCK_THREADS + threadIdx.x; i < n16;
         i += (long long)gridDim.x * BMI_BLOCK_THREADS)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_zero_each(int *const *ptrs, int n) {
    const int i = blockIdx.x * BMI_BLOCK_THREADS + threadIdx.x;
    if (i < n) *ptrs[i] = 0;
}

__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_zero_ints(int *p, int n) {
    for (int i = blockIdx.x * BMI_BLOCK_THREADS + threadIdx.x; i < n; i += gridDim.x * BMI_BLOCK_THREADS) p[i] = 0;
}
}  // namespace

hipError_t launch_panels_items(const PanelArgs *args, const int4 *items, int nitems, int ncm, int op_kind, int mixed,
                               int fast, int c32, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    dispatch4<ItemsK>(op_kind, mixed, fast, c32, args, items, nitems, ncm, s);
    return hipGetLastError();
}

hipError_t launch_panels_persist(const PanelArgs *args, const int2 *items, int nitems, int ncm, int *head, int grid,
                                 int op_kind, int mixed, int fast, int c32, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    dispatch4<PersistK>(op_kind, mixed, fast, c32, args, items, nitems, ncm, head, grid, s);
    return hipGetLastError();
}

// p must be 16-byte aligned and bytes a multiple of 8 (matrices are).
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
    const long long n16 = (long long)(bytes / 16);
    if (n16 > 0) {
        const long long g = std::min<long long>((n16 + BMI_BLOCK_THREADS - 1) / BMI_BLOCK_THREADS, 65536);
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)g), dim3(BMI_BLOCK_THREADS), 0, s, (uint4 *)p, n16);
    }
    if (bytes % 16) {  // an 8-byte tail
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(BMI_BLOCK_THREADS), 0, s, (int *)((char *)p + n16 * 16),
                           (int)((bytes % 16) / 4));
    }
    return hipGetLastError();
}

hipError_t launch_zero_each(int *const *ptrs, int n, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_zero_each, dim3((n + BMI_BLOCK_THREADS - 1) / BMI_BLOCK_THREADS), dim3(BMI_BLOCK_THREADS),
                           0, s, ptrs, n);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int *p, int n, hipStream_t s) {
    if (n > 0) {
        const int g = std::min((n + BMI_BLOCK_THREADS - 1) / BMI_BLOCK_THREADS, 4096);
        hipLaunchKernelGGL(k_zero_ints, dim3(g), dim3(BMI_BLOCK_THREADS), 0, s, p, n);
    }
    return hipGetLastError();
}

// ----------------------------------------------------- apply and compare

// y = Z x for up to BMI_MAX_RHS right-hand sides.  Grid: (row blocks of 256,
// column chunks); each thread owns one row of one chunk, reads its column
// entries coalesced across the wave (x staged in LDS BMI_APPLY_CHUNK columns
// at a time), and writes a partial sum; k_apply_reduce adds the chunk
// partials in chunk order (deterministic).
template <typename T>
__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_apply(const T *Z, long long ld, int rows, int cols,
                                                             const double2 *x, int nrhs, int chunk,
                                                             double2 *partial) {
    __shared__ double2 sx[BMI_APPLY_CHUNK * BMI_MAX_RHS];
    const int row = blockIdx.x * BMI_BLOCK_THREADS + threadIdx.x;
    const int c0 = blockIdx.y * chunk;
    const int c1 = min(c0 + chunk, cols);
    double ar[BMI_MAX_RHS], ai[BMI_MAX_RHS];
#pragma unroll
    for (int r = 0; r < BMI_MAX_RHS; ++r) ar[r] = ai[r] = 0.0;
    for (int s0 = c0; s0 < c1; s0 += BMI_APPLY_CHUNK) {
        const int s1 = min(s0 + BMI_APPLY_CHUNK, c1);
        __syncthreads();  // the previous sub-chunk's x is consumed
        for (int i = threadIdx.x; i < (s1 - s0) * nrhs; i += BMI_BLOCK_THREADS) {
            const int c = i / nrhs, r = i - c * nrhs;
            sx[i] = x[(long long)r * cols + s0 + c];
        }
        __syncthreads();
        if (row >= rows) continue;
        // eight columns' entries are loaded before any is used, so that each
        // wave keeps 8 KB of the matrix in flight
        constexpr int U = 8;
        for (int cb = s0; cb < s1; cb += U) {
            T zz[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // ComplexF64 entries stream past the caches (10 % faster);
                // ComplexF32 ones read slower that way (measured 2.7x)
                if constexpr (std::is_same<T, double2>::value)
                    zz[u] = cb + u < s1 ? load_stream(&Z[row + (long long)(cb + u) * ld]) : T{};
                else
                    zz[u] = cb + u < s1 ? Z[row + (long long)(cb + u) * ld] : T{};
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (cb + u >= s1) break;
                const double zr = (double)zz[u].x, zi = (double)zz[u].y;
                const double2 *xc = &sx[(cb + u - s0) * nrhs];
#pragma unroll
                for (int r = 0; r < BMI_MAX_RHS; ++r) {
                    if (r >= nrhs) break;
                    const double2 v = xc[r];
                    ar[r] = fma(zr, v.x, fma(-zi, v.y, ar[r]));
                    ai[r] = fma(zr, v.y, fma(zi, v.x, ai[r]));
                }
            }
        }
    }
    if (row >= rows) return;
    for (int r = 0; r < nrhs; ++r)
        partial[((long long)blockIdx.y * nrhs + r) * rows + row] = make_double2(ar[r], ai[r]);
}

// Column chunks of the apply: enough (row block, chunk) work-groups to fill
// the device (at least 2048), chunks of at least 256 columns.
int apply_chunks(int rows, int cols, int *chunk) {
    const int rb = (rows + BMI_BLOCK_THREADS - 1) / BMI_BLOCK_THREADS;
    const int most = std::max(1, (cols + 255) / 256);
    const int n = std::max(1, std::min(most, (2048 + rb - 1) / rb));
    *chunk = (cols + n - 1) / n;
    return (cols + *chunk - 1) / *chunk;
}

__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_apply_reduce(const double2 *partial, int nchunk, int rows,
                                                                    int nrhs, double2 *y) {
    const long long i = (long long)blockIdx.x * BMI_BLOCK_THREADS + threadIdx.x;
    if (i >= (long long)rows * nrhs) return;
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < nchunk; ++k) {
        const double2 v = partial[(long long)k * rows * nrhs + i];
        sr += v.x;
        si += v.y;
    }
    y[i] = make_double2(sr, si);
}

hipError_t launch_apply(const void *Z, int c32, long long ld, int rows, int cols, const double2 *x, int nrhs,
                        double2 *partial, double2 *y, hipStream_t s) {
    int chunk = 0;
    const int nchunk = apply_chunks(rows, cols, &chunk);
    dim3 grid((rows + BMI_BLOCK_THREADS - 1) / BMI_BLOCK_THREADS, nchunk);
    if (c32)
        hipLaunchKernelGGL(k_apply<float2>, grid, dim3(BMI_BLOCK_THREADS), 0, s, (const float2 *)Z, ld, rows, cols, x,
                           nrhs, chunk, partial);
    else
        hipLaunchKernelGGL(k_apply<double2>, grid, dim3(BMI_BLOCK_THREADS), 0, s, (const double2 *)Z, ld, rows, cols,
                           x, nrhs, chunk, partial);
    const long long n = (long long)rows * nrhs;
    hipLaunchKernelGGL(k_apply_reduce, dim3((unsigned)((n + BMI_BLOCK_THREADS - 1) / BMI_BLOCK_THREADS)),
                       dim3(BMI_BLOCK_THREADS), 0, s, partial, nchunk, rows, nrhs, y);
    return hipGetLastError();
}

// Sums of |a - b|^2, |b|^2 and max |a - b| over a matrix, per work-group;
// the host adds the per-group values in order.
template <typename TA, typename TB>
__global__ __launch_bounds__(BMI_BLOCK_THREADS) void k_compare(const TA *a, const TB *b, long long n,
                                                               double *out /* 3 per WG */) {
    __shared__ double s0[BMI_BLOCK_THREADS], s1[BMI_BLOCK_THREADS], s2[BMI_BLOCK_THREADS];
    const long long per = (n + (long long)gridDim.x - 1) / gridDim.x;
    const long long i0 = (long long)blockIdx.x * per;
    const long long i1 = min(i0 + per, n);
    double d2 = 0.0, r2 = 0.0, dm = 0.0;
    for (long long i = i0 + threadIdx.x; i < i1; i += BMI_BLOCK_THREADS) {
        const double ax = (double)a[i].x, ay = (double)a[i].y;
        const double bx = (double)b[i].x, by = (double)b[i].y;
        const double dx = ax - bx, dy = ay - by;
        const double dd = dx * dx + dy * dy;
        d2 += dd;
        r2 += bx * bx + by * by;
        dm = fmax(dm, dd);
    }
    s0[threadIdx.x] = d2;
    s1[threadIdx.x] = r2;
    s2[threadIdx.x] = dm;
    __syncthreads();
    for (int o = BMI_BLOCK_THREADS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x] = s0[0];
        out[3 * blockIdx.x + 1] = s1[0];
        out[3 * blockIdx.x + 2] = s2[0];
    }
}

hipError_t launch_compare(const void *a, int a32, const void *b, int b32, long long n, int groups, double *out,
                          hipStream_t s) {
    dim3 g(groups), t(BMI_BLOCK_THREADS);
    if (!a32 && !b32) hipLaunchKernelGGL((k_compare<double2, double2>), g, t, 0, s, (const double2 *)a, (const double2 *)b, n, out);
    else if (a32 && !b32) hipLaunchKernelGGL((k_compare<float2, double2>), g, t, 0, s, (const float2 *)a, (const double2 *)b, n, out);
    else if (!a32 && b32) hipLaunchKernelGGL((k_compare<double2, float2>), g, t, 0, s, (const double2 *)a, (const float2 *)b, n, out);
    else hipLaunchKernelGGL((k_compare<float2, float2>), g, t, 0, s, (const float2 *)a, (const float2 *)b, n, out);
    return hipGetLastError();
}
