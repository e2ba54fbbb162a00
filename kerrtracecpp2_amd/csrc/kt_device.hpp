// ============================================================================
//  kt_device.hpp -- per-ray physics of the KNdS geodesic tracer (HIP, gfx950).
//
//  Everything here is templated on the arithmetic type T (double or float) so
//  the binary64 and the mixed-precision kernels share one source.  Functions
//  are __host__ __device__ so the identical code can also be compiled for the
//  host by the test-only CPU check library.
//
//  Reference algorithm anchors (KerrTraceCpp2 tree):
//    BL metric / Hamiltonian ........ knds_metric.hpp:45-132
//    RHS (dH/dp, -dH/dq) ............ geodesic.hpp:143-166 (finite differences there;
//                                     analytic here, same Hamiltonian)
//    RK4 step doubling .............. geodesic.hpp:169-257
//    DOPRI5 + FSAL .................. geodesic.hpp:261-414
//    Hermite event location ......... geodesic.hpp:49-140
//    disk redshift .................. main.cpp:329-341
//    camera tetrad .................. camera.hpp:38-145
// ============================================================================
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define KT_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define KT_HD inline
namespace kt {
using std::isfinite; using std::fma; using std::sqrt; using std::fabs; using std::fmin; using std::fmax;
using std::floor; using std::frexp; using std::ldexp; using std::cbrt; using std::exp; using std::log;
using std::log10; using std::pow; using std::fmod; using std::fmaf; using std::atan2; using std::rint; using std::ceil;
KT_HD float sqrtf(float x) { return std::sqrt(x); }
}
#endif

#include <stdint.h>

// Scheduling fence: the machine scheduler may not move instructions across it.  Placed
// between phases of a long straight-line computation whose independent parts it would
// otherwise interleave, multiplying the registers live at once (a host build ignores it).
#if defined(__HIP_DEVICE_COMPILE__)
#define KT_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define KT_SCHED_FENCE() ((void)0)
#endif

namespace kt {

static constexpr double kPi = 3.14159265358979323846;
// cos(M_PI/2) and sin(M_PI/2) as evaluated by the reference (glibc, binary64)
static constexpr double kCosHalfPi = 6.123233995736766e-17;

// ---------------------------------------------------------------------------
//  small math layer
// ---------------------------------------------------------------------------
KT_HD double kt_sqrt(double x) { return sqrt(x); }
KT_HD float kt_sqrt(float x) { return sqrtf(x); }
KT_HD double kt_abs(double x) { return fabs(x); }
KT_HD float kt_abs(float x) { return fabsf(x); }
// sin and cos in binary64, short enough to inline in every stage.  x = k pi/2 + r
// with pi/2 in three parts (161 bits; under fma the reduction error stays below
// 2^-100 for |x| < 2^40, which covers every angle of a ray), then Taylor series on
// |r| <= pi/4 through r^17 (sin) and r^18 (cos): the first omitted terms are below
// 1.3e-19 relative.  NaN and infinity give NaN.  Error ~1 ulp.
KT_HD void kt_sincos(double x, double* s, double* c) {
    const double k = rint(x * 0.6366197723675814);                    // x * 2/pi
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    r = fma(-k, -1.4973849048591698e-33, r);
    if (k == 0.0) r = x;                                               // keeps -0 and tiny x exact
    const double r2 = r*r;
    double ps = 1.0 / 355687428096000.0;                               // 1/17!
    ps = fma(ps, r2, -1.0 / 1307674368000.0);
    ps = fma(ps, r2, 1.0 / 6227020800.0);
    ps = fma(ps, r2, -1.0 / 39916800.0);
    ps = fma(ps, r2, 1.0 / 362880.0);
    ps = fma(ps, r2, -1.0 / 5040.0);
    ps = fma(ps, r2, 1.0 / 120.0);
    ps = fma(ps, r2, -1.0 / 6.0);
    const double sr = (r == 0.0) ? r : fma(r*r2, ps, r);              // sin(-0) = -0
    double pc = -1.0 / 6402373705728000.0;                             // -1/18!
    pc = fma(pc, r2, 1.0 / 20922789888000.0);
    pc = fma(pc, r2, -1.0 / 87178291200.0);
    pc = fma(pc, r2, 1.0 / 479001600.0);
    pc = fma(pc, r2, -1.0 / 3628800.0);
    pc = fma(pc, r2, 1.0 / 40320.0);
    pc = fma(pc, r2, -1.0 / 720.0);
    pc = fma(pc, r2, 1.0 / 24.0);                                     // cos r = 1 - r^2/2 + r^4 pc
    const double hr = 0.5*r2, w = 1.0 - hr;                            // 1 - r^2/2, compensated
    const double cr = w + (((1.0 - w) - hr) + r2*r2*pc);
    const int q = (int)((long long)k & 3);
    const double ss = (q & 1) ? cr : sr, cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}
KT_HD void kt_sincos(float x, float* s, float* c) {
    double sd, cd;
    kt_sincos((double)x, &sd, &cd);
    *s = (float)sd;
    *c = (float)cd;
}
KT_HD bool kt_finite(double x) { return isfinite(x); }
KT_HD bool kt_finite(float x) { return isfinite(x); }

// reciprocal: hardware estimate refined by Newton steps (full precision for
// the normal-range operands that occur here; no denormal/inf special cases)
KT_HD double kt_rcp(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double x = __builtin_amdgcn_rcp(d);
    double e = fma(-d, x, 1.0);
    x = fma(x, e, x);
    e = fma(-d, x, 1.0);
    x = fma(x, e, x);
    return x;
#else
    return 1.0 / d;
#endif
}
KT_HD float kt_rcp(float d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(d);
#else
    return 1.0f / d;
#endif
}

// (sin, cos) of theta0 + delta from (sin theta0, cos theta0) and a short
// Taylor series in delta: sin d = d + d^3 P(d2), 1 - cos d = d2 Q(d2).
//   |delta| <= 1/32 (most steps): P to d^7, Q to d^8, the first omitted terms below 2.5e-18
//                  of sin d and 2.5e-22 absolute in cos d;
//   |delta| <= 1/8: P to d^11, Q to d^12, the first omitted terms below 3e-21 relative;
//   |delta| <= 0.5: P to d^15, Q to d^16, truncation below 3e-20.
// The choice is per lane (a lane's result never depends on its wave), and a wave
// whose lanes all take the short series skips the long one.
template <class T> struct SinCos { T s, c; };

KT_HD SinCos<double> rotate(const SinCos<double>& b, double d) {
    const double d2 = d*d;
    double ps, pc;
    if (fabs(d) <= 0.03125) {
        ps = fma(d2, -1.0/5040.0, 1.0/120.0);
        ps = fma(d2, ps, -1.0/6.0);
        pc = fma(d2, -1.0/40320.0, 1.0/720.0);
        pc = fma(d2, pc, -1.0/24.0);
        pc = fma(d2, pc, 0.5);
    } else if (fabs(d) <= 0.125) {
        ps = fma(d2, -1.0/39916800.0, 1.0/362880.0);
        ps = fma(d2, ps, -1.0/5040.0);
        ps = fma(d2, ps, 1.0/120.0);
        ps = fma(d2, ps, -1.0/6.0);
        pc = fma(d2, -1.0/479001600.0, 1.0/3628800.0);
        pc = fma(d2, pc, -1.0/40320.0);
        pc = fma(d2, pc, 1.0/720.0);
        pc = fma(d2, pc, -1.0/24.0);
        pc = fma(d2, pc, 0.5);
    } else {
        ps = fma(d2, -1.0/1307674368000.0, 1.0/6227020800.0);
        ps = fma(d2, ps, -1.0/39916800.0);
        ps = fma(d2, ps, 1.0/362880.0);
        ps = fma(d2, ps, -1.0/5040.0);
        ps = fma(d2, ps, 1.0/120.0);
        ps = fma(d2, ps, -1.0/6.0);
        pc = fma(d2, -1.0/20922789888000.0, 1.0/87178291200.0);
        pc = fma(d2, pc, -1.0/479001600.0);
        pc = fma(d2, pc, 1.0/3628800.0);
        pc = fma(d2, pc, -1.0/40320.0);
        pc = fma(d2, pc, 1.0/720.0);
        pc = fma(d2, pc, -1.0/24.0);
        pc = fma(d2, pc, 0.5);
    }
    const double sd = fma(d*d2, ps, d);
    const double omc = d2*pc;                          // 1 - cos d, no cancellation
    SinCos<double> o;
    o.s = fma(b.c, sd, fma(-b.s, omc, b.s));           // s cos d + c sin d
    o.c = fma(-b.s, sd, fma(-b.c, omc, b.c));          // c cos d - s sin d
    return o;
}
KT_HD SinCos<float> rotate(const SinCos<float>& b, float d) {
    const float d2 = d*d;
    float ps = fmaf(d2, -1.0f/39916800.0f, 1.0f/362880.0f);
    ps = fmaf(d2, ps, -1.0f/5040.0f);
    ps = fmaf(d2, ps, 1.0f/120.0f);
    ps = fmaf(d2, ps, -1.0f/6.0f);
    const float sd = fmaf(d*d2, ps, d);
    float pc = fmaf(d2, 1.0f/479001600.0f, -1.0f/3628800.0f);
    pc = fmaf(d2, pc, 1.0f/40320.0f);
    pc = fmaf(d2, pc, -1.0f/720.0f);
    pc = fmaf(d2, pc, 1.0f/24.0f);
    pc = fmaf(d2, pc, -0.5f);
    const float omc = -d2*pc;
    SinCos<float> o;
    o.s = fmaf(b.c, sd, fmaf(-b.s, omc, b.s));
    o.c = fmaf(-b.s, sd, fmaf(-b.c, omc, b.c));
    return o;
}
static constexpr double kRotateMax = 0.5;

// ---------------------------------------------------------------------------
//  per-frame physics constants, in the arithmetic type of the RHS
// ---------------------------------------------------------------------------
template <class T> struct Knds {
    T a, a2, M, M2, Q2, lam3, lam3a2, xi, xi2;   // M2 = 2M, lam3 = Lambda/3, lam3a2 = Lambda a^2/3
};

template <class T, class S> KT_HD Knds<T> knds_cast(const Knds<S>& k) {
    Knds<T> o;
    o.a = (T)k.a; o.a2 = (T)k.a2; o.M = (T)k.M; o.M2 = (T)k.M2; o.Q2 = (T)k.Q2;
    o.lam3 = (T)k.lam3; o.lam3a2 = (T)k.lam3a2; o.xi = (T)k.xi; o.xi2 = (T)k.xi2;
    return o;
}

// Right-hand side of Hamilton's equations for H = 1/2 g^{mu nu} p_mu p_nu.
//   out[0] = dr/dl, out[1] = dtheta/dl, out[2] = dp_r/dl, out[3] = dp_theta/dl, out[4] = dphi/dl
//
// BL chart (knds_metric.hpp:99-132), written in the separable form
//   2 Sigma H = Dr pr^2 + Dt pth^2 - Xi^2 K^2 / Dr + Xi^2 L^2 / (Dt s^2)
//   K = (r^2+a^2) pt + a pphi,   L = a s^2 pt + pphi
// which expands to exactly the reference's g^{tt}, g^{t phi}, g^{phi phi}.
//
// KS chart: ingoing Kerr(-Schild) coordinates (v, r, theta, phi~) with
//   dv = dt + Xi (r^2+a^2)/Dr dr, dphi~ = dphi + Xi a / Dr dr, for which
//   2 Sigma H = Dr pr^2 + 2 Xi pr K + Dt pth^2 + Xi^2 L^2 / (Dt s^2),
// regular through the horizon for every Lambda; for Lambda = 0 this is the
// reference's Cartesian Kerr-Schild chart (knds_metric.hpp:135-230) with
// X + iY = (r + i a) sin(theta) e^{i phi~}, Z = r cos(theta), T = v - r.
//
// Arithmetic: one reciprocal of Sigma * Dr * (Dt s^2) (BL) / Sigma * Dt s^2 (KS).  The
// derivatives are formed at half scale, which takes the factors 2 of the quadratic form out
// of every product:  with N = 2 Sigma H and N_s = N / Sigma,
//   dr/dl = (1/2 dN/dp_r) / Sigma,            dtheta/dl = Dt pth / Sigma,
//   -dp_r/dl = (1/2 dN/dr - r N_s) / Sigma,   -dp_theta/dl = s c (X + a^2 N_s) / Sigma,
// where 1/2 dN/dtheta = s c X (every theta-derivative of the metric functions carries s c):
//   X = -(Lambda a^2/3) pth^2 + 2 a p_t B1 - B1 L/(Dt s^2) (Dt - (Lambda a^2/3) s^2),
//   B1 = Xi^2 L / (Dt s^2),
// and 1/2 dDelta_r/dr = r (1 - Lambda r^2/3) - (Lambda/3) r (r^2+a^2) - M.
enum { CHART_BL = 0, CHART_KS = 1 };

// T: type of the state variables (double, float, or the dual numbers of the
// bundle kernel); S: type of the spacetime constants and the conserved p_t, p_phi.
// SCALED: every output multiplied by hscale (the integrator's step size), folded into the
// final scaling by 1/Sigma, so that stage derivatives come out as h k at the cost of one
// multiply instead of one per component.
template <int CHART, class T, class S = T, bool SCALED = false>
KT_HD void rhs_impl(const Knds<S>& P, T r, const SinCos<T>& sc, T pr, T pth, S pt, S pp, T hscale, T out[5]) {
    const T s = sc.s, c = sc.c;
    const T s2 = s*s, c2 = c*c, sc2 = s*c;
    const T r2 = r*r;
    const T ra = r2 + P.a2;
    const T sig = fma(P.a2, c2, r2);
    const T one_m = fma(-P.lam3, r2, S(1));                  // 1 - Lambda r^2/3
    const T dR = fma(ra, one_m, fma(-P.M2, r, P.Q2));        // Delta_r
    const T hdR_r = fma(r, one_m, fma(-(P.lam3*r), ra, -P.M));   // 1/2 dDelta_r/dr
    const T dT = fma(P.lam3a2, c2, S(1));                    // Delta_theta
    const T K = fma(ra, pt, P.a*pp);
    const T L = fma(P.a*s2, pt, pp);
    const T dTs2 = dT*s2;
    const T dTpth = dT*pth;
    T inv_sig, inv_ts, B1, N, o0, hdNr;
    if (CHART == CHART_BL) {
        const T p12 = sig*dR;
        const T inv = kt_rcp(p12*dTs2);
        inv_sig = inv*dR*dTs2;
        const T inv_dR = inv*sig*dTs2;
        inv_ts = inv*p12;
        const T A1 = P.xi2*K*inv_dR;                          // Xi^2 K / Dr
        B1 = P.xi2*L*inv_ts;                                  // Xi^2 L / (Dt s^2)
        N = fma(dR*pr, pr, fma(dTpth, pth, fma(-A1, K, B1*L)));
        o0 = dR*pr;
        out[4] = fma(-A1, P.a, B1);
        // 1/2 dN/dr = 1/2 dDelta_r/dr (pr^2 + Xi^2 K^2/Dr^2) - 2 A1 p_t r
        hdNr = fma(hdR_r, fma(pr, pr, A1*K*inv_dR), -(((pt + pt)*r)*A1));
    } else {
        const T inv = kt_rcp(sig*dTs2);
        inv_sig = inv*dTs2;
        inv_ts = inv*sig;
        B1 = P.xi2*L*inv_ts;                                  // Xi^2 L / (Dt s^2)
        const T XK = P.xi*K;
        N = fma(dR*pr, pr, fma(XK + XK, pr, fma(dTpth, pth, B1*L)));
        o0 = fma(dR, pr, XK);
        out[4] = fma(P.xi*P.a, pr, B1);
        // 1/2 dN/dr = 1/2 dDelta_r/dr pr^2 + 2 Xi p_t r pr
        hdNr = pr*fma(hdR_r, pr, ((P.xi + P.xi)*pt)*r);
    }
    const T inv_out = SCALED ? inv_sig*hscale : inv_sig;
    const T Ns = N*inv_sig;
    const T X = fma(-(P.lam3a2*pth), pth,
                    fma((P.a + P.a)*pt, B1, -(B1*L*inv_ts)*fma(-P.lam3a2, s2, dT)));
    out[0] = o0*inv_out;
    out[1] = dTpth*inv_out;
    out[4] = out[4]*inv_out;
    out[2] = -fma(-r, Ns, hdNr)*inv_out;
    out[3] = -(fma(P.a2, Ns, X)*sc2)*inv_out;
}

template <int CHART, class T, class S = T>
KT_HD void rhs(const Knds<S>& P, T r, const SinCos<T>& sc, T pr, T pth, S pt, S pp, T out[5]) {
    rhs_impl<CHART, T, S, false>(P, r, sc, pr, pth, pt, pp, T(), out);   // scale unused
}

// the right-hand side times h: (h dr, h dtheta, h dp_r, h dp_theta, h dphi)
template <int CHART, class T, class S = T>
KT_HD void rhs_h(const Knds<S>& P, T r, const SinCos<T>& sc, T pr, T pth, S pt, S pp, T h, T out[5]) {
    rhs_impl<CHART, T, S, true>(P, r, sc, pr, pth, pt, pp, h, out);
}

// 2 Sigma H (the numerator N above), for constraint monitoring / renormalisation
template <int CHART, class T>
KT_HD T hamiltonian_num(const Knds<T>& P, T r, const SinCos<T>& sc, T pr, T pth, T pt, T pp, T* sig_out) {
    const T s2 = sc.s*sc.s, c2 = sc.c*sc.c;
    const T r2 = r*r, ra = r2 + P.a2;
    const T sig = fma(P.a2, c2, r2);
    const T dR = fma(ra, fma(-P.lam3, r2, (T)1), fma(-P.M2, r, P.Q2));
    const T dT = fma(P.lam3a2, c2, (T)1);
    const T K = fma(ra, pt, P.a*pp);
    const T L = fma(P.a*s2, pt, pp);
    if (sig_out) *sig_out = sig;
    if (CHART == CHART_BL)
        return dR*pr*pr + dT*pth*pth - P.xi2*K*K/dR + P.xi2*L*L/(dT*s2);
    return dR*pr*pr + (T)2*P.xi*pr*K + dT*pth*pth + P.xi2*L*L/(dT*s2);
}

}  // namespace kt
