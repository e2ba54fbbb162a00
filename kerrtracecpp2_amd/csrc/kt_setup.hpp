// ============================================================================
//  kt_setup.hpp -- per-frame scalars: horizon, ISCO, camera constants.
//
//  Reference anchors (KerrTraceCpp2 tree):
//    r_horizon ........ knds_metric.hpp:237-315 (closed form for Lambda = 0; 20000-sample
//                       log scan of Delta_r + 160-step bisection otherwise)
//    r_isco ........... knds_metric.hpp:318-349 (Bardeen closed form for Q = Lambda = 0;
//                       scan r += 1e-4 M from 1.05 r_+ for the first decrease of L_circ)
//    camera constants . camera.hpp:66-142 (metric at the observer, static tetrad)
//    frame scalars .... main.cpp:2078-2082 (r_disk_in = r_isco, r_escape = 1.05 r_obs)
//
//  The ISCO scan and the Lambda = 0 horizon reproduce the reference bit for bit:
//  the scan abscissae are the exact binary64 partial sums of r += dr (computed in
//  closed form per binade) and L_circ is evaluated with the reference's operation
//  order and no FMA contraction.
// ============================================================================
#pragma once
#include "kt_trace.hpp"

#include <cmath>
#include <limits>

namespace kt {

// Delta_r with the reference's operation order (knds_metric.hpp:51-53)
KT_HD double delta_r_ref(const FrameDev& F, double r) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    return (r*r + F.a*F.a) * (1.0 - F.Lambda*r*r/3.0) - 2.0*F.M*r + F.Q*F.Q;
}

// -L_circ of the prograde-by-convention Keplerian orbit (knds_metric.hpp:333-341)
KT_HD double isco_L(const FrameDev& F, double rr) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double s = (F.a < 0.0) ? 1.0 : -1.0;
    const double Me = F.M - F.Q*F.Q/(2.0*rr) + F.Lambda*F.a*rr*rr/3.0;
    const double sq = sqrt(Me > 0.0 ? Me : 0.0);
    const double den = rr*sqrt(rr) + s*F.a*sq;
    const double Om = (fabs(den) < 1e-14) ? 0.0 : s * sq / den;
    const double a = F.a, ct = kCosHalfPi;
    const double sg = rr*rr + a*a*ct*ct;
    const double dr = (rr*rr + a*a) * (1.0 - F.Lambda*rr*rr/3.0) - 2.0*F.M*rr + F.Q*F.Q;
    const double dt = 1.0 + F.Lambda*a*a*ct*ct/3.0;
    const double xi = 1.0 + F.Lambda*a*a/3.0;
    const double x2 = xi*xi, s2 = 1.0*1.0, ra = rr*rr + a*a, den2 = sg*x2;
    const double g00 = (-dr + dt*a*a*s2) / den2;
    const double g03 = a*s2*(dr - dt*ra) / den2;
    const double g33 = s2*(dt*ra*ra - dr*a*a*s2) / den2;
    const double N2 = -(g00 + 2.0*g03*Om + g33*Om*Om);
    if (N2 <= 0.0) return -1e30;
    const double ut = 1.0/sqrt(N2);
    return -(g03 + g33*Om)*ut;
}

// Bardeen-Press-Teukolsky ISCO (knds_metric.hpp:320-327)
KT_HD double isco_bardeen(double M, double a) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    if (fabs(a) < 1e-10) return 6.0*M;
    const double ap = a/M;
    const double z1 = 1.0 + cbrt(1.0-ap*ap) * (cbrt(1.0+ap)+cbrt(1.0-ap));
    const double z2 = sqrt(3.0*ap*ap + z1*z1);
    return M*(3.0 + z2 - sqrt((3.0-z1)*(3.0+z1+2.0*z2)));
}

// Exact representation of the partial sums r_k of "r += dr": within one binade
// [2^e, 2^(e+1)) the increment is the constant multiple d*u of the ulp u.
struct ScanSeg {
    int64_t k0, n;        // first index, count
    int64_t m0, d;        // r_k = (m0 + (k-k0) d) * 2^ex
    int32_t ex, pad;
};
static constexpr int kMaxSeg = 24;

// Builds the segments of r_0 = r0, r_{k+1} = fl(r_k + dr) up to r >= rmax.
// Returns the segment count, or -1 if the constant-increment property fails
// (ties), in which case callers fall back to the sequential loop.
KT_HD int isco_segments(double r0, double dr, double rmax, ScanSeg* seg) {
    double r = r0;
    int64_t k = 0;
    int ns = 0;
    while (r < rmax && ns < kMaxSeg) {
        int e;
        frexp(r, &e);                      // r = f 2^e, f in [0.5, 1)
        const int ex = e - 53;             // ulp of the binade [2^(e-1), 2^e)
        const double u = ldexp(1.0, ex);
        const double mr = r / u;           // exact integer < 2^53
        const double q = dr / u;
        const double qf = floor(q);
        if (q - qf == 0.5) return -1;
        const double dq = (q - qf > 0.5) ? qf + 1.0 : qf;
        if (dq < 1.0) return -1;
        const int64_t m0 = (int64_t)mr, d = (int64_t)dq;
        // verify the rounding model on the first increment inside the binade
        const int64_t lim = (int64_t)1 << 53;
        if (m0 + d < lim && r + dr != (double)(m0 + d) * u) return -1;
        const int64_t n = (lim - m0 + d - 1) / d;       // indices with m0 + j d < 2^53
        seg[ns].k0 = k; seg[ns].n = n; seg[ns].m0 = m0; seg[ns].d = d; seg[ns].ex = ex; seg[ns].pad = 0;
        ++ns;
        const double rlast = (double)(m0 + (n - 1)*d) * u;
        r = rlast + dr;                   // binade crossing: one real binary64 addition
        k += n;
    }
    return ns;
}

KT_HD double isco_abscissa(const ScanSeg* seg, int ns, int64_t k) {
    for (int i = 0; i < ns; ++i)
        if (k >= seg[i].k0 && k < seg[i].k0 + seg[i].n)
            return ldexp((double)(seg[i].m0 + (k - seg[i].k0)*seg[i].d), seg[i].ex);
    return 1e300;
}

// Camera constants at (r_obs, theta_obs): BL metric (knds_metric.hpp:71-117) and
// the static-observer tetrad of camera.hpp:84-114.
KT_HD void frame_camera_consts(FrameDev& F) {
    const double r = F.r_obs, th = F.th_obs, a = F.a;
    double st, ct;
    kt_sincos(th, &st, &ct);
    const double sg = r*r + a*a*ct*ct;
    const double dr = (r*r + a*a) * (1.0 - F.Lambda*r*r/3.0) - 2.0*F.M*r + F.Q*F.Q;
    const double dt = 1.0 + F.Lambda*a*a*ct*ct/3.0;
    const double xi = 1.0 + F.Lambda*a*a/3.0;
    const double x2 = xi*xi, s2 = st*st, ra = r*r + a*a;
    const double den = sg*x2;
    F.gtt = (-dr + dt*a*a*s2) / den;
    F.gtp = a*s2*(dr - dt*ra) / den;
    F.grr = sg/dr;
    F.gth = sg/dt;
    F.gpp = s2*(dt*ra*ra - dr*a*a*s2) / den;
    const double pre = sg*dr*dt;
    F.U00 = -x2*(dt*ra*ra - dr*a*a*s2) / pre;
    F.U03 = a*x2*(dr - dt*ra) / pre;
    F.U11 = dr/sg;
    F.U22 = dt/sg;
    F.U33 = (s2 > 1e-14) ? x2*(dr - dt*a*a*s2) / (pre*s2) : 0.0;
    F.sq_rr = sqrt(fabs(F.grr));
    F.sq_th = sqrt(fabs(F.gth));
    F.dphi = F.gpp - (F.gtp*F.gtp)/F.gtt;
    F.frame_ok = (F.gtt < -1e-14) && isfinite(F.dphi) && (F.dphi > 1e-14) && isfinite(F.sq_rr) &&
                 (F.sq_rr > 1e-14) && isfinite(F.sq_th) && (F.sq_th > 1e-14);
    F.u0 = F.frame_ok ? 1.0 / sqrt(-F.gtt) : 0.0;
    F.ep = F.frame_ok ? 1.0 / sqrt(F.dphi) : 0.0;
    F.et = F.frame_ok ? -F.gtp/F.gtt * F.ep : 0.0;
}

// Constants that depend only on (M, a, Q, Lambda) in the RHS's form.
KT_HD void frame_knds(FrameDev& F) {
    Knds<double>& P = F.P;
    P.a = F.a; P.a2 = F.a*F.a; P.M = F.M; P.M2 = 2.0*F.M; P.Q2 = F.Q*F.Q;
    P.lam3 = F.Lambda/3.0; P.lam3a2 = F.Lambda*F.a*F.a/3.0;
    P.xi = 1.0 + F.Lambda*F.a*F.a/3.0; P.xi2 = P.xi*P.xi;
}

// Sequential horizon / ISCO (used by the host check build; the device uses the
// parallel kernels of kt_hip.hip, which must agree with these).
inline double horizon_seq(const FrameDev& F) {
    if (std::fabs(F.Lambda) < 1e-15) {
        const double disc = F.M*F.M - F.a*F.a - F.Q*F.Q;
        return F.M + std::sqrt(disc > 0.0 ? disc : 0.0);
    }
    const double scale = std::fabs(F.M) > 1.0 ? std::fabs(F.M) : 1.0;
    const double rmin = (1e-6*scale > 1e-9) ? 1e-6*scale : 1e-9;
    double rmax;
    if (F.Lambda > 0.0) { const double t = 1.25*std::sqrt(3.0/F.Lambda); rmax = 20.0*scale > t ? 20.0*scale : t; }
    else rmax = 200.0*scale + 20.0*std::fabs(F.a) + 20.0*std::fabs(F.Q);
    if (!(rmax > rmin)) rmax = rmin * 10.0;
    double roots[16];
    int nr = 0;
    auto add = [&](double rr) {
        if (!std::isfinite(rr) || rr <= 0.0) return;
        for (int i = 0; i < nr; ++i) {
            double rel = 1.0;
            if (std::fabs(roots[i]) > rel) rel = std::fabs(roots[i]);
            if (std::fabs(rr) > rel) rel = std::fabs(rr);
            if (std::fabs(roots[i] - rr) <= 1e-8 * rel) return;
        }
        if (nr < 16) roots[nr++] = rr;
    };
    auto bis = [&](double lo, double hi) {
        double flo = delta_r_ref(F, lo), fhi = delta_r_ref(F, hi);
        if (!std::isfinite(flo) || !std::isfinite(fhi)) return std::numeric_limits<double>::quiet_NaN();
        for (int i = 0; i < 160; ++i) {
            const double mid = 0.5*(lo + hi);
            // fixed point while flo, fhi have strictly opposite signs: the remaining steps
            // change nothing (with flo an exact zero they still move lo)
            if ((mid == lo || mid == hi) && ((flo > 0.0 && fhi < 0.0) || (flo < 0.0 && fhi > 0.0))) break;
            const double fm = delta_r_ref(F, mid);
            if (!std::isfinite(fm)) break;
            if ((flo > 0.0 && fm < 0.0) || (flo < 0.0 && fm > 0.0)) { hi = mid; fhi = fm; }
            else { lo = mid; flo = fm; }
        }
        return 0.5*(lo + hi);
    };
    const int N = 20000;
    const double span = std::log(rmax / rmin);
    double r0 = rmin, f0 = delta_r_ref(F, r0);
    for (int i = 1; i <= N; ++i) {
        const double t = double(i) / double(N);
        const double r1 = rmin * std::exp(span * t);
        const double f1 = delta_r_ref(F, r1);
        if (std::isfinite(f0) && std::isfinite(f1)) {
            if (f0 == 0.0) add(r0);
            if (f1 == 0.0) add(r1);
            if ((f0 > 0.0 && f1 < 0.0) || (f0 < 0.0 && f1 > 0.0)) add(bis(r0, r1));
        }
        r0 = r1; f0 = f1;
    }
    if (nr == 0) {
        const double disc = F.M*F.M - F.a*F.a - F.Q*F.Q;
        return F.M + std::sqrt(disc > 0.0 ? disc : 0.0);
    }
    for (int i = 1; i < nr; ++i)
        for (int j = i; j > 0 && roots[j-1] > roots[j]; --j) { const double t = roots[j]; roots[j] = roots[j-1]; roots[j-1] = t; }
    if (F.Lambda > 0.0 && nr >= 2) return roots[nr - 2];
    return roots[nr - 1];
}

inline double isco_seq(const FrameDev& F, double rh) {
    if (std::fabs(F.Lambda) < 1e-15 && std::fabs(F.Q) < 1e-15) return isco_bardeen(F.M, F.a);
    double r = rh * 1.05;
    const double dr = 1e-4*F.M;
    double Lp = isco_L(F, r);
    for (; r < 30.0*F.M; r += dr) {
        const double L = isco_L(F, r + dr);
        if (L < Lp && Lp > 0.0) return r;
        Lp = L;
    }
    return r;
}

}  // namespace kt
