// ============================================================================
//  kt_trace.hpp -- per-ray tracer state machine (camera -> adaptive steps ->
//  disk / horizon / escape events -> GeoPixel), shared by the persistent HIP
//  kernel and the host-compiled check build.
//
//  Reference anchors (KerrTraceCpp2 tree):
//    camera ............ camera.hpp:38-145
//    BL tracer loop .... main.cpp:343-434 (reject cap 64, max_steps, event arbitration)
//    KS tracer loop .... main.cpp:1530-1660 (initial covector :1404-1430)
//    RK4 doubling ...... geodesic.hpp:204-257
//    DOPRI5 ............ geodesic.hpp:312-414
//    events ............ geodesic.hpp:49-140, main.cpp:363-431
//    disk redshift ..... main.cpp:329-341
// ============================================================================
#pragma once
#include "kt_device.hpp"

namespace kt {

enum { METHOD_RK4 = 0, METHOD_DP5 = 1 };
enum { PREC_FP64 = 0, PREC_MIXED = 1 };

// quirk bits (same values as the KT_QUIRK_* macros of include/kt_hip.h)
enum : uint32_t {
    Q_DP5_PHI_FROZEN = 1u << 0,
    Q_ESC_POSITIONAL = 1u << 1,
    Q_KS_INIT_NO_SHIFT = 1u << 2,
    Q_BUNDLE_HORIZON_ESC = 1u << 3,
    Q_BUNDLE_ROOT_SKIP = 1u << 4,
    Q_KS_LAMBDA_TO_BL = 1u << 5,
    // internal mode of the differentiable path: converge the disk root (below)
    Q_EXACT_ROOT = 1u << 30,
};

// Per-frame constants, produced by the setup kernel (device memory, 64-bit aligned).
struct FrameDev {
    // inputs
    double M, a, Q, Lambda;
    double r_obs, th_obs, ph_obs, fov;     // radians
    double disk_out, off_x, off_y;
    int32_t W, H, tiles_x, tiles_y;
    int64_t ray_base;                      // first global ray index (8x8 tile padded)
    int64_t pix_base;                      // first output pixel
    const uint32_t* tile_order;            // queue rank -> 8x8 tile (row-major id); null: row-major
    // derived
    double r_h, r_isco, r_in, r_out, r_esc, r_cut;
    Knds<double> P;
    int32_t chart, ks_ref_init;            // effective chart; KS covector per reference (Lambda == 0)
    // camera constants at the observer (camera.hpp:66-142, pixel independent)
    double gtt, gtp, grr, gth, gpp, sq_rr, sq_th, dphi;
    double u0, ep, et;
    int32_t frame_ok, pad0;
    double U00, U03, U11, U22, U33;
    double isco_closed;    // host-evaluated Bardeen ISCO when Q = Lambda = 0 (host libm, as the reference)
};

// default island parameters of the mixed-precision path
static constexpr double kMixedNoiseDefault = 1e-9;
static constexpr double kMixedRIslandDefault = 2.5;
static constexpr int kMixedHoldDefault = 0;   // retry off unless asked for

struct RunOpts {
    int32_t max_steps;
    uint32_t quirks;
    double step_init, rtol, atol;
    double isl_noise;      // mixed: FP32 stage noise coefficient for the island test
    double isl_r;          // mixed: binary64 below r < isl_r * r_horizon
    int32_t isl_hold = 0;  // mixed: accepted binary64 steps owed after a binary32 rejection
                           // (which is retried at the same step size); 0: no retry
    int32_t pad = 0;
};

// result fields of one ray before GeoPixel narrowing (TraceResult, main.cpp:187-191)
struct RayOut {
    int32_t fate;          // 0 escaped, 1 disk, 2 horizon
    double r, g, ph_disk, th_esc, ph_esc;
};

// ---------------------------------------------------------------------------
//  camera: pixel -> BL state with p_t re-solved from H = 0   camera.hpp:38-145
// ---------------------------------------------------------------------------
KT_HD void pixel_angles(const FrameDev& F, double px, double py, double& al, double& be) {
    const int span = (F.W > 1) ? (F.W - 1) : 1;
    al = F.fov * (px + F.off_x - 0.5*(F.W-1)) / span;
    be = F.fov * (0.5*(F.H-1) - (py + F.off_y)) / span;
}

// screen angles (alpha, beta) -> BL state (camera.hpp:50-145)
KT_HD void camera_shoot(const FrameDev& F, double al, double be, double st[7]) {
    double sa, ca, sb, cb;
    kt_sincos(al, &sa, &ca);
    kt_sincos(be, &sb, &cb);
    double ut, ur, uth, uph;
    if (F.frame_ok) {
        const double nr = -ca*cb, nt = -sb, np = -sa*cb;
        ut = F.u0 + np*F.et;
        ur = nr / F.sq_rr;
        uth = nt / F.sq_th;
        uph = np*F.ep;
    } else {
        ut = 1.0;
        ur = -ca*cb / fmax(F.sq_rr, 1e-14);
        uth = -sb / fmax(F.sq_th, 1e-14);
        uph = -sa*cb / sqrt(fmax(fabs(F.gpp), 1e-14));
    }
    double pt = F.gtt*ut + F.gtp*uph;
    const double pr = F.grr*ur;
    const double pth = F.gth*uth;
    const double pp = F.gtp*ut + F.gpp*uph;
    const double A = F.U00, B = 2.0*F.U03*pp;
    const double C = F.U11*pr*pr + F.U22*pth*pth + F.U33*pp*pp;
    const double disc = B*B - 4.0*A*C;
    if (disc >= 0.0 && fabs(A) > 1e-15) {
        const double sq = sqrt(disc);
        const double p1 = (-B - sq) / (2.0*A), p2 = (-B + sq) / (2.0*A);
        pt = (p1 < 0.0) ? p1 : p2;
        if (pt > 0.0) pt = fmin(p1, p2);
    }
    st[0] = F.r_obs; st[1] = F.th_obs; st[2] = F.ph_obs;
    st[3] = pr; st[4] = pth; st[5] = pt; st[6] = pp;
}

KT_HD void camera_ray(const FrameDev& F, double px, double py, double st[7]) {
    double al, be;
    pixel_angles(F, px, py, al, be);
    camera_shoot(F, al, be, st);
}

// BL state -> ingoing Kerr-Schild covector (v, r, theta, phi~), same event point.
//  reference form (Lambda == 0): spatial covector kept, p_T re-solved from H = 0
//  (main.cpp:1404-1430); p_r(v chart) = p_r(T chart) - p_T.
//  exact form: p_v = p_t, p_r(v) = p_r - Xi K / Delta_r.
KT_HD void bl_to_ks_covector(const FrameDev& F, double st[7]) {
    const Knds<double>& P = F.P;
    const double r = st[0], th = st[1];
    double s, c;
    kt_sincos(th, &s, &c);
    const double s2 = s*s, c2 = c*c, ra = r*r + P.a2;
    const double dR = ra*(1.0 - P.lam3*r*r) - P.M2*r + P.Q2;
    const double dT = 1.0 + P.lam3a2*c2;
    const double pr = st[3], pth = st[4], pt = st[5], pp = st[6];
    if (F.ks_ref_init) {
        const double Pr = pr;
        const double A = dR - 2.0*P.xi*ra + P.xi2*P.a2*s2/dT;
        const double B = -2.0*dR*Pr + 2.0*P.xi*Pr*ra - 2.0*P.xi*P.a*pp + 2.0*P.xi2*P.a*pp/dT;
        const double C = dR*Pr*Pr + 2.0*P.xi*Pr*P.a*pp + dT*pth*pth + P.xi2*pp*pp/(dT*s2);
        const double disc = B*B - 4.0*A*C;
        if (disc >= 0.0 && fabs(A) > 1e-15) {
            const double sq = sqrt(disc);
            const double p1 = (-B - sq) / (2.0*A), p2 = (-B + sq) / (2.0*A);
            double pT = (p1 < 0.0) ? p1 : p2;
            if (pT > 0.0) pT = fmin(p1, p2);
            st[5] = pT;
            st[3] = Pr - pT;
            return;
        }
    }
    const double K = ra*pt + P.a*pp;
    st[3] = pr - P.xi*K/dR;
}

// ---------------------------------------------------------------------------
//  event helpers (geodesic.hpp:49-140)
// ---------------------------------------------------------------------------
KT_HD double hermite(double y0, double y1, double d0, double d1, double h, double al) {
    al = al < 0.0 ? 0.0 : (al > 1.0 ? 1.0 : al);
    const double a2 = al*al, a3 = a2*al;
    const double h00 = 2.0*a3 - 3.0*a2 + 1.0;
    const double h10 = a3 - 2.0*a2 + al;
    const double h01 = -2.0*a3 + 3.0*a2;
    const double h11 = a3 - a2;
    return h00*y0 + h10*(h*d0) + h01*y1 + h11*(h*d1);
}
KT_HD bool crosses(double f0, double f1) { return (f0 <= 0.0 && f1 >= 0.0) || (f0 >= 0.0 && f1 <= 0.0); }

// Can the Hermite cubic minus target come within 1e-12 of zero on [0,1]?  Exact
// bound from the endpoint values and the interior critical points; a cheap
// pre-test before the 8-bin / 8-bisection search of the reference.
KT_HD bool hermite_may_touch(double y0, double y1, double d0, double d1, double h, double tgt) {
    const double f0 = y0 - tgt, f1 = y1 - tgt;
    if (crosses(f0, f1)) return true;
    const double m0 = h*d0, m1 = h*d1;
    // generous margin for the rounding of the monomial form
    const double margin = 1e-12 + 1e-12*(fabs(f0) + fabs(f1) + fabs(m0) + fabs(m1));
    // Bernstein form: the cubic lies in the hull of f0, f0 + m0/3, f1 - m1/3, f1; a hull
    // clear of [-margin, margin] settles most steps near the plane without the extrema
    const double b1 = f0 + m0*(1.0/3.0), b2 = f1 - m1*(1.0/3.0);
    if (fmin(fmin(f0, f1), fmin(b1, b2)) > margin || fmax(fmax(f0, f1), fmax(b1, b2)) < -margin) return false;
    // cubic in alpha: f(al) = c3 al^3 + c2 al^2 + c1 al + c0
    const double c3 = 2.0*f0 + m0 - 2.0*f1 + m1;
    const double c2 = -3.0*f0 - 2.0*m0 + 3.0*f1 - m1;
    const double c1 = m0;
    double lo = fmin(f0, f1), hi = fmax(f0, f1);
    // roots of 3 c3 al^2 + 2 c2 al + c1 (an extremum's position only needs to be close:
    // the cubic is flat there)
    const double qa = 3.0*c3, qb = 2.0*c2, qc = c1;
    double ex[2];
    int n = 0;
    if (fabs(qa) > 1e-300) {
        const double disc = qb*qb - 4.0*qa*qc;
        if (disc >= 0.0) {
            const double sq = sqrt(disc), inv = kt_rcp(2.0*qa);
            ex[n++] = (-qb - sq) * inv;
            ex[n++] = (-qb + sq) * inv;
        }
    } else if (fabs(qb) > 1e-300) {
        ex[n++] = -qc * kt_rcp(qb);
    }
    for (int i = 0; i < n; ++i) {
        const double al = ex[i];
        if (al > 0.0 && al < 1.0) {
            const double v = ((c3*al + c2)*al + c1)*al + f0;
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
    }
    return lo <= margin && hi >= -margin;
}

// Can the Hermite interpolant of r over the step reach [r_in, r_out]?  The cubic lies in
// the hull of its Bernstein control points r0, r0 + m0/3, r1 - m1/3, r1 (m = h dr/dl), so a
// hull clear of the disk's radii (with a margin for rounding) rules out a disk hit at any
// root the search could return; the search is then skipped with the same outcome.
KT_HD bool hermite_may_reach(double r0, double r1, double dr0, double dr1, double h, double r_in, double r_out) {
    const double m0 = h*dr0, m1 = h*dr1;
    const double b1 = r0 + m0*(1.0/3.0), b2 = r1 - m1*(1.0/3.0);
    const double lo = fmin(fmin(r0, r1), fmin(b1, b2)), hi = fmax(fmax(r0, r1), fmax(b1, b2));
    const double margin = 1e-9*(fabs(r0) + fabs(r1) + fabs(m0) + fabs(m1));
    return hi >= r_in - margin && lo <= r_out + margin;
}

// first root of the Hermite cubic on [0,1]: 8 bins, 8 bisections, |f| <= 1e-12 short cut
KT_HD bool hermite_first_root(double y0, double y1, double d0, double d1, double h, double tgt, double& out) {
    const double feps = 1e-12;
    double a0 = 0.0, f0 = y0 - tgt;
    if (fabs(f0) <= feps) { out = 0.0; return true; }
    for (int i = 1; i <= 8; ++i) {
        const double a1 = double(i) / 8.0;
        const double f1 = (i == 8) ? (y1 - tgt) : (hermite(y0, y1, d0, d1, h, a1) - tgt);
        if (fabs(f1) <= feps) { out = a1; return true; }
        if (crosses(f0, f1)) {
            double lo = a0, hi = a1;
            for (int k = 0; k < 8; ++k) {
                const double mid = 0.5*(lo + hi);
                const double fm = hermite(y0, y1, d0, d1, h, mid) - tgt;
                if (crosses(f0, fm)) hi = mid; else lo = mid;
            }
            out = 0.5*(lo + hi);
            return true;
        }
        a0 = a1;
        f0 = f1;
    }
    return false;
}

// Newton steps on the Hermite cubic from the bisection result al, kept inside its final
// bracket (half-width 2^-12): the root to rounding, so that the hit point is a smooth
// function of the ray (the 8 bisections alone leave it piecewise constant).
KT_HD double hermite_root_polish(double y0, double y1, double d0, double d1, double h, double tgt, double al) {
    const double lo = al - 1.0/4096.0 > 0.0 ? al - 1.0/4096.0 : 0.0;
    const double hi = al + 1.0/4096.0 < 1.0 ? al + 1.0/4096.0 : 1.0;
    for (int it = 0; it < 4; ++it) {
        const double a2 = al*al;
        const double f = hermite(y0, y1, d0, d1, h, al) - tgt;
        const double df = (6.0*a2 - 6.0*al)*y0 + (3.0*a2 - 4.0*al + 1.0)*(h*d0) + (-6.0*a2 + 6.0*al)*y1 +
                          (3.0*a2 - 2.0*al)*(h*d1);
        if (!(fabs(df) > 0.0)) break;
        const double an = al - f/df;
        if (!(an >= lo && an <= hi)) break;
        al = an;
    }
    return al;
}

// Keplerian angular velocity, reference sign convention   knds_metric.hpp:354-363
KT_HD double omega_k(const FrameDev& F, double r) {
    const double s = (F.a < 0.0) ? 1.0 : -1.0;
    const double Me = F.M - F.Q*F.Q/(2.0*r) + F.Lambda*F.a*r*r/3.0;
    const double sq = sqrt(fmax(Me, 0.0));
    const double den = r*sqrt(r) + s*F.a*sq;
    if (fabs(den) < 1e-14) return 0.0;
    return s * sq / den;
}

// g_tt, g_tphi, g_phiphi of BL at theta = pi/2 (knds_metric.hpp:71-88 with the
// reference's cos(M_PI/2), sin(M_PI/2))
KT_HD void lower_equator(const FrameDev& F, double r, double& gtt, double& gtp, double& gpp) {
    const double a = F.a, ct = kCosHalfPi;
    const double sg = r*r + a*a*ct*ct;
    const double dr = (r*r + a*a) * (1.0 - F.Lambda*r*r/3.0) - 2.0*F.M*r + F.Q*F.Q;
    const double dt = 1.0 + F.Lambda*a*a*ct*ct/3.0;
    const double xi = 1.0 + F.Lambda*a*a/3.0;
    const double x2 = xi*xi, s2 = 1.0, ra = r*r + a*a, den = sg*x2;
    gtt = (-dr + dt*a*a*s2) / den;
    gtp = a*s2*(dr - dt*ra) / den;
    gpp = s2*(dt*ra*ra - dr*a*a*s2) / den;
}

// redshift of the thin disk at r (main.cpp:329-341), clamped to [0,20] (main.cpp:390)
KT_HD double disk_redshift(const FrameDev& F, double r, double pt, double pp) {
    const double Om = omega_k(F, r);
    const double b = -pp / (-pt);
    double gtt, gtp, gpp;
    lower_equator(F, r, gtt, gtp, gpp);
    const double d2 = -(gtt + 2.0*gtp*Om + gpp*Om*Om);
    double red;
    if (d2 <= 0.0) red = 1.0;
    else {
        const double dv = fabs(1.0 - Om*b);
        if (dv < 1e-8) red = 20.0;
        else {
            red = sqrt(d2) / dv;
            if (!isfinite(red)) red = 1.0;
        }
    }
    return red < 0.0 ? 0.0 : (red > 20.0 ? 20.0 : red);
}

// ---------------------------------------------------------------------------
//  DOPRI5 tableau (Dormand & Prince 1980; geodesic.hpp:261-274)
// ---------------------------------------------------------------------------
struct DP5 {
    static constexpr double a21 = 1.0/5.0;
    static constexpr double a31 = 3.0/40.0, a32 = 9.0/40.0;
    static constexpr double a41 = 44.0/45.0, a42 = -56.0/15.0, a43 = 32.0/9.0;
    static constexpr double a51 = 19372.0/6561.0, a52 = -25360.0/2187.0, a53 = 64448.0/6561.0,
                            a54 = -212.0/729.0;
    static constexpr double a61 = 9017.0/3168.0, a62 = -355.0/33.0, a63 = 46732.0/5247.0,
                            a64 = 49.0/176.0, a65 = -5103.0/18656.0;
    static constexpr double b1 = 35.0/384.0, b3 = 500.0/1113.0, b4 = 125.0/192.0,
                            b5 = -2187.0/6784.0, b6 = 11.0/84.0;
    // b5 - b4 (error weights; b4 of geodesic.hpp:337-340)
    static constexpr double e1 = 35.0/384.0 - 5179.0/57600.0;
    static constexpr double e3 = 500.0/1113.0 - 7571.0/16695.0;
    static constexpr double e4 = 125.0/192.0 - 393.0/640.0;
    static constexpr double e5 = -2187.0/6784.0 + 92097.0/339200.0;
    static constexpr double e6 = 11.0/84.0 - 187.0/2100.0;
    static constexpr double e7 = -1.0/40.0;
};

// ---------------------------------------------------------------------------
//  one ray
// ---------------------------------------------------------------------------
struct RayState {
    double y[4];           // r, theta, p_r, p_theta
    double ph, pt, pp;     // phi (phi~ in KS), conserved p_t (p_v), p_phi
    double h;
    double k1[5];          // RHS at y (FSAL / first stage); [4] = dphi/dl
    SinCos<double> sc;     // sin, cos of y[1]
    double r_cut, r_esc;   // the frame's horizon cut and escape radius, tested every step:
                           // per lane in registers, not re-read from the frame table
    int32_t n_acc, n_rej;
    int32_t rhs32, rhs64, hermites, rej_total;
    int32_t hold;          // mixed: binary64 steps still owed after a binary32 rejection
};

// Wave-level vote policy: the kernel passes one backed by __ballot, the host check
// build evaluates lanes one at a time.
struct LaneVote {
    KT_HD static bool any(bool b) { return b; }
};

// stage derivative at y0 + d (NC = 4: r, theta, p_r, p_theta; NC = 5: + phi)
template <int CHART, int NC>
KT_HD void stage_f64(const Knds<double>& P, const double y0[4], const double d[4], const SinCos<double>& sc0,
                     double pt, double pp, double k[NC], SinCos<double>& sc) {
    // per lane (not a wave vote): a ray's result must not depend on its wave-mates
    if (fabs(d[1]) <= kRotateMax) sc = rotate(sc0, d[1]);
    else kt_sincos(y0[1] + d[1], &sc.s, &sc.c);
    double o[5];
    rhs<CHART>(P, y0[0] + d[0], sc, y0[2] + d[2], y0[3] + d[3], pt, pp, o);
    for (int i = 0; i < NC; ++i) k[i] = o[i];
}

// The binary32 copies of p_t and p_phi, made opaque at each stage: otherwise the compiler
// forms the per-ray products (2 p_t, 2 a p_t, a p_phi) once per attempt and keeps them live
// through all stages, which at 3 waves per SIMD (168 VGPRs) spills them and reloads them in
// every stage; recomputing them costs three multiplies per stage.
KT_HD void opaque_f32(float& x, float& y) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(x), "+v"(y));
#else
    (void)x; (void)y;
#endif
}

template <int CHART, int NC>
KT_HD void stage_f32(const Knds<float>& Pf, const double y0[4], const double d[4], const SinCos<double>& sc0,
                     double pt, double pp, double k[NC]) {
    SinCos<float> scf;
    if (fabs(d[1]) <= kRotateMax) scf = rotate(SinCos<float>{(float)sc0.s, (float)sc0.c}, (float)d[1]);
    else kt_sincos((float)(y0[1] + d[1]), &scf.s, &scf.c);
    float kf[5];
    float ptf = (float)pt, ppf = (float)pp;
    opaque_f32(ptf, ppf);
    rhs<CHART>(Pf, (float)(y0[0] + d[0]), scf, (float)(y0[2] + d[2]), (float)(y0[3] + d[3]),
               ptf, ppf, kf);
    for (int i = 0; i < NC; ++i) k[i] = (double)kf[i];
}

template <int CHART, int PREC, int NC>
KT_HD void stage(const Knds<double>& P, const Knds<float>& Pf, bool use64, const double y0[4], const double d[4],
                 const SinCos<double>& sc0, double pt, double pp, double k[NC]) {
    if (PREC == PREC_FP64 || use64) {
        SinCos<double> sc;
        stage_f64<CHART, NC>(P, y0, d, sc0, pt, pp, k, sc);
    } else {
        stage_f32<CHART, NC>(Pf, y0, d, sc0, pt, pp, k);
    }
}

// right-hand-side evaluations of an attempt, counted once per attempt (all stages of an
// attempt run in one precision) so that no counter is live through the stages
template <int PREC>
KT_HD void count_rhs(RayState& S, bool use64, int n) {
    if (PREC == PREC_FP64 || use64) S.rhs64 += n;
    else S.rhs32 += n;
}

// stage derivative at y0 + d with the sin/cos of theta given (integrated stage values)
template <int CHART, int PREC, int NC>
KT_HD void stage_given(const Knds<double>& P, const Knds<float>& Pf, bool use64, const double y0[4],
                       const double d[4], const SinCos<double>& sc, double pt, double pp, double k[NC]) {
    if (PREC == PREC_FP64 || use64) {
        double o[5];
        rhs<CHART>(P, y0[0] + d[0], sc, y0[2] + d[2], y0[3] + d[3], pt, pp, o);
        for (int i = 0; i < NC; ++i) k[i] = o[i];
    } else {
        const SinCos<float> scf = {(float)sc.s, (float)sc.c};
        float kf[5];
        float ptf = (float)pt, ppf = (float)pp;
        opaque_f32(ptf, ppf);
        rhs<CHART>(Pf, (float)(y0[0] + d[0]), scf, (float)(y0[2] + d[2]), (float)(y0[3] + d[3]),
                   ptf, ppf, kf);
        for (int i = 0; i < NC; ++i) k[i] = (double)kf[i];
    }
}

// Step-size factor x^(-q) for a finite x > 0 of any binary64 magnitude: the binary exponent
// is split off exactly (x = m 2^e, m in [0.5, 1)), log2 of the mantissa and the final exp2
// are binary32 (~1e-7 relative).  The factor only scales the next trial step; the accept
// test itself is binary64.
KT_HD double pow_neg(double x, float q) {
    int e;
    const double m = frexp(x, &e);
#if defined(__HIP_DEVICE_COMPILE__)
    const float l = (float)e + __builtin_amdgcn_logf((float)m);
    return (double)__builtin_amdgcn_exp2f(-q * l);
#else
    const float l = (float)e + std::log2((float)m);
    return (double)std::exp2(-q * l);
#endif
}

KT_HD SinCos<double> sc_at(const SinCos<double>& sc0, double th0, double dth) {
    SinCos<double> o;
    if (fabs(dth) <= kRotateMax) o = rotate(sc0, dth);
    else kt_sincos(th0 + dth, &o.s, &o.c);
    return o;
}

// ---------------------------------------------------------------------------
//  events after an accepted step (main.cpp:363-431)
// ---------------------------------------------------------------------------
// The reference's KS tracer reports azimuths through atan2 (KS_to_BL_spatial,
// knds_metric.hpp:389-411), i.e. wrapped into [-pi, pi]; BL reports the integrated
// phi.  The wrap is phi - 2 pi rint(phi / 2 pi) with 2 pi in two parts.
template <int CHART> KT_HD double out_phi(double ph) {
    if (CHART == CHART_KS) {
        const double k = rint(ph * 0.15915494309189535);
        return fma(-k, 2.4492935982947064e-16, fma(-k, 6.283185307179586, ph));
    }
    return ph;
}

// Step from (r0, th0, ph0) with derivatives (dr0, dth0) to the new state S over h.
// Returns true if the ray is finished (o filled).
template <int CHART>
KT_HD bool step_events(const FrameDev& F, double r0, double th0, double ph0, double dr0, double dth0,
                       const RayState& S, double hu, uint32_t quirks, RayOut& o, int& hermites) {
    const double halfpi = kPi / 2.0;
    const double r1 = S.y[0], th1 = S.y[1], dr1 = S.k1[0], dth1 = S.k1[1];
    double best = 2.0;
    int ev = 0;
    double d_r = 0.0, d_g = 1.0, d_ph = 0.0;
    const double q0 = th0 - halfpi, q1 = th1 - halfpi;
    const bool near_eq = crosses(q0, q1) || (fmin(fabs(q0), fabs(q1)) < 0.35);
    if (near_eq && hermite_may_reach(r0, r1, dr0, dr1, hu, F.r_in, F.r_out) &&
        hermite_may_touch(th0, th1, dth0, dth1, hu, halfpi)) {
        hermites++;
        double al = 0.0;
        if (hermite_first_root(th0, th1, dth0, dth1, hu, halfpi, al)) {
            if (quirks & Q_EXACT_ROOT) al = hermite_root_polish(th0, th1, dth0, dth1, hu, halfpi, al);
            const double rh = hermite(r0, r1, dr0, dr1, hu, al);
            if (rh >= F.r_in && rh <= F.r_out) {
                d_r = rh;
                d_g = disk_redshift(F, rh, S.pt, S.pp);
                d_ph = out_phi<CHART>(ph0 + al*(S.ph - ph0));
                best = al;
                ev = 1;
            }
        }
    }
    const double r_cut = S.r_cut, r_esc = S.r_esc;
    if (((r0 > r_cut) && (r1 <= r_cut)) || (r1 <= r_cut)) {
        const double dn = r0 - r1;
        double ah = (fabs(dn) > 1e-12) ? ((r0 - r_cut) / dn) : 0.0;
        ah = ah < 0.0 ? 0.0 : (ah > 1.0 ? 1.0 : ah);
        if (ah < best) { best = ah; ev = 2; }
    }
    if (((r0 < r_esc) && (r1 >= r_esc)) || (r1 >= r_esc)) {
        const double dn = r1 - r0;
        double ae = (fabs(dn) > 1e-12) ? ((r_esc - r0) / dn) : 1.0;
        ae = ae < 0.0 ? 0.0 : (ae > 1.0 ? 1.0 : ae);
        if (ae < best) { best = ae; ev = 3; }
    }
    if (ev == 0) return false;
    o.ph_disk = o.th_esc = o.ph_esc = 0.0;
    if (ev == 1) { o.fate = 1; o.r = d_r; o.g = d_g; o.ph_disk = d_ph; }
    else if (ev == 2) { o.fate = 2; o.r = r0 + best*(r1 - r0); o.g = 0.0; }
    else {
        o.fate = 0; o.r = r_esc; o.g = 1.0;
        const double te = th0 + best*(th1 - th0);
        const double pe = out_phi<CHART>(ph0 + best*(S.ph - ph0));
        if (quirks & Q_ESC_POSITIONAL) { o.ph_disk = te; o.th_esc = pe; o.ph_esc = 0.0; }
        else { o.th_esc = te; o.ph_esc = pe; }
    }
    return true;
}

// result of a ray stopped by the reject cap or max_steps (main.cpp:360,433)
template <int CHART>
KT_HD void stopped_result(const RayState& S, uint32_t quirks, RayOut& o) {
    o.fate = 0; o.r = S.y[0]; o.g = 1.0;
    const double ph = out_phi<CHART>(S.ph);
    if (quirks & Q_ESC_POSITIONAL) { o.ph_disk = S.y[1]; o.th_esc = ph; o.ph_esc = 0.0; }
    else { o.ph_disk = 0.0; o.th_esc = S.y[1]; o.ph_esc = ph; }
}

// Constraint renormalisation (binary64) for the mixed-precision path: one Newton
// step of minimal norm in (p_r, p_theta) onto N = 2 Sigma H = 0,
//   dp = -N grad_p N / |grad_p N|^2,
// applied only when the drift exceeds 1e-12 of the constraint's scale.
template <int CHART>
KT_HD void renormalise(const Knds<double>& P, RayState& S) {
    const double r = S.y[0], s2 = S.sc.s*S.sc.s, c2 = S.sc.c*S.sc.c;
    const double ra = r*r + P.a2;
    const double dR = ra*(1.0 - P.lam3*r*r) - P.M2*r + P.Q2;
    const double dT = 1.0 + P.lam3a2*c2;
    const double K = ra*S.pt + P.a*S.pp;
    const double L = P.a*s2*S.pt + S.pp;
    const double pr = S.y[2], pth = S.y[3];
    const double tL = P.xi2*L*L*kt_rcp(dT*s2);
    double N, gr, scale;
    if (CHART == CHART_BL) {
        const double tK = P.xi2*K*K*kt_rcp(dR);
        N = dR*pr*pr + dT*pth*pth - tK + tL;
        gr = 2.0*dR*pr;
        scale = fabs(dR*pr*pr) + dT*pth*pth + fabs(tK) + tL;
    } else {
        const double xk = 2.0*P.xi*K*pr;
        N = dR*pr*pr + xk + dT*pth*pth + tL;
        gr = 2.0*(dR*pr + P.xi*K);
        scale = fabs(dR*pr*pr) + fabs(xk) + dT*pth*pth + tL;
    }
    const double gt = 2.0*dT*pth;
    const double g2 = gr*gr + gt*gt;
    if (!(fabs(N) > 1e-12*scale) || !(g2 > 0.0)) return;
    const double f = N * kt_rcp(g2);
    if (!isfinite(f)) return;
    S.y[2] = pr - f*gr;
    S.y[3] = pth - f*gt;
}

enum { STEP_REJECTED = 0, STEP_ACCEPTED = 1, STEP_DONE = 2 };

// bookkeeping after an accepted step: constraint projection, events, max_steps
template <int CHART, int PREC>
KT_HD int after_accept(const FrameDev& F, const Knds<double>& P, const RunOpts& O, RayState& S, double r0,
                       double th0, double ph0, double dr0, double dth0, double hu, RayOut& o, bool step64) {
    S.n_rej = 0;
    S.n_acc++;
    // binary32 steps only: a step the wave took in binary64 drifts off H = 0 no more than
    // the binary64 path, which projects nothing
    if (PREC == PREC_MIXED && !step64) renormalise<CHART>(P, S);
    if (PREC == PREC_MIXED && step64 && S.hold > 0) S.hold--;
    if (step_events<CHART>(F, r0, th0, ph0, dr0, dth0, S, hu, O.quirks, o, S.hermites)) return STEP_DONE;
    if (S.n_acc >= O.max_steps) { stopped_result<CHART>(S, O.quirks, o); return STEP_DONE; }
    return STEP_ACCEPTED;
}

// Mixed precision: a rejected binary32 attempt says little about the step size (its error
// estimate may be binary32 stage noise, which does not shrink with h as the truncation error
// does), so the same step is retried in binary64 and the lane stays in binary64 for
// O.isl_hold accepted steps, instead of shrinking h through a cascade of noisy rejections.
template <int PREC>
KT_HD void mixed_retry(const RunOpts& O, bool use64, double h, RayState& S) {
    if (PREC == PREC_MIXED && !use64 && O.isl_hold > 0) { S.h = h; S.hold = O.isl_hold; }
}

KT_HD int after_reject(const RunOpts& O, RayState& S) {
    S.rej_total++;
    if (!isfinite(S.h) || ++S.n_rej > 64) return STEP_DONE;
    return STEP_REJECTED;
}

// stage derivative times h at the stage point z (r, -, p_r, p_theta; z[1] is not read), the
// sin/cos of theta given (integrated stage values)
template <int CHART, int PREC, int NC>
KT_HD void stage_at_h(const Knds<double>& P, const Knds<float>& Pf, bool use64, const double z[4],
                      const SinCos<double>& sc, double pt, double pp, double h, double hk[NC]) {
    if (PREC == PREC_FP64 || use64) {
        double o[5];
        rhs_h<CHART>(P, z[0], sc, z[2], z[3], pt, pp, h, o);
        for (int i = 0; i < NC; ++i) hk[i] = o[i];
    } else {
        const SinCos<float> scf = {(float)sc.s, (float)sc.c};
        float kf[5];
        float ptf = (float)pt, ppf = (float)pp;
        opaque_f32(ptf, ppf);
        rhs_h<CHART>(Pf, (float)z[0], scf, (float)z[2], (float)z[3], ptf, ppf, (float)h, kf);
        for (int i = 0; i < NC; ++i) hk[i] = (double)kf[i];
    }
}

// FSAL stage at the 5th-order point z (theta = th0 + dth): f(z) unscaled, and sin/cos of its theta
template <int CHART, int PREC>
KT_HD void fsal_stage(const Knds<double>& P, const Knds<float>& Pf, bool use64, const double z[4], double th0,
                      double dth, const SinCos<double>& sc0, double pt, double pp, double k[5], SinCos<double>& sc) {
    if (PREC == PREC_FP64 || use64) {
        if (fabs(dth) <= kRotateMax) sc = rotate(sc0, dth);
        else kt_sincos(th0 + dth, &sc.s, &sc.c);
        rhs<CHART>(P, z[0], sc, z[2], z[3], pt, pp, k);
    } else {
        SinCos<float> scf;
        if (fabs(dth) <= kRotateMax) scf = rotate(SinCos<float>{(float)sc0.s, (float)sc0.c}, (float)dth);
        else kt_sincos((float)(th0 + dth), &scf.s, &scf.c);
        float kf[5];
        float ptf = (float)pt, ppf = (float)pp;
        opaque_f32(ptf, ppf);
        rhs<CHART>(Pf, (float)z[0], scf, (float)z[2], (float)z[3], ptf, ppf, kf);
        for (int i = 0; i < 5; ++i) k[i] = (double)kf[i];
        sc = sc_at(sc0, th0, dth);
    }
}

// ---------------------------------------------------------------------------
//  DOPRI5 attempt (geodesic.hpp:312-414) with integrated stage sin/cos.
//
//  The metric depends on theta only through sin and cos, so stages 2-6 take (sin, cos) as
//  two more integrated components,
//    d(sin theta)/dl = cos theta * theta',   d(cos theta)/dl = -sin theta * theta',
//  combined with the same tableau as the other components (a consistent 5th-order method
//  for the augmented system), instead of a Taylor rotation per stage.  The FSAL stage at
//  the accepted point keeps the exact rotation, so every step starts from sin/cos of theta.
//  Stages 2-6 are carried as h k_i (h folded into each right-hand side's final scaling) and
//  the k1 terms take their tableau coefficient times h.  Each stage point, and the 5th-order
//  solution, is accumulated onto y by fused multiply-adds (one per tableau term, no separate
//  sum and add); theta enters the stages only through its sin/cos, so its stage values are
//  not formed, and its 5th-order increment feeds the FSAL rotation.  k1 and the FSAL stage
//  stay unscaled: they are the event derivatives and the next attempt's (or a retry's) k1.
//  The adjoint's replay runs this same function, so it retraces the forward pass bit for bit.
// ---------------------------------------------------------------------------
// One attempt from (y, k1 = f(y), sc = sin/cos of y[1]) over h: the 5th-order solution
// (z[0], z[2], z[3]; z[1] is not formed), its theta and phi increments, the error sum ae
// (without the FSAL term), the FSAL derivative k7 = f(solution) and its sin/cos.
template <int CHART, int PREC, int NC>
KT_HD void dp5_attempt_h(const Knds<double>& P, const Knds<float>& Pf, bool use64, const double y[4],
                         const double k1[5], const SinCos<double>& sc0, double h, double pt, double pp,
                         double z[4], double& dth5, double& dph5, double ae[4], double k7[5],
                         SinCos<double>& sc7) {
    typedef DP5 C;
    double k2[NC], k3[NC], k4[NC];              // h k_i
    dph5 = 0.0;
    z[1] = 0.0;
    // stage sin/cos and their h-scaled derivatives u = cos h theta', v = -sin h theta'
    const double s1 = sc0.s, c1 = sc0.c;
    const double hth1 = h*k1[1];
    const double u1 = c1*hth1, v1 = -s1*hth1;
    double u2, v2, u3, v3, u4, v4;
    SinCos<double> sc;
    double ch = C::a21*h;
    for (int i = 0; i < 4; ++i) if (i != 1) z[i] = fma(ch, k1[i], y[i]);
    sc.s = fma(C::a21, u1, s1); sc.c = fma(C::a21, v1, c1);
    stage_at_h<CHART, PREC, NC>(P, Pf, use64, z, sc, pt, pp, h, k2);
    u2 = sc.c*k2[1]; v2 = -sc.s*k2[1];
    ch = C::a31*h;
    for (int i = 0; i < 4; ++i) if (i != 1) z[i] = fma(C::a32, k2[i], fma(ch, k1[i], y[i]));
    sc.s = fma(C::a32, u2, fma(C::a31, u1, s1)); sc.c = fma(C::a32, v2, fma(C::a31, v1, c1));
    stage_at_h<CHART, PREC, NC>(P, Pf, use64, z, sc, pt, pp, h, k3);
    u3 = sc.c*k3[1]; v3 = -sc.s*k3[1];
    ch = C::a41*h;
    for (int i = 0; i < 4; ++i) if (i != 1) z[i] = fma(C::a43, k3[i], fma(C::a42, k2[i], fma(ch, k1[i], y[i])));
    sc.s = fma(C::a43, u3, fma(C::a42, u2, fma(C::a41, u1, s1)));
    sc.c = fma(C::a43, v3, fma(C::a42, v2, fma(C::a41, v1, c1)));
    stage_at_h<CHART, PREC, NC>(P, Pf, use64, z, sc, pt, pp, h, k4);
    u4 = sc.c*k4[1]; v4 = -sc.s*k4[1];
    {
        double k5[NC];
        ch = C::a51*h;
        for (int i = 0; i < 4; ++i)
            if (i != 1) z[i] = fma(C::a54, k4[i], fma(C::a53, k3[i], fma(C::a52, k2[i], fma(ch, k1[i], y[i]))));
        sc.s = fma(C::a54, u4, fma(C::a53, u3, fma(C::a52, u2, fma(C::a51, u1, s1))));
        sc.c = fma(C::a54, v4, fma(C::a53, v3, fma(C::a52, v2, fma(C::a51, v1, c1))));
        stage_at_h<CHART, PREC, NC>(P, Pf, use64, z, sc, pt, pp, h, k5);
        const double u5 = sc.c*k5[1], v5 = -sc.s*k5[1];
        sc.s = fma(C::a65, u5, fma(C::a64, u4, fma(C::a63, u3, fma(C::a62, u2, fma(C::a61, u1, s1)))));
        sc.c = fma(C::a65, v5, fma(C::a64, v4, fma(C::a63, v3, fma(C::a62, v2, fma(C::a61, v1, c1)))));
        ch = C::a61*h;
        for (int i = 0; i < 4; ++i)
            if (i != 1)
                z[i] = fma(C::a65, k5[i], fma(C::a64, k4[i], fma(C::a63, k3[i], fma(C::a62, k2[i], fma(ch, k1[i], y[i])))));
        // stage 6 is evaluated at z; the 5th-order solution is accumulated in a5 meanwhile
        double a5[4];
        ch = C::b1*h;
        for (int i = 0; i < 4; ++i)
            if (i != 1) a5[i] = fma(C::b5, k5[i], fma(C::b4, k4[i], fma(C::b3, k3[i], fma(ch, k1[i], y[i]))));
        dth5 = fma(ch, k1[1], C::b3*k3[1] + C::b4*k4[1] + C::b5*k5[1]);
        if (NC == 5) dph5 = fma(ch, k1[NC - 1], C::b3*k3[NC - 1] + C::b4*k4[NC - 1] + C::b5*k5[NC - 1]);
        ch = C::e1*h;
        for (int i = 0; i < 4; ++i) ae[i] = fma(ch, k1[i], C::e3*k3[i] + C::e4*k4[i] + C::e5*k5[i]);
        double k6[NC];
        stage_at_h<CHART, PREC, NC>(P, Pf, use64, z, sc, pt, pp, h, k6);
        for (int i = 0; i < 4; ++i) if (i != 1) z[i] = fma(C::b6, k6[i], a5[i]);
        dth5 = fma(C::b6, k6[1], dth5);
        if (NC == 5) dph5 = fma(C::b6, k6[NC - 1], dph5);
        for (int i = 0; i < 4; ++i) ae[i] = fma(C::e6, k6[i], ae[i]);
    }
    // FSAL stage at the 5th-order point, unscaled; its sin/cos becomes the next base
    fsal_stage<CHART, PREC>(P, Pf, use64, z, y[1], dth5, sc0, pt, pp, k7, sc7);
}

template <int CHART, int PREC, int NC>
KT_HD int dp5_step_h(const FrameDev& F, const Knds<double>& P, const Knds<float>& Pf, const RunOpts& O, bool use64,
                     RayState& S, RayOut& o) {
    typedef DP5 C;
    const double h = S.h;
    const double* y = S.y;
    double z[4], ae[4], k7[5], dth5, dph5;
    SinCos<double> sc7;
    dp5_attempt_h<CHART, PREC, NC>(P, Pf, use64, S.y, S.k1, S.sc, h, S.pt, S.pp, z, dth5, dph5, ae, k7, sc7);
    count_rhs<PREC>(S, use64, 6);      // stages 2-6 and the FSAL stage
    // Scaled RMS of the embedded error, err^2 = mean_i (dy_i / (atol + rtol |y_i|))^2, in
    // binary64 in both precisions: each quotient through the refined reciprocal, and the
    // accept test err <= 1 taken on err^2 (no square root; sqrt is monotone, so the test is
    // the same).
    const double e7h = C::e7*h;
    double e2 = 0.0;
    for (int i = 0; i < 4; ++i) {
        const double q = fma(e7h, k7[i], ae[i]) * kt_rcp(O.atol + O.rtol*fabs(y[i]));
        e2 = fma(q, q, e2);
    }
    const double err2 = e2 * 0.25;
    if (!isfinite(err2)) {
        const double hn = (isfinite(h) && h > 1e-10) ? h*0.5 : 1e-6;
        S.h = hn < 1e-10 ? 1e-10 : hn;
        mixed_retry<PREC>(O, use64, h, S);
        return after_reject(O, S);
    }
    if (err2 <= 1.0 || h < 1e-10) {
        const double r0 = y[0], th0 = y[1], ph0 = S.ph, dr0 = S.k1[0], dth0 = S.k1[1];
        if (NC == 5) S.ph += dph5;
        S.y[0] = z[0]; S.y[1] = th0 + dth5; S.y[2] = z[2]; S.y[3] = z[3];
        for (int i = 0; i < 5; ++i) S.k1[i] = k7[i];
        S.sc = sc7;
        double hn = h * ((err2 > 1e-24) ? 0.9*pow_neg(err2, 0.1f) : 4.0);
        if (!isfinite(hn)) hn = h;
        S.h = hn > 100.0 ? 100.0 : (hn < 1e-10 ? 1e-10 : hn);
        return after_accept<CHART, PREC>(F, P, O, S, r0, th0, ph0, dr0, dth0, h, o, use64);
    }
    double hn = h * 0.9 * pow_neg(err2, 0.125f);
    if (!isfinite(hn)) hn = h*0.5;
    S.h = hn < 1e-10 ? 1e-10 : hn;
    mixed_retry<PREC>(O, use64, h, S);
    return after_reject(O, S);
}

// Classic RK4 over [0, h] from y0 with f(y0) = k0 (geodesic.hpp:169-194): returns
// the increment of (r, theta, p_r, p_theta, phi).
// SCI = false: each stage rotates the base sin/cos to its theta (the trace kernels' RK4, whose
//   step sequence then follows the reference's closely enough that escape angles, interpolated
//   inside the escaping step, match its images);
// SCI = true: as in the DOPRI5 attempt, stages 2-4 take (sin theta, cos theta) as two more
//   integrated components, d sin/dl = cos theta', d cos/dl = -sin theta', with the RK4 stage
//   offsets (a 4th-order method for the augmented system) instead of a rotation series per
//   stage (the bundle kernel).  The accepted point's sin/cos is rotated exactly by the caller.
template <int CHART, int PREC, bool SCI = false>
KT_HD void rk4_incr(const Knds<double>& P, const Knds<float>& Pf, bool use64, const double y0[4],
                    const SinCos<double>& sc0, const double k0[5], double h, double pt, double pp,
                    double inc[5]) {
    double k2[5], k3[5], k4[5], d[4];
    if (SCI) {
        const double hh = 0.5*h;
        SinCos<double> sc, sp;
        for (int i = 0; i < 4; ++i) d[i] = hh*k0[i];
        sc.s = fma(hh, sc0.c*k0[1], sc0.s); sc.c = fma(-hh, sc0.s*k0[1], sc0.c);
        stage_given<CHART, PREC, 5>(P, Pf, use64, y0, d, sc, pt, pp, k2);
        sp = sc;
        for (int i = 0; i < 4; ++i) d[i] = hh*k2[i];
        sc.s = fma(hh, sp.c*k2[1], sc0.s); sc.c = fma(-hh, sp.s*k2[1], sc0.c);
        stage_given<CHART, PREC, 5>(P, Pf, use64, y0, d, sc, pt, pp, k3);
        sp = sc;
        for (int i = 0; i < 4; ++i) d[i] = h*k3[i];
        sc.s = fma(h, sp.c*k3[1], sc0.s); sc.c = fma(-h, sp.s*k3[1], sc0.c);
        stage_given<CHART, PREC, 5>(P, Pf, use64, y0, d, sc, pt, pp, k4);
    } else {
        for (int i = 0; i < 4; ++i) d[i] = 0.5*h*k0[i];
        stage<CHART, PREC, 5>(P, Pf, use64, y0, d, sc0, pt, pp, k2);
        for (int i = 0; i < 4; ++i) d[i] = 0.5*h*k2[i];
        stage<CHART, PREC, 5>(P, Pf, use64, y0, d, sc0, pt, pp, k3);
        for (int i = 0; i < 4; ++i) d[i] = h*k3[i];
        stage<CHART, PREC, 5>(P, Pf, use64, y0, d, sc0, pt, pp, k4);
    }
    for (int i = 0; i < 5; ++i) inc[i] = h/6.0*(k0[i] + 2.0*k2[i] + 2.0*k3[i] + k4[i]);
}

// RK4 step doubling attempt (geodesic.hpp:204-257); phi is always advanced.
template <int CHART, int PREC>
KT_HD int rk4_step(const FrameDev& F, const Knds<double>& P, const Knds<float>& Pf, const RunOpts& O, bool use64,
                   RayState& S, RayOut& o) {
    const double h = S.h;
    const double tol = O.rtol;
    double incA[5], incB1[5], incB2[5];
    rk4_incr<CHART, PREC>(P, Pf, use64, S.y, S.sc, S.k1, h, S.pt, S.pp, incA);
    rk4_incr<CHART, PREC>(P, Pf, use64, S.y, S.sc, S.k1, 0.5*h, S.pt, S.pp, incB1);
    double yB1[4];
    for (int i = 0; i < 4; ++i) yB1[i] = S.y[i] + incB1[i];
    const SinCos<double> scB1 = sc_at(S.sc, S.y[1], incB1[1]);
    double kB1[5];
    {
        const double z[4] = {0.0, 0.0, 0.0, 0.0};
        stage<CHART, PREC, 5>(P, Pf, use64, yB1, z, scB1, S.pt, S.pp, kB1);
    }
    rk4_incr<CHART, PREC>(P, Pf, use64, yB1, scB1, kB1, 0.5*h, S.pt, S.pp, incB2);
    count_rhs<PREC>(S, use64, 10);     // three RK4 increments and f at the B1 point
    double e2 = 0.0;
    for (int i = 0; i < 4; ++i) { const double t = (S.y[i] + incA[i]) - (yB1[i] + incB2[i]); e2 += t*t; }
    const double err = sqrt(e2) / 15.0;
    if (!isfinite(err)) {
        const double hn = (isfinite(h) && h > 1e-10) ? h*0.5 : 1e-6;
        S.h = hn < 1e-10 ? 1e-10 : hn;
        mixed_retry<PREC>(O, use64, h, S);
        return after_reject(O, S);
    }
    if (err < tol || h < 1e-10) {
        const double r0 = S.y[0], th0 = S.y[1], ph0 = S.ph, dr0 = S.k1[0], dth0 = S.k1[1];
        S.ph = S.ph + incB1[4] + incB2[4];
        const double dth = (yB1[1] - S.y[1]) + incB2[1];
        for (int i = 0; i < 4; ++i) S.y[i] = yB1[i] + incB2[i];
        S.sc = sc_at(S.sc, th0, dth);
        double hn = h * ((err > 1e-14) ? 0.9*pow_neg(err / tol, 0.2f) : 4.0);
        if (!isfinite(hn)) hn = h;
        S.h = hn > 100.0 ? 100.0 : (hn < 1e-10 ? 1e-10 : hn);
        // f(y_new): event derivatives and first stage of the next attempt
        const double z[4] = {0.0, 0.0, 0.0, 0.0};
        stage<CHART, PREC, 5>(P, Pf, use64, S.y, z, S.sc, S.pt, S.pp, S.k1);
        count_rhs<PREC>(S, use64, 1);
        return after_accept<CHART, PREC>(F, P, O, S, r0, th0, ph0, dr0, dth0, h, o, use64);
    }
    const double half = h*0.5;
    double hn = h * 0.9 * pow_neg(err / tol, 0.25f);
    if (!isfinite(hn)) hn = half;
    if (hn > half) hn = half;
    S.h = hn < 1e-10 ? 1e-10 : hn;
    mixed_retry<PREC>(O, use64, h, S);
    return after_reject(O, S);
}

// Initialise a ray at pixel (px, py) of frame F.
template <int CHART, int PREC>
KT_HD void ray_init(const FrameDev& F, const RunOpts& O, int px, int py, RayState& S) {
    double st[7];
    camera_ray(F, (double)px, (double)py, st);
    if (CHART == CHART_KS) bl_to_ks_covector(F, st);
    S.y[0] = st[0]; S.y[1] = st[1]; S.y[2] = st[3]; S.y[3] = st[4];
    S.ph = st[2]; S.pt = st[5]; S.pp = st[6];
    S.h = O.step_init < 1e-10 ? 1e-10 : O.step_init;
    kt_sincos(S.y[1], &S.sc.s, &S.sc.c);
    rhs<CHART>(F.P, S.y[0], S.sc, S.y[2], S.y[3], S.pt, S.pp, S.k1);
    S.r_cut = F.r_cut; S.r_esc = F.r_esc;
    S.n_acc = 0; S.n_rej = 0;
    S.rhs32 = 0; S.rhs64 = 1; S.hermites = 0; S.rej_total = 0;
    S.hold = 0;
}

// Initial state of a ray as the init kernel leaves it in HBM: one 128-byte record per
// pixel, so that the persistent loop refills a lane with loads instead of running the
// camera (tetrad, two square-root solves, the Kerr-Schild covector, sin/cos, first RHS)
// for the whole wave every time one lane needs a new ray.
struct RayInit {
    double y[4], ph, pt, pp, k1[5], s, c, r_cut, r_esc;
};
static_assert(sizeof(RayInit) == 128, "init record");

KT_HD void ray_store_init(const RayState& S, RayInit& R) {
    for (int i = 0; i < 4; ++i) R.y[i] = S.y[i];
    R.ph = S.ph; R.pt = S.pt; R.pp = S.pp;
    for (int i = 0; i < 5; ++i) R.k1[i] = S.k1[i];
    R.s = S.sc.s; R.c = S.sc.c;
    R.r_cut = S.r_cut; R.r_esc = S.r_esc;
}

KT_HD void ray_load_init(const RayInit& R, const RunOpts& O, RayState& S) {
    for (int i = 0; i < 4; ++i) S.y[i] = R.y[i];
    S.ph = R.ph; S.pt = R.pt; S.pp = R.pp;
    for (int i = 0; i < 5; ++i) S.k1[i] = R.k1[i];
    S.sc.s = R.s; S.sc.c = R.c;
    S.r_cut = R.r_cut; S.r_esc = R.r_esc;
    S.h = O.step_init < 1e-10 ? 1e-10 : O.step_init;
    S.n_acc = 0; S.n_rej = 0;
    S.rhs32 = 0; S.rhs64 = 1; S.hermites = 0; S.rej_total = 0;
    S.hold = 0;
}

// Does this lane need a binary64 step in the mixed-precision path?  The FP32
// stage noise of the embedded error estimate, ~isl_noise * h |k|, must stay
// well below the error scale atol + rtol |y|; near the horizon and the polar
// axis the FP32 Hamiltonian loses too many digits.
KT_HD bool wants_fp64(const FrameDev& F, const RunOpts& O, const RayState& S) {
    if (S.hold > 0) return true;
    if (S.y[0] < O.isl_r * F.r_h) return true;
    if (S.sc.s*S.sc.s < 1e-3) return true;
    for (int i = 0; i < 4; ++i) {
        const double sc = O.atol + O.rtol*fabs(S.y[i]);
        if (O.isl_noise * S.h * fabs(S.k1[i]) > 0.25*sc) return true;
    }
    return false;
}

// Pack a result into the 28-byte GeoPixel (main.cpp:2130-2147).
KT_HD void pack_geopixel(const RayOut& o, uint32_t w[7]) {
    union { float f; uint32_t u; } cv;
    w[0] = (uint32_t)(o.fate & 0xFF);
    cv.f = (float)o.r; w[1] = cv.u;
    cv.f = (float)o.g; w[2] = cv.u;
    cv.f = 1.0f; w[3] = cv.u;
    cv.f = (float)o.ph_disk; w[4] = cv.u;
    cv.f = (float)o.th_esc; w[5] = cv.u;
    cv.f = (float)o.ph_esc; w[6] = cv.u;
}

KT_HD void pack_geopixel_m(const RayOut& o, double magnif, uint32_t w[7]) {
    pack_geopixel(o, w);
    union { float f; uint32_t u; } cv;
    cv.f = (float)magnif;
    w[3] = cv.u;
}

// One attempt + bookkeeping for an active ray; returns true when the ray is done.
// PHI: DOPRI5 advances phi (false reproduces the reference's frozen phi).
template <int CHART, int METHOD, int PREC, bool PHI>
KT_HD bool ray_advance(const FrameDev& F, const Knds<double>& P, const Knds<float>& Pf, const RunOpts& O,
                       bool use64, RayState& S, RayOut& o) {
    int rc;
    if (METHOD == METHOD_DP5) rc = dp5_step_h<CHART, PREC, PHI ? 5 : 4>(F, P, Pf, O, use64, S, o);
    else rc = rk4_step<CHART, PREC>(F, P, Pf, O, use64, S, o);
    if (rc != STEP_DONE) return false;
    if (S.n_rej > 64 || !isfinite(S.h)) stopped_result<CHART>(S, O.quirks, o);   // reject cap
    return true;
}

}  // namespace kt
