// ============================================================================
//  kt_bundle.hpp -- Jacobi ray bundles: the geodesic plus its 4x2 deviation
//  matrix W, dW/dl = M W with M = J_s Hess(H) = d f / d z (f = J_s grad H).
//
//  The reference forms M from a 33-point finite-difference Hessian of H
//  (ray_bundle.hpp:56-126); here f and M W come out of ONE evaluation of the
//  analytic right-hand side in forward-mode dual numbers carrying the two
//  columns of W as tangents (exact Jacobian-vector products).
//
//  Reference anchors (KerrTraceCpp2 tree):
//    bundle state / RHS ....... ray_bundle.hpp:131-164
//    bundle RK4 + doubling .... ray_bundle.hpp:167-263 (error on the geodesic part)
//    initial W (half-pixel FD of the camera ray) .... ray_bundle.hpp:278-298
//    trace + events ........... ray_bundle.hpp:301-436 (magnification |det| of the
//                               interpolated (dr, dtheta) rows, floor 1e-12; signed
//                               Doppler factor; horizon / reject cap -> empty result)
//    GeoPixel fill ............ main.cpp:2102-2111
// ============================================================================
#pragma once
#include "kt_trace.hpp"

namespace kt {

// scalar fused multiply-add, bound before the dual overloads below hide the name
KT_HD double dfma(double a, double b, double c) { return fma(a, b, c); }

// value + two tangent directions
struct D2 {
    double v, a, b;
    KT_HD D2() : v(0.0), a(0.0), b(0.0) {}
    KT_HD D2(double v_, double a_, double b_) : v(v_), a(a_), b(b_) {}
};
KT_HD D2 operator+(const D2& x, const D2& y) { return D2(x.v + y.v, x.a + y.a, x.b + y.b); }
KT_HD D2 operator+(const D2& x, double y) { return D2(x.v + y, x.a, x.b); }
KT_HD D2 operator+(double x, const D2& y) { return D2(x + y.v, y.a, y.b); }
KT_HD D2 operator-(const D2& x, const D2& y) { return D2(x.v - y.v, x.a - y.a, x.b - y.b); }
KT_HD D2 operator-(const D2& x) { return D2(-x.v, -x.a, -x.b); }
KT_HD D2 operator*(const D2& x, const D2& y) {
    return D2(x.v*y.v, dfma(x.a, y.v, x.v*y.a), dfma(x.b, y.v, x.v*y.b));
}
KT_HD D2 operator*(const D2& x, double y) { return D2(x.v*y, x.a*y, x.b*y); }
KT_HD D2 operator*(double x, const D2& y) { return D2(x*y.v, x*y.a, x*y.b); }
KT_HD D2 fma(const D2& x, const D2& y, const D2& z) {
    return D2(dfma(x.v, y.v, z.v), dfma(x.a, y.v, dfma(x.v, y.a, z.a)), dfma(x.b, y.v, dfma(x.v, y.b, z.b)));
}
KT_HD D2 fma(double x, const D2& y, const D2& z) { return D2(dfma(x, y.v, z.v), dfma(x, y.a, z.a), dfma(x, y.b, z.b)); }
KT_HD D2 fma(const D2& x, double y, const D2& z) { return D2(dfma(x.v, y, z.v), dfma(x.a, y, z.a), dfma(x.b, y, z.b)); }
KT_HD D2 fma(const D2& x, const D2& y, double z) {
    return D2(dfma(x.v, y.v, z), dfma(x.a, y.v, x.v*y.a), dfma(x.b, y.v, x.v*y.b));
}
KT_HD D2 fma(double x, const D2& y, double z) { return D2(dfma(x, y.v, z), x*y.a, x*y.b); }
KT_HD D2 fma(const D2& x, double y, double z) { return D2(dfma(x.v, y, z), x.a*y, x.b*y); }
KT_HD D2 kt_rcp(const D2& x) {
    const double r = kt_rcp(x.v), m = -r*r;
    return D2(r, x.a*m, x.b*m);
}

// Bundle right-hand side at (r, theta, p_r, p_theta) with sin/cos sc of theta:
// f (4 comps + dphi/dl) and M W (4 x 2).
KT_HD void bundle_rhs(const Knds<double>& P, const double z[4], const SinCos<double>& sc, double pt, double pp,
                      const double W[4][2], double f[5], double dW[4][2]) {
    const D2 r(z[0], W[0][0], W[0][1]);
    const D2 pr(z[2], W[2][0], W[2][1]);
    const D2 pth(z[3], W[3][0], W[3][1]);
    SinCos<D2> scd;
    scd.s = D2(sc.s, sc.c*W[1][0], sc.c*W[1][1]);
    scd.c = D2(sc.c, -sc.s*W[1][0], -sc.s*W[1][1]);
    D2 o[5];
    rhs<CHART_BL, D2, double>(P, r, scd, pr, pth, pt, pp, o);
    for (int i = 0; i < 4; ++i) { f[i] = o[i].v; dW[i][0] = o[i].a; dW[i][1] = o[i].b; }
    f[4] = o[4].v;
}

struct BundleState {
    RayState g;            // geodesic (y, phi, p_t, p_phi, h, f(y) in g.k1, sin/cos, counters)
    double W[4][2];        // deviation matrix, columns: d/d alpha, d/d beta
    double dW[4][2];       // M W at the current state (first stage of the next attempt)
};

// classic RK4 of the coupled system over h from (y0, W0) with first stage (k0, dW0).
// One stage per trip of a loop that is not unrolled: the weighted stage sum is
// accumulated as the stages come (same order of additions as h/6 (k0 + 2 k2 + 2 k3 + k4)),
// so one stage's derivatives are live at a time and the compiler cannot interleave the
// dual-number right-hand sides of different stages (which exhausts the register file).
// SCI: the stage sin/cos of theta are integrated with the stages (rk4_incr<..., SCI = true> in
// kt_trace.hpp), here and in the full step; otherwise each stage rotates the base sin/cos.
template <bool SCI>
KT_HD void bundle_rk4_incr(const Knds<double>& P, const double y0[4], const SinCos<double>& sc0,
                           const double W0[4][2], const double k0[5], const double dW0[4][2], double h, double pt,
                           double pp, double inc[5], double incW[4][2], int& nrhs) {
    double k[5], dW[4][2], acc[5], accW[4][2];
    for (int i = 0; i < 5; ++i) { k[i] = k0[i]; acc[i] = k0[i]; }
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 2; ++j) { dW[i][j] = dW0[i][j]; accW[i][j] = dW0[i][j]; }
    SinCos<double> sp = sc0;                           // sin/cos the previous stage was evaluated at
#pragma unroll 1
    for (int st = 0; st < 3; ++st) {
        const double c = (st == 2) ? h : 0.5*h;        // stage offset
        const double wgt = (st == 2) ? 1.0 : 2.0;      // stage weight
        double y[4], W[4][2];
        for (int i = 0; i < 4; ++i) { y[i] = y0[i] + c*k[i]; W[i][0] = W0[i][0] + c*dW[i][0]; W[i][1] = W0[i][1] + c*dW[i][1]; }
        SinCos<double> sc;
        if (SCI) {
            sc.s = fma(c, sp.c*k[1], sc0.s);
            sc.c = fma(-c, sp.s*k[1], sc0.c);
            sp = sc;
        } else {
            sc = sc_at(sc0, y0[1], c*k[1]);
        }
        bundle_rhs(P, y, sc, pt, pp, W, k, dW);
        for (int i = 0; i < 5; ++i) acc[i] = acc[i] + wgt*k[i];
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 2; ++j) accW[i][j] = accW[i][j] + wgt*dW[i][j];
    }
    nrhs += 3;
    for (int i = 0; i < 5; ++i) inc[i] = h/6.0*acc[i];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 2; ++j) incW[i][j] = h/6.0*accW[i][j];
}

// Doppler factor of the bundle hit (ray_bundle.hpp:368-377): signed 1 - Omega b
KT_HD double bundle_redshift(const FrameDev& F, double r, double pt, double pp) {
    const double Om = omega_k(F, r);
    const double b = -pp / (-pt);
    double gtt, gtp, gpp;
    lower_equator(F, r, gtt, gtp, gpp);
    const double d2 = -(gtt + 2.0*gtp*Om + gpp*Om*Om);
    const double dv = 1.0 - Om*b;
    double red = (d2 > 0.0 && fabs(dv) > 1e-10) ? sqrt(d2)/dv : 1.0;
    return red < 0.0 ? 0.0 : (red > 20.0 ? 20.0 : red);
}

// empty BundleResult{} (ray_bundle.hpp:266-274) as a GeoPixel, or a black horizon pixel
KT_HD void bundle_empty(uint32_t quirks, bool horizon, RayOut& o, double& magnif) {
    o.fate = (horizon && !(quirks & Q_BUNDLE_HORIZON_ESC)) ? 2 : 0;
    o.r = 0.0; o.g = (o.fate == 2) ? 0.0 : 1.0; o.ph_disk = 0.0; o.th_esc = 0.0; o.ph_esc = 0.0;
    magnif = 1.0;
}

// initial state: geodesic from the camera, W from the half-pixel central differences
// of the camera ray in alpha and beta (ray_bundle.hpp:278-298)
KT_HD void bundle_init(const FrameDev& F, const RunOpts& O, int px, int py, BundleState& B) {
    double al, be, st[7], sp[7], sm[7];
    pixel_angles(F, (double)px, (double)py, al, be);
    camera_shoot(F, al, be, st);
    const double eps = F.fov / F.W * 0.5;          // half a pixel in radians
    const int idx[4] = {0, 1, 3, 4};               // r, theta, p_r, p_theta
    camera_shoot(F, al + eps, be, sp);
    camera_shoot(F, al - eps, be, sm);
    for (int i = 0; i < 4; ++i) B.W[i][0] = (sp[idx[i]] - sm[idx[i]]) / (2.0*eps);
    camera_shoot(F, al, be + eps, sp);
    camera_shoot(F, al, be - eps, sm);
    for (int i = 0; i < 4; ++i) B.W[i][1] = (sp[idx[i]] - sm[idx[i]]) / (2.0*eps);
    RayState& S = B.g;
    S.y[0] = st[0]; S.y[1] = st[1]; S.y[2] = st[3]; S.y[3] = st[4];
    S.ph = st[2]; S.pt = st[5]; S.pp = st[6];
    S.h = O.step_init < 1e-10 ? 1e-10 : O.step_init;
    kt_sincos(S.y[1], &S.sc.s, &S.sc.c);
    bundle_rhs(F.P, S.y, S.sc, S.pt, S.pp, B.W, S.k1, B.dW);
    S.n_acc = 0; S.n_rej = 0;
    S.rhs32 = 0; S.rhs64 = 1; S.hermites = 0; S.rej_total = 0; S.hold = 0;
}

// One RK4-doubling attempt of the bundle + the reference's event logic.
// Returns STEP_REJECTED / STEP_ACCEPTED / STEP_DONE (o, magnif filled).
// SCI (integrated stage sin/cos, 16 % faster at 1920x1080) is taken only with the root-skip quirk
// cleared: with it (ray_bundle.hpp:357-365) an escape is recorded at the first step outside the
// equatorial band that passes r_escape, a point that follows the reference's step sequence, and
// the per-stage rotation tracks that sequence (its images match the reference's; the integrated
// form moves half the escapes of the 64x36 reference frame by more than 1e-3 rad).  With the quirk
// cleared both forms agree with the oracle's escape angles to 4e-4.
template <bool SCI>
KT_HD int bundle_step(const FrameDev& F, const RunOpts& O, BundleState& B, RayOut& o, double& magnif) {
    RayState& S = B.g;
    const Knds<double>& P = F.P;
    const double h = S.h;
    const double tol = O.rtol;
    int nrhs = 0;
    // The full step A only enters the error estimate, which is taken on the geodesic part
    // (ray_bundle.hpp:226-237), and the geodesic does not depend on W: A is the plain
    // binary64 RK4 increment, without the deviation matrix.
    double incA[5];
    rk4_incr<CHART_BL, PREC_FP64, SCI>(P, knds_cast<float>(P), true, S.y, S.sc, S.k1, h, S.pt, S.pp, incA);
    nrhs += 3;
    // the two half steps B1, B2, one per trip (not unrolled); B2 starts from the
    // derivatives at the B1 point
    double yb[4], Wb[4][2], kb[5], dWb[4][2], inc[5], incW[4][2];
    SinCos<double> scb = S.sc;
    for (int i = 0; i < 4; ++i) { yb[i] = S.y[i]; Wb[i][0] = B.W[i][0]; Wb[i][1] = B.W[i][1]; }
    for (int i = 0; i < 5; ++i) kb[i] = S.k1[i];
    for (int i = 0; i < 4; ++i) { dWb[i][0] = B.dW[i][0]; dWb[i][1] = B.dW[i][1]; }
    double yB1_th = 0.0, ph_inc = 0.0;
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
        if (half == 1) {
            for (int i = 0; i < 4; ++i) { yb[i] = yb[i] + inc[i]; Wb[i][0] = Wb[i][0] + incW[i][0]; Wb[i][1] = Wb[i][1] + incW[i][1]; }
            scb = sc_at(S.sc, S.y[1], inc[1]);
            bundle_rhs(P, yb, scb, S.pt, S.pp, Wb, kb, dWb);
            nrhs++;
            yB1_th = yb[1];
        }
        bundle_rk4_incr<SCI>(P, yb, scb, Wb, kb, dWb, 0.5*h, S.pt, S.pp, inc, incW, nrhs);
        ph_inc = (half == 0) ? inc[4] : ph_inc;
    }
    // here yb, Wb = the B1 point and inc, incW = the B2 increment
    S.rhs64 += nrhs;
    double e2 = 0.0;
    for (int i = 0; i < 4; ++i) { const double t = (S.y[i] + incA[i]) - (yb[i] + inc[i]); e2 += t*t; }
    const double err = sqrt(e2) / 15.0;
    if (!isfinite(err)) {
        const double hn = (isfinite(h) && h > 1e-10) ? h*0.5 : 1e-6;
        S.h = hn < 1e-10 ? 1e-10 : hn;
        const int rc = after_reject(O, S);
        if (rc == STEP_DONE) bundle_empty(O.quirks, false, o, magnif);
        return rc;
    }
    if (!(err < tol || h < 1e-10)) {
        const double half = h*0.5;
        double hn = h * 0.9 * pow_neg(err / tol, 0.25f);
        if (!isfinite(hn)) hn = half;
        if (hn > half) hn = half;
        S.h = hn < 1e-10 ? 1e-10 : hn;
        const int rc = after_reject(O, S);
        if (rc == STEP_DONE) bundle_empty(O.quirks, false, o, magnif);
        return rc;
    }
    // accepted: commit
    const double r0 = S.y[0], th0 = S.y[1], ph0 = S.ph, dr0 = S.k1[0], dth0 = S.k1[1];
    double W0[4][2];
    for (int i = 0; i < 4; ++i) { W0[i][0] = B.W[i][0]; W0[i][1] = B.W[i][1]; }
    S.ph = S.ph + ph_inc + inc[4];
    const double dth = (yB1_th - S.y[1]) + inc[1];
    for (int i = 0; i < 4; ++i) {
        S.y[i] = yb[i] + inc[i];
        B.W[i][0] = Wb[i][0] + incW[i][0];
        B.W[i][1] = Wb[i][1] + incW[i][1];
    }
    S.sc = sc_at(S.sc, th0, dth);
    double hn = h * ((err > 1e-14) ? 0.9*pow_neg(err / tol, 0.2f) : 4.0);
    if (!isfinite(hn)) hn = h;
    S.h = hn > 100.0 ? 100.0 : (hn < 1e-10 ? 1e-10 : hn);
    bundle_rhs(P, S.y, S.sc, S.pt, S.pp, B.W, S.k1, B.dW);
    S.rhs64++;
    S.n_rej = 0;
    S.n_acc++;

    // events (ray_bundle.hpp:335-433)
    const double halfpi = kPi / 2.0;
    const double r1 = S.y[0], th1 = S.y[1];
    double best = 2.0;
    int ev = 0;
    double d_r = 0.0, d_g = 1.0, d_det = 1.0, d_ph = 0.0;
    const double q0 = th0 - halfpi, q1 = th1 - halfpi;
    const bool near_eq = crosses(q0, q1) || (fmin(fabs(q0), fabs(q1)) < 0.35);
    bool skip = false;
    // without the root-skip quirk a search that cannot land on the disk changes nothing
    const bool q_skip = (O.quirks & Q_BUNDLE_ROOT_SKIP) != 0;
    if (near_eq && (q_skip || hermite_may_reach(r0, r1, dr0, S.k1[0], h, F.r_in, F.r_out))) {
        double al = 0.0;
        bool found = false;
        if (hermite_may_touch(th0, th1, dth0, S.k1[1], h, halfpi)) {
            S.hermites++;
            found = hermite_first_root(th0, th1, dth0, S.k1[1], h, halfpi, al);
        }
        if (!found) {
            skip = q_skip;                                      // `continue` of ray_bundle.hpp:360-364
        } else {
            const double rh = hermite(r0, r1, dr0, S.k1[0], h, al);
            if (rh >= F.r_in && rh <= F.r_out) {
                const double J00 = W0[0][0] + al*(B.W[0][0] - W0[0][0]);
                const double J01 = W0[0][1] + al*(B.W[0][1] - W0[0][1]);
                const double J10 = W0[1][0] + al*(B.W[1][0] - W0[1][0]);
                const double J11 = W0[1][1] + al*(B.W[1][1] - W0[1][1]);
                double det = fabs(J00*J11 - J01*J10);
                d_det = det < 1e-12 ? 1e-12 : det;
                d_r = rh;
                d_g = bundle_redshift(F, rh, S.pt, S.pp);
                d_ph = ph0 + al*(S.ph - ph0);
                best = al;
                ev = 1;
            }
        }
    }
    if (!skip) {
        if (((r0 > F.r_cut) && (r1 <= F.r_cut)) || (r1 <= F.r_cut)) {
            const double dn = r0 - r1;
            double ah = (fabs(dn) > 1e-12) ? ((r0 - F.r_cut) / dn) : 0.0;
            ah = ah < 0.0 ? 0.0 : (ah > 1.0 ? 1.0 : ah);
            if (ah < best) { best = ah; ev = 2; }
        }
        if (((r0 < F.r_esc) && (r1 >= F.r_esc)) || (r1 >= F.r_esc)) {
            const double dn = r1 - r0;
            double ae = (fabs(dn) > 1e-12) ? ((F.r_esc - r0) / dn) : 1.0;
            ae = ae < 0.0 ? 0.0 : (ae > 1.0 ? 1.0 : ae);
            if (ae < best) { best = ae; ev = 3; }
        }
        if (ev == 1) {
            o.fate = 1; o.r = d_r; o.g = d_g; o.ph_disk = d_ph; o.th_esc = 0.0; o.ph_esc = 0.0;
            magnif = d_det;
            return STEP_DONE;
        }
        if (ev == 2) { bundle_empty(O.quirks, true, o, magnif); return STEP_DONE; }
        if (ev == 3) {
            o.fate = 0; o.r = 0.0; o.g = 1.0; o.ph_disk = 0.0;
            o.th_esc = th0 + best*(th1 - th0);
            o.ph_esc = ph0 + best*(S.ph - ph0);
            magnif = 1.0;
            return STEP_DONE;
        }
    }
    if (S.n_acc >= O.max_steps) {       // {false, r, 1, 1}: outcome 0, r cleared, angles 0
        o.fate = 0; o.r = 0.0; o.g = 1.0; o.ph_disk = 0.0; o.th_esc = 0.0; o.ph_esc = 0.0;
        magnif = 1.0;
        return STEP_DONE;
    }
    return STEP_ACCEPTED;
}

}  // namespace kt
