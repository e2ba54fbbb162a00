// ============================================================================
//  kt_adjoint.hpp -- reverse-mode derivatives of the per-pixel disk intensity
//
//    I(pixel) = PT(r_hit)/PT(3 r_isco) * gc^4 * (1 + 0.85 clamp(1 - gc, 0, 1)),
//    gc = clamp(g, 0.1, 10)    (the shading intensity of main.cpp:1799-1810 with
//                               magnification 1; 0 off the disk)
//
//  with respect to p = (a, Q, Lambda, theta_obs [deg], r_obs), for the binary64
//  DOPRI5 path.  The forward pass records every accepted step size and a state
//  checkpoint (r, theta, p_r, p_theta, sin theta, cos theta) every kCkptEvery
//  accepted steps.  The reverse pass walks the checkpoints backwards: it replays
//  one segment from its checkpoint (the same arithmetic as the forward step),
//  then propagates the adjoint through each step in reverse:
//
//    y+ = y + h sum_i b_i k_i,   k_i = f(Y_i),   Y_i = y + h sum_{j<i} a_ij k_j
//    mu_i   = h b_i lam + h sum_{m>i} a_mi nu_m        (adjoint of k_i)
//    nu_i   = J_y(Y_i)^T mu_i                          (adjoint of Y_i)
//    lam   <- lam + sum_i nu_i,  grad_c += J_c^T mu_i,  grad_P += J_P^T mu_i
//
//  with the accepted step sizes held fixed (the discrete adjoint of the taken
//  steps; the controller's dependence on p enters only at the tolerance level).
//  The products J^T mu come from a hand-written reverse mode of the right-hand side
//  (rhs_vjp), pinned against the forward-mode Jacobian in dual numbers
//  (rhs_jacobian).  The seed differentiates the disk event: the Hermite root alpha of
//  theta(alpha) = pi/2 by the implicit function theorem, r_hit = Hermite r at
//  alpha, then redshift and intensity.  The differentiable path converges that
//  root (Q_EXACT_ROOT); the reference's 8 bisections stop 2^-12 short, which
//  leaves r_hit piecewise constant in p.  The initial state's Jacobian (camera
//  tetrad, p_t from H = 0, the Kerr-Schild covector) comes from the templated
//  camera map in dual numbers.  r_isco, which normalises the emissivity, moves
//  with p through r_+ (isco_gradient).
// ============================================================================
#pragma once
#include "kt_dual.hpp"
#include "kt_shade.hpp"
#include "kt_trace.hpp"

namespace kt {

static constexpr int kCkptEvery = 16;
enum { AP_A = 0, AP_Q, AP_L, AP_TH, AP_R, AP_N };

// ---------------------------------------------------------------------------
//  templated mirrors of the camera / covector / redshift maps (the binary64
//  kernels' own functions are these with T = double; a host test pins that)
// ---------------------------------------------------------------------------
template <class T> struct ObsT {
    T gtt, gtp, grr, gth, gpp, U00, U03, U11, U22, U33, sq_rr, sq_th, u0, ep, et;
    bool ok;
};

template <class T> KT_HD Knds<T> knds_t(double M, const T& a, const T& Q, const T& L) {
    Knds<T> P;
    P.a = a; P.a2 = a*a; P.M = T(M); P.M2 = T(2.0*M); P.Q2 = Q*Q;
    P.lam3 = L/3.0; P.lam3a2 = L*a*a/3.0;
    P.xi = 1.0 + L*a*a/3.0; P.xi2 = P.xi*P.xi;
    return P;
}

// static observer constants at (r, theta) (frame_camera_consts, camera.hpp:66-142)
template <class T> KT_HD void observer_t(double M, const T& a, const T& Q, const T& L, const T& r, const T& th,
                                         ObsT<T>& C) {
    T st, ct;
    kt_sincos(th, &st, &ct);
    const T sg = r*r + a*a*ct*ct;
    const T dr = (r*r + a*a) * (1.0 - L*r*r/3.0) - 2.0*M*r + Q*Q;
    const T dt = 1.0 + L*a*a*ct*ct/3.0;
    const T xi = 1.0 + L*a*a/3.0;
    const T x2 = xi*xi, s2 = st*st, ra = r*r + a*a;
    const T den = sg*x2;
    C.gtt = (-dr + dt*a*a*s2) / den;
    C.gtp = a*s2*(dr - dt*ra) / den;
    C.grr = sg/dr;
    C.gth = sg/dt;
    C.gpp = s2*(dt*ra*ra - dr*a*a*s2) / den;
    const T pre = sg*dr*dt;
    C.U00 = -x2*(dt*ra*ra - dr*a*a*s2) / pre;
    C.U03 = a*x2*(dr - dt*ra) / pre;
    C.U11 = dr/sg;
    C.U22 = dt/sg;
    C.U33 = (val(s2) > 1e-14) ? x2*(dr - dt*a*a*s2) / (pre*s2) : T(0.0);
    C.sq_rr = sqrt(fabs(C.grr));
    C.sq_th = sqrt(fabs(C.gth));
    const T dphi = C.gpp - (C.gtp*C.gtp)/C.gtt;
    C.ok = (val(C.gtt) < -1e-14) && isfinite(val(dphi)) && (val(dphi) > 1e-14) && isfinite(val(C.sq_rr)) &&
           (val(C.sq_rr) > 1e-14) && isfinite(val(C.sq_th)) && (val(C.sq_th) > 1e-14);
    if (C.ok) {
        C.u0 = 1.0 / sqrt(-C.gtt);
        C.ep = 1.0 / sqrt(dphi);
        C.et = -C.gtp/C.gtt * C.ep;
    } else {
        C.u0 = T(0.0); C.ep = T(0.0); C.et = T(0.0);
    }
}

// screen angles -> BL state (camera_shoot): st = r, theta, phi(0), p_r, p_theta, p_t, p_phi
template <class T> KT_HD void camera_t(const ObsT<T>& C, const T& r_obs, const T& th_obs, double al, double be,
                                       T st[7]) {
    double sa, ca, sb, cb;
    kt_sincos(al, &sa, &ca);
    kt_sincos(be, &sb, &cb);
    T ut, ur, uth, uph;
    if (C.ok) {
        const double nr = -ca*cb, nt = -sb, np = -sa*cb;
        ut = C.u0 + np*C.et;
        ur = nr / C.sq_rr;
        uth = nt / C.sq_th;
        uph = np*C.ep;
    } else {
        ut = T(1.0);
        ur = -ca*cb / vmax(C.sq_rr, T(1e-14));
        uth = -sb / vmax(C.sq_th, T(1e-14));
        uph = -sa*cb / sqrt(vmax(fabs(C.gpp), T(1e-14)));
    }
    T pt = C.gtt*ut + C.gtp*uph;
    const T pr = C.grr*ur;
    const T pth = C.gth*uth;
    const T pp = C.gtp*ut + C.gpp*uph;
    const T A = C.U00, B = 2.0*C.U03*pp;
    const T Cq = C.U11*pr*pr + C.U22*pth*pth + C.U33*pp*pp;
    const T disc = B*B - 4.0*A*Cq;
    if (val(disc) >= 0.0 && fabs(val(A)) > 1e-15) {
        const T sq = sqrt(disc);
        const T p1 = (-B - sq) / (2.0*A), p2 = (-B + sq) / (2.0*A);
        pt = (val(p1) < 0.0) ? p1 : p2;
        if (val(pt) > 0.0) pt = vmin(p1, p2);
    }
    st[0] = r_obs; st[1] = th_obs; st[2] = T(0.0);
    st[3] = pr; st[4] = pth; st[5] = pt; st[6] = pp;
}

// BL state -> Kerr-Schild covector (bl_to_ks_covector)
template <class T> KT_HD void bl_to_ks_t(const Knds<T>& P, bool ref_init, T st[7]) {
    const T r = st[0], th = st[1];
    T s, c;
    kt_sincos(th, &s, &c);
    const T s2 = s*s, c2 = c*c, ra = r*r + P.a2;
    const T dR = ra*(1.0 - P.lam3*r*r) - P.M2*r + P.Q2;
    const T dT = 1.0 + P.lam3a2*c2;
    const T pr = st[3], pth = st[4], pt = st[5], pp = st[6];
    if (ref_init) {
        const T Pr = pr;
        const T A = dR - 2.0*P.xi*ra + P.xi2*P.a2*s2/dT;
        const T B = -2.0*dR*Pr + 2.0*P.xi*Pr*ra - 2.0*P.xi*P.a*pp + 2.0*P.xi2*P.a*pp/dT;
        const T C = dR*Pr*Pr + 2.0*P.xi*Pr*P.a*pp + dT*pth*pth + P.xi2*pp*pp/(dT*s2);
        const T disc = B*B - 4.0*A*C;
        if (val(disc) >= 0.0 && fabs(val(A)) > 1e-15) {
            const T sq = sqrt(disc);
            const T p1 = (-B - sq) / (2.0*A), p2 = (-B + sq) / (2.0*A);
            T pT = (val(p1) < 0.0) ? p1 : p2;
            if (val(pT) > 0.0) pT = vmin(p1, p2);
            st[5] = pT;
            st[3] = Pr - pT;
            return;
        }
    }
    const T K = ra*pt + P.a*pp;
    st[3] = pr - P.xi*K/dR;
}

// thin-disk redshift (omega_k, lower_equator, disk_redshift), clamped to [0, 20]
template <class T> KT_HD T disk_redshift_t(double M, const T& a, const T& Q, const T& L, const T& r, const T& pt,
                                           const T& pp) {
    const double s = (val(a) < 0.0) ? 1.0 : -1.0;
    const T Me = M - Q*Q/(2.0*r) + L*a*r*r/3.0;
    const T sqm = sqrt(vmax(Me, T(0.0)));
    const T den = r*sqrt(r) + s*a*sqm;
    const T Om = (fabs(val(den)) < 1e-14) ? T(0.0) : s*sqm/den;
    const T b = -pp / (-pt);
    const double ct = kCosHalfPi;
    const T sg = r*r + a*a*ct*ct;
    const T dr = (r*r + a*a) * (1.0 - L*r*r/3.0) - 2.0*M*r + Q*Q;
    const T dt = 1.0 + L*a*a*ct*ct/3.0;
    const T xi = 1.0 + L*a*a/3.0;
    const T x2 = xi*xi, ra = r*r + a*a, den2 = sg*x2;
    const T gtt = (-dr + dt*a*a) / den2;
    const T gtp = a*(dr - dt*ra) / den2;
    const T gpp = (dt*ra*ra - dr*a*a) / den2;
    const T d2 = -(gtt + 2.0*gtp*Om + gpp*Om*Om);
    T red;
    if (val(d2) <= 0.0) red = T(1.0);
    else {
        const T dv = fabs(1.0 - Om*b);
        if (val(dv) < 1e-8) red = T(20.0);
        else {
            red = sqrt(d2) / dv;
            if (!isfinite(val(red))) red = T(1.0);
        }
    }
    return val(red) < 0.0 ? T(0.0) : (val(red) > 20.0 ? T(20.0) : red);
}

// disk intensity of the shading (main.cpp:1799-1810), magnification 1
template <class T> KT_HD T disk_intensity_t(const T& r, const T& g, const T& risco) {
    if (!(val(r) > val(risco))) return T(0.0);
    const T rp = 3.0*risco;
    const T pk = (1.0 - sqrt(risco/rp)) / (rp*rp*rp);
    if (!(val(pk) > 0.0)) return T(0.0);
    const T x = sqrt(risco/r);
    const T pt = (1.0 - x) / (r*r*r);
    const T gc = val(g) < 0.1 ? T(0.1) : (val(g) > 10.0 ? T(10.0) : g);
    const T g2 = gc*gc;
    const T lift = val(gc) < 1.0 ? 1.0 + 0.85*(1.0 - gc) : T(1.0);
    return pt / pk * (g2*g2) * lift;
}

// Bardeen ISCO (isco_bardeen) in the arithmetic type T
template <class T> KT_HD T isco_bardeen_t(double M, const T& a) {
    if (fabs(val(a)) < 1e-10) return T(6.0*M);
    const T ap = a/M;
    const T z1 = 1.0 + cbrt(1.0 - ap*ap) * (cbrt(1.0 + ap) + cbrt(1.0 - ap));
    const T z2 = sqrt(3.0*ap*ap + z1*z1);
    return M*(3.0 + z2 - sqrt((3.0 - z1)*(3.0 + z1 + 2.0*z2)));
}

// d r_isco / d(a, Q, Lambda) of the frame's ISCO.  The scan starts at 1.05 r_+ and moves in
// fixed steps of 1e-4 M, so between its jumps r_isco follows 1.05 dr_+/dp, with dr_+/dp from
// Delta_r(r_+) = 0; with Q = Lambda = 0 the Bardeen closed form is used (its a-derivative).
KT_HD void isco_gradient(const FrameDev& F, double d[3]) {
    if (fabs(F.Lambda) < 1e-15 && fabs(F.Q) < 1e-15) {
        const DN<1> r = isco_bardeen_t<DN<1>>(F.M, DN<1>::var(F.a, 0));
        d[0] = r.d[0]; d[1] = 0.0; d[2] = 0.0;
        return;
    }
    const double r = F.r_h, a = F.a, L = F.Lambda, ra = r*r + a*a;
    const double dDr = 2.0*r*(1.0 - L*r*r/3.0) - ra*(2.0*L*r/3.0) - 2.0*F.M;
    if (!(fabs(dDr) > 0.0)) { d[0] = d[1] = d[2] = 0.0; return; }
    d[0] = -1.05 * 2.0*a*(1.0 - L*r*r/3.0) / dDr;
    d[1] = -1.05 * 2.0*F.Q / dDr;
    d[2] = 1.05 * ra*r*r/3.0 / dDr;
}

// ---------------------------------------------------------------------------
//  local Jacobians
// ---------------------------------------------------------------------------
// J[i][k] = d f_i / d x_k,  x = (r, theta, p_r, p_theta, p_t, p_phi, a, Q, Lambda), i < 4;
// three passes of three tangent directions (register economy)
template <int CHART>
KT_HD void rhs_jacobian(double M, const double par[3], const double y[4], double pt, double pp, double J[4][9]) {
    typedef DN<3> D;
    for (int g = 0; g < 3; ++g) {
        const int o = 3*g;
        const D a = D::var(par[0], 6 - o), Q = D::var(par[1], 7 - o), L = D::var(par[2], 8 - o);
        const D r = D::var(y[0], 0 - o), th = D::var(y[1], 1 - o);
        const D pr = D::var(y[2], 2 - o), pth = D::var(y[3], 3 - o);
        const D ptd = D::var(pt, 4 - o), ppd = D::var(pp, 5 - o);
        const Knds<D> P = knds_t<D>(M, a, Q, L);
        SinCos<D> sc;
        kt_sincos(th, &sc.s, &sc.c);
        D out[5];
        rhs<CHART, D, D>(P, r, sc, pr, pth, ptd, ppd, out);
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 3; ++k) J[i][o + k] = out[i].d[k];
    }
}

// Vector-Jacobian product of the right-hand side (rhs<CHART>, kt_device.hpp), hand-written
// reverse mode: g[k] = sum_i mu[i] d out_i / d x_k over x = (r, theta, p_r, p_theta, p_t,
// p_phi, a, Q, Lambda), the columns of rhs_jacobian, for about three right-hand sides of work
// instead of nine forward directions.  The forward sweep below restates rhs<CHART>'s
// intermediates (1/Sigma, 1/Delta_r and 1/(Delta_theta sin^2) taken as such); the reverse
// sweep visits them in the opposite order.
//   P: the frame's constants (knds_t of a, Q, Lambda); sc: sin, cos of y[1].
//   gsc (optional): the adjoints of sin and cos taken as independent inputs (g[1] is their
//   combination along theta, gsc[0] cos - gsc[1] sin).
template <int CHART, class T = double>
KT_HD void rhs_vjp_sc(const Knds<T>& P, T Q, T Lam, const T y[4], const SinCos<T>& sc, T pt, T pp, const T mu[4],
                      T g[9], T* gsc = nullptr) {
    const T a = P.a;
    const T r = y[0], pr = y[2], pth = y[3];
    const T s = sc.s, c = sc.c;
    // ---- forward ----
    const T s2 = s*s, c2 = c*c, sc2 = s*c, r2 = r*r;
    const T ra = r2 + P.a2;
    const T sig = fma(P.a2, c2, r2);
    const T one_m = fma(-P.lam3, r2, 1.0);
    const T dR = fma(ra, one_m, fma(-P.M2, r, P.Q2));
    const T dR_r = fma(2.0*r, one_m, fma(-2.0*P.lam3*r, ra, -P.M2));
    const T dT = fma(P.lam3a2, c2, 1.0);
    const T dT_th = -2.0*P.lam3a2*sc2;
    const T K = fma(ra, pt, P.a*pp);
    const T L = fma(P.a*s2, pt, pp);
    const T Kr = 2.0*pt*r;
    const T Lth = 2.0*P.a*pt*sc2;
    const T dTs2 = dT*s2;
    const T inv_sig = kt_rcp(sig);
    const T inv_ts = kt_rcp(dTs2);
    const T B1 = P.xi2*L*inv_ts;
    const T W = fma(dT_th, s2, 2.0*dT*sc2);
    const T dNth = fma(dT_th*pth, pth, fma(2.0*B1, Lth, -B1*L*inv_ts*W));
    const T hs = 0.5*inv_sig;
    T N, dNr, T0, inv_dR = 0.0, A1 = 0.0, XK = 0.0;
    if (CHART == CHART_BL) {
        inv_dR = kt_rcp(dR);
        A1 = P.xi2*K*inv_dR;
        N = fma(dR*pr, pr, fma(dT*pth, pth, fma(-A1, K, B1*L)));
        dNr = fma(dR_r*pr, pr, fma(-2.0*A1, Kr, A1*K*inv_dR*dR_r));
        T0 = dR*pr;
    } else {
        XK = P.xi*K;
        N = fma(dR*pr, pr, fma(2.0*XK, pr, fma(dT*pth, pth, B1*L)));
        dNr = fma(dR_r*pr, pr, 2.0*P.xi*pr*Kr);
        T0 = fma(dR, pr, XK);
    }
    const T T2 = 2.0*r*N*inv_sig - dNr;            // out2 = hs T2
    const T T3 = 2.0*P.a2*sc2*N*inv_sig + dNth;    // out3 = -hs T3
    // ---- reverse ----
    T b_r = 0.0, b_s = 0.0, b_c = 0.0, b_pr = 0.0, b_pth = 0.0, b_pt = 0.0, b_pp = 0.0;
    T b_a = 0.0, b_a2 = 0.0, b_Q2 = 0.0, b_lam3 = 0.0, b_lam3a2 = 0.0, b_xi = 0.0, b_xi2 = 0.0;
    T b_inv_sig = 0.0, b_inv_ts = 0.0, b_N = 0.0, b_dR = 0.0, b_dT = 0.0, b_K = 0.0, b_L = 0.0;
    T b_sc2 = 0.0, b_s2 = 0.0, b_dT_th = 0.0, b_Kr = 0.0, b_Lth = 0.0, b_dR_r = 0.0, b_B1 = 0.0;
    // out0 = T0 inv_sig, out1 = dT pth inv_sig, out2 = hs T2, out3 = -hs T3
    const T b_hs = mu[2]*T2 - mu[3]*T3;
    const T b_T2 = mu[2]*hs, b_T3 = -mu[3]*hs;
    const T b_T0 = mu[0]*inv_sig;
    b_inv_sig += mu[0]*T0 + mu[1]*dT*pth + 0.5*b_hs;
    b_dT += mu[1]*pth*inv_sig;
    b_pth += mu[1]*dT*inv_sig;
    // T2 = 2 r N inv_sig - dNr
    b_N += b_T2*2.0*r*inv_sig;
    b_r += b_T2*2.0*N*inv_sig;
    b_inv_sig += b_T2*2.0*r*N;
    const T b_dNr = -b_T2;
    // T3 = 2 a2 sc2 N inv_sig + dNth
    b_a2 += b_T3*2.0*sc2*N*inv_sig;
    b_sc2 += b_T3*2.0*P.a2*N*inv_sig;
    b_N += b_T3*2.0*P.a2*sc2*inv_sig;
    b_inv_sig += b_T3*2.0*P.a2*sc2*N;
    const T b_dNth = b_T3;
    // dNth = dT_th pth^2 + 2 B1 Lth - B1 L inv_ts W
    b_dT_th += b_dNth*pth*pth;
    b_pth += b_dNth*2.0*dT_th*pth;
    b_B1 += b_dNth*(2.0*Lth - L*inv_ts*W);
    b_Lth += b_dNth*2.0*B1;
    b_L -= b_dNth*B1*inv_ts*W;
    b_inv_ts -= b_dNth*B1*L*W;
    const T b_W = -b_dNth*B1*L*inv_ts;
    // W = dT_th s2 + 2 dT sc2
    b_dT_th += b_W*s2;
    b_s2 += b_W*dT_th;
    b_dT += b_W*2.0*sc2;
    b_sc2 += b_W*2.0*dT;
    if (CHART == CHART_BL) {
        T b_A1 = 0.0, b_inv_dR = 0.0;
        // T0 = dR pr
        b_dR += b_T0*pr;
        b_pr += b_T0*dR;
        // dNr = dR_r pr^2 - 2 A1 Kr + A1 K inv_dR dR_r
        b_dR_r += b_dNr*(pr*pr + A1*K*inv_dR);
        b_pr += b_dNr*2.0*dR_r*pr;
        b_A1 += b_dNr*(K*inv_dR*dR_r - 2.0*Kr);
        b_Kr -= b_dNr*2.0*A1;
        b_K += b_dNr*A1*inv_dR*dR_r;
        b_inv_dR += b_dNr*A1*K*dR_r;
        // N = dR pr^2 + dT pth^2 - A1 K + B1 L
        b_dR += b_N*pr*pr;
        b_pr += b_N*2.0*dR*pr;
        b_dT += b_N*pth*pth;
        b_pth += b_N*2.0*dT*pth;
        b_A1 -= b_N*K;
        b_K -= b_N*A1;
        b_B1 += b_N*L;
        b_L += b_N*B1;
        // A1 = xi2 K inv_dR, inv_dR = 1/dR
        b_xi2 += b_A1*K*inv_dR;
        b_K += b_A1*P.xi2*inv_dR;
        b_inv_dR += b_A1*P.xi2*K;
        b_dR -= b_inv_dR*inv_dR*inv_dR;
    } else {
        // T0 = dR pr + XK
        b_dR += b_T0*pr;
        b_pr += b_T0*dR;
        T b_XK = b_T0;
        // dNr = dR_r pr^2 + 2 xi pr Kr
        b_dR_r += b_dNr*pr*pr;
        b_pr += b_dNr*(2.0*dR_r*pr + 2.0*P.xi*Kr);
        b_xi += b_dNr*2.0*pr*Kr;
        b_Kr += b_dNr*2.0*P.xi*pr;
        // N = dR pr^2 + 2 XK pr + dT pth^2 + B1 L
        b_dR += b_N*pr*pr;
        b_pr += b_N*(2.0*dR*pr + 2.0*XK);
        b_XK += b_N*2.0*pr;
        b_dT += b_N*pth*pth;
        b_pth += b_N*2.0*dT*pth;
        b_B1 += b_N*L;
        b_L += b_N*B1;
        // XK = xi K
        b_xi += b_XK*K;
        b_K += b_XK*P.xi;
    }
    // B1 = xi2 L inv_ts
    b_xi2 += b_B1*L*inv_ts;
    b_L += b_B1*P.xi2*inv_ts;
    b_inv_ts += b_B1*P.xi2*L;
    // inv_ts = 1/dTs2, inv_sig = 1/sig, dTs2 = dT s2
    const T b_dTs2 = -b_inv_ts*inv_ts*inv_ts;
    const T b_sig = -b_inv_sig*inv_sig*inv_sig;
    b_dT += b_dTs2*s2;
    b_s2 += b_dTs2*dT;
    // Lth = 2 a pt sc2, Kr = 2 pt r
    b_a += b_Lth*2.0*pt*sc2;
    b_pt += b_Lth*2.0*P.a*sc2 + b_Kr*2.0*r;
    b_sc2 += b_Lth*2.0*P.a*pt;
    b_r += b_Kr*2.0*pt;
    // L = a s2 pt + pp, K = ra pt + a pp
    b_a += b_L*s2*pt + b_K*pp;
    b_s2 += b_L*P.a*pt;
    b_pt += b_L*P.a*s2 + b_K*ra;
    b_pp += b_L + b_K*P.a;
    T b_ra = b_K*pt;
    // dT_th = -2 lam3a2 sc2, dT = 1 + lam3a2 c2
    b_lam3a2 += -2.0*b_dT_th*sc2 + b_dT*c2;
    b_sc2 -= 2.0*b_dT_th*P.lam3a2;
    T b_c2 = b_dT*P.lam3a2;
    // dR_r = 2 r one_m - 2 lam3 r ra - M2, dR = ra one_m - M2 r + Q2
    b_r += b_dR_r*(2.0*one_m - 2.0*P.lam3*ra) - b_dR*P.M2;
    T b_one_m = b_dR_r*2.0*r + b_dR*ra;
    b_lam3 -= b_dR_r*2.0*r*ra;
    b_ra += -b_dR_r*2.0*P.lam3*r + b_dR*one_m;
    b_Q2 += b_dR;
    // one_m = 1 - lam3 r2, sig = a2 c2 + r2, ra = r2 + a2
    b_lam3 -= b_one_m*r2;
    T b_r2 = -b_one_m*P.lam3 + b_sig + b_ra;
    b_a2 += b_sig*c2 + b_ra;
    b_c2 += b_sig*P.a2;
    // r2 = r r, sc2 = s c, c2 = c c, s2 = s s
    b_r += 2.0*r*b_r2;
    b_s += b_sc2*c + 2.0*s*b_s2;
    b_c += b_sc2*s + 2.0*c*b_c2;
    if (gsc) { gsc[0] = b_s; gsc[1] = b_c; }
    // (s, c) = (sin, cos) theta
    const T b_th = b_s*c - b_c*s;
    // constants: a2 = a^2, Q2 = Q^2, lam3 = L/3, lam3a2 = L a^2/3, xi = 1 + L a^2/3, xi2 = xi^2
    b_xi += 2.0*P.xi*b_xi2;
    g[0] = b_r; g[1] = b_th; g[2] = b_pr; g[3] = b_pth; g[4] = b_pt; g[5] = b_pp;
    const T third = 1.0/3.0;
    g[6] = b_a + 2.0*a*b_a2 + (b_lam3a2 + b_xi)*(2.0*third*Lam*a);
    g[7] = 2.0*Q*b_Q2;
    g[8] = (b_lam3 + (b_lam3a2 + b_xi)*(a*a))*third;
}

// rhs_vjp_sc from (M, a, Q, Lambda) and theta itself
template <int CHART>
KT_HD void rhs_vjp(double M, const double par[3], const double y[4], double pt, double pp, const double mu[4],
                   double g[9]) {
    const Knds<double> P = knds_t<double>(M, par[0], par[1], par[2]);
    SinCos<double> sc;
    kt_sincos(y[1], &sc.s, &sc.c);
    rhs_vjp_sc<CHART>(P, par[1], par[2], y, sc, pt, pp, mu, g);
}

// initial state y0 = (r, theta, p_r, p_theta), c = (p_t, p_phi) of a pixel and their
// derivatives with respect to p (theta_obs in degrees)
template <int CHART>
KT_HD void init_jacobian(const FrameDev& F, double al, double be, double y0[4], double c[2], double Jy[4][AP_N],
                         double Jc[2][AP_N]) {
    typedef DN<AP_N> D;
    const D a = D::var(F.a, AP_A), Q = D::var(F.Q, AP_Q), L = D::var(F.Lambda, AP_L);
    const D th = D::var(F.th_obs * 180.0 / kPi, AP_TH) * kPi / 180.0;
    const D r = D::var(F.r_obs, AP_R);
    D thv = th;
    thv.v = F.th_obs;                       // the frame's own rounding of the angle
    ObsT<D> C;
    observer_t<D>(F.M, a, Q, L, r, thv, C);
    D st[7];
    camera_t<D>(C, r, thv, al, be, st);
    if (CHART == CHART_KS) bl_to_ks_t<D>(knds_t<D>(F.M, a, Q, L), F.ks_ref_init != 0, st);
    const D* ys[4] = {&st[0], &st[1], &st[3], &st[4]};
    for (int i = 0; i < 4; ++i) {
        y0[i] = ys[i]->v;
        for (int k = 0; k < AP_N; ++k) Jy[i][k] = ys[i]->d[k];
    }
    for (int k = 0; k < AP_N; ++k) { Jc[0][k] = st[5].d[k]; Jc[1][k] = st[6].d[k]; }
    c[0] = st[5].v;
    c[1] = st[6].v;
}

// ---------------------------------------------------------------------------
//  replay of an accepted DOPRI5 step: the forward pass's own attempt (dp5_attempt_h,
//  binary64, phi not carried), so the replay retraces it bit for bit: y, sc at the step
//  start and the FSAL derivative k1 = f(y) -> the step end, its sin/cos and its derivative
// ---------------------------------------------------------------------------
template <int CHART>
KT_HD void dp5_replay(const Knds<double>& P, double y[4], SinCos<double>& sc, double k1[5], double h, double pt,
                      double pp) {
    double z[4], ae[4], k7[5], dth5, dph5;
    SinCos<double> sc7;
    dp5_attempt_h<CHART, PREC_FP64, 4>(P, knds_cast<float>(P), true, y, k1, sc, h, pt, pp, z, dth5, dph5, ae, k7,
                                       sc7);
    y[0] = z[0]; y[1] = y[1] + dth5; y[2] = z[2]; y[3] = z[3];
    for (int i = 0; i < 5; ++i) k1[i] = k7[i];
    sc = sc7;
}

// Adjoint of one accepted step of dp5_attempt_h: lam (adjoint of the step end) -> adjoint of
// the step start; accumulates the parameter adjoints gc (p_t, p_phi) and gP (a, Q, Lambda).
// The step, with F_i = h f(z_i, s_i, c_i) the scaled stages:
//   z_i = y + sum_j a_ij F_j          (r, p_r, p_theta)
//   s_i = s + sum_j a_ij u_j,  c_i = c + sum_j a_ij v_j,   u_j = c_j F_j[theta], v_j = -s_j F_j[theta]
//   y+  = y + sum_i b_i F_i           (r, theta, p_r, p_theta)
// with (s, c) = (sin, cos) of theta at the step start (the forward pass's FSAL rotation makes
// them so).  Reverse, i = 6 .. 1, with bars for adjoints:
//   ubar_i = sum_m a_mi sbar_m,  vbar_i = sum_m a_mi cbar_m
//   Fbar_i = b_i lam + sum_m a_mi zbar_m   (r, p_r, p_theta),
//   Fbar_i[theta] = b_i lam[theta] + ubar_i c_i - vbar_i s_i
//   (zbar_i, sbar_i, cbar_i, p_t, p_phi, a, Q, Lambda) += h J_f^T Fbar_i   (rhs_vjp_sc)
//   sbar_i -= vbar_i F_i[theta],  cbar_i += ubar_i F_i[theta]
// then lam[theta] <- lam[theta] + cos theta sum_i sbar_i - sin theta sum_i cbar_i: the exact
// adjoint of the integrated stage sin/cos, not of a rotation per stage.
template <int CHART>
KT_HD void dp5_step_adjoint(double M, const double par[3], const Knds<double>& P, const double y[4],
                            const SinCos<double>& sc, double h, double pt, double pp, double lam[4], double gc[2],
                            double gP[3]) {
    typedef DP5 C;
    const double A[6][6] = {{0, 0, 0, 0, 0, 0},
                            {C::a21, 0, 0, 0, 0, 0},
                            {C::a31, C::a32, 0, 0, 0, 0},
                            {C::a41, C::a42, C::a43, 0, 0, 0},
                            {C::a51, C::a52, C::a53, C::a54, 0, 0},
                            {C::a61, C::a62, C::a63, C::a64, C::a65, 0}};
    const double B[6] = {C::b1, 0.0, C::b3, C::b4, C::b5, C::b6};
    // stage points (r, p_r, p_theta: theta enters f only through sin/cos), sin/cos and
    // scaled stages (forward arithmetic of the attempt)
    const int ix[3] = {0, 2, 3};
    double Z[6][3], Fs[6][4];
    SinCos<double> scs[6];
    for (int i = 0; i < 3; ++i) Z[0][i] = y[ix[i]];
    scs[0] = sc;
    {
        double k[5];
        rhs<CHART>(P, y[0], sc, y[2], y[3], pt, pp, k);
        for (int i = 0; i < 4; ++i) Fs[0][i] = h*k[i];
    }
    for (int st = 1; st < 6; ++st) {
        double zs = sc.s, zc = sc.c;
        for (int i = 0; i < 3; ++i) Z[st][i] = y[ix[i]];
        for (int j = 0; j < st; ++j) {
            for (int i = 0; i < 3; ++i) Z[st][i] = fma(A[st][j], Fs[j][ix[i]], Z[st][i]);
            zs = fma(A[st][j], scs[j].c*Fs[j][1], zs);
            zc = fma(-A[st][j], scs[j].s*Fs[j][1], zc);
        }
        scs[st].s = zs;
        scs[st].c = zc;
        double k[5];
        rhs_h<CHART>(P, Z[st][0], scs[st], Z[st][1], Z[st][2], pt, pp, h, k);
        for (int i = 0; i < 4; ++i) Fs[st][i] = k[i];
    }
    // reverse over the stages
    double zb[6][3], sb[6], cb[6];
    double lam_out[4] = {lam[0], lam[1], lam[2], lam[3]};
    double sb_sum = 0.0, cb_sum = 0.0;
    for (int st = 5; st >= 0; --st) {
        double ub = 0.0, vb = 0.0, Fb[4];
        for (int m = st + 1; m < 6; ++m) { ub = fma(A[m][st], sb[m], ub); vb = fma(A[m][st], cb[m], vb); }
        for (int i = 0; i < 4; ++i) Fb[i] = B[st]*lam[i];
        for (int i = 0; i < 3; ++i)
            for (int m = st + 1; m < 6; ++m) Fb[ix[i]] = fma(A[m][st], zb[m][i], Fb[ix[i]]);
        Fb[1] = fma(ub, scs[st].c, fma(-vb, scs[st].s, Fb[1]));
        double mu[4], gv[9], gsc[2];
        for (int i = 0; i < 4; ++i) mu[i] = h*Fb[i];
        const double zs[4] = {Z[st][0], 0.0, Z[st][1], Z[st][2]};    // theta is not an input of f
        rhs_vjp_sc<CHART>(P, par[1], par[2], zs, scs[st], pt, pp, mu, gv, gsc);
        for (int i = 0; i < 3; ++i) zb[st][i] = gv[ix[i]];
        sb[st] = fma(-vb, Fs[st][1], gsc[0]);
        cb[st] = fma(ub, Fs[st][1], gsc[1]);
        sb_sum += sb[st];
        cb_sum += cb[st];
        lam_out[0] += gv[0];
        lam_out[2] += gv[2];
        lam_out[3] += gv[3];
        gc[0] += gv[4];
        gc[1] += gv[5];
        gP[0] += gv[6];
        gP[1] += gv[7];
        gP[2] += gv[8];
    }
    lam_out[1] = fma(sb_sum, sc.c, fma(-cb_sum, sc.s, lam_out[1]));
    for (int i = 0; i < 4; ++i) lam[i] = lam_out[i];
}

// Hermite basis on [0,1] and its alpha-derivative (the interpolant of hermite())
KT_HD void hermite_basis(double al, double hb[4], double db[4]) {
    const double a2 = al*al, a3 = a2*al;
    hb[0] = 2.0*a3 - 3.0*a2 + 1.0; hb[1] = a3 - 2.0*a2 + al; hb[2] = -2.0*a3 + 3.0*a2; hb[3] = a3 - a2;
    db[0] = 6.0*a2 - 6.0*al; db[1] = 3.0*a2 - 4.0*al + 1.0; db[2] = -6.0*a2 + 6.0*al; db[3] = 3.0*a2 - 2.0*al;
}

// Seed of the disk event on the last step [y0 -> y1] of size h: the intensity and its
// adjoints on y0, y1 (through the interpolation and the end-point derivatives),
// on (p_t, p_phi) and on (a, Q, Lambda).  Returns false when the step has no disk root.
template <int CHART>
KT_HD bool event_seed(const FrameDev& F, const double y0[4], const double y1[4], double h, double pt, double pp,
                      double lam0[4], double lam1[4], double gc[2], double gP[3], double* inten) {
    const double par[3] = {F.a, F.Q, F.Lambda};
    const Knds<double>& P = F.P;
    // end-point derivatives (values) and Jacobians
    double f0[5], f1[5];
    SinCos<double> s0, s1;
    kt_sincos(y0[1], &s0.s, &s0.c);
    kt_sincos(y1[1], &s1.s, &s1.c);
    rhs<CHART>(P, y0[0], s0, y0[2], y0[3], pt, pp, f0);
    rhs<CHART>(P, y1[0], s1, y1[2], y1[3], pt, pp, f1);
    const double halfpi = kPi / 2.0;
    double al = 0.0;
    if (!hermite_first_root(y0[1], y1[1], f0[1], f1[1], h, halfpi, al)) return false;
    al = hermite_root_polish(y0[1], y1[1], f0[1], f1[1], h, halfpi, al);
    double hb[4], db[4];
    hermite_basis(al, hb, db);
    const double rh = hb[0]*y0[0] + hb[1]*(h*f0[0]) + hb[2]*y1[0] + hb[3]*(h*f1[0]);
    // redshift and intensity in dual numbers over (r_hit, p_t, p_phi, a, Q, Lambda, r_isco)
    typedef DN<7> D;
    const D r = D::var(rh, 0), ptd = D::var(pt, 1), ppd = D::var(pp, 2);
    const D a = D::var(F.a, 3), Q = D::var(F.Q, 4), L = D::var(F.Lambda, 5);
    const D g = disk_redshift_t<D>(F.M, a, Q, L, r, ptd, ppd);
    const D I = disk_intensity_t<D>(r, g, D::var(F.r_isco, 6));
    double dri[3];
    isco_gradient(F, dri);
    *inten = I.v;
    gc[0] += I.d[1];
    gc[1] += I.d[2];
    gP[0] += I.d[3] + I.d[6]*dri[0];
    gP[1] += I.d[4] + I.d[6]*dri[1];
    gP[2] += I.d[5] + I.d[6]*dri[2];
    const double lr = I.d[0];                       // dI / dr_hit
    // r_hit(alpha; r0, r1, f0r, f1r), alpha from theta(alpha) = pi/2
    const double th_al = db[0]*y0[1] + db[1]*(h*f0[1]) + db[2]*y1[1] + db[3]*(h*f1[1]);
    const double r_al = db[0]*y0[0] + db[1]*(h*f0[0]) + db[2]*y1[0] + db[3]*(h*f1[0]);
    const double la = (fabs(th_al) > 1e-300) ? -lr*r_al/th_al : 0.0;   // d alpha = -(dTheta)/Theta_alpha
    // adjoints of the interpolation inputs
    double e0[4] = {lr*hb[0], la*hb[0], 0.0, 0.0};          // direct on y0 (r, theta)
    double e1[4] = {lr*hb[2], la*hb[2], 0.0, 0.0};          // direct on y1
    const double df0[2] = {lr*hb[1]*h, la*hb[1]*h};          // on f0 (r, theta components)
    const double df1[2] = {lr*hb[3]*h, la*hb[3]*h};
    // through the end-point derivatives f0, f1 (their r and theta rows): J^T (df, 0, 0)
    const double m0[4] = {df0[0], df0[1], 0.0, 0.0}, m1[4] = {df1[0], df1[1], 0.0, 0.0};
    double g0[9], g1[9];
    rhs_vjp<CHART>(F.M, par, y0, pt, pp, m0, g0);
    rhs_vjp<CHART>(F.M, par, y1, pt, pp, m1, g1);
    for (int k = 0; k < 4; ++k) {
        lam0[k] = e0[k] + g0[k];
        lam1[k] = e1[k] + g1[k];
    }
    for (int k = 0; k < 2; ++k) gc[k] += g0[4 + k] + g1[4 + k];
    for (int k = 0; k < 3; ++k) gP[k] += g0[6 + k] + g1[6 + k];
    return true;
}

// Reverse pass of one pixel from its records: N >= 1 accepted steps of sizes hs[0..N-1]
// and checkpoints ck (6 doubles each, the state before steps 0, 16, 32, ...).  The ray
// must have ended on the disk in step N-1.  grad = dI/dp; false if the replayed last
// step has no disk root.
template <int CHART>
KT_HD bool adjoint_pixel(const FrameDev& F, int px, int py, int N, const double* hs, const double* ck,
                         double grad[AP_N]) {
    const double par[3] = {F.a, F.Q, F.Lambda};
    const Knds<double>& P = F.P;
    double al, be;
    pixel_angles(F, (double)px, (double)py, al, be);
    double y0[4], cc[2], Jy[4][AP_N], Jc[2][AP_N];
    init_jacobian<CHART>(F, al, be, y0, cc, Jy, Jc);
    const double pt = cc[0], pp = cc[1];
    double lam[4] = {0.0, 0.0, 0.0, 0.0}, gc[2] = {0.0, 0.0}, gP[3] = {0.0, 0.0, 0.0};
    double segy[kCkptEvery + 1][4];
    SinCos<double> segsc[kCkptEvery + 1];
    for (int k = (N - 1) / kCkptEvery; k >= 0; --k) {
        const int n0 = k*kCkptEvery;
        const int n1 = (n0 + kCkptEvery < N) ? n0 + kCkptEvery : N;
        double y[4] = {ck[6*k + 0], ck[6*k + 1], ck[6*k + 2], ck[6*k + 3]};
        SinCos<double> sc;
        sc.s = ck[6*k + 4];
        sc.c = ck[6*k + 5];
        double k1[5];
        rhs<CHART>(P, y[0], sc, y[2], y[3], pt, pp, k1);
        for (int i = 0; i < 4; ++i) segy[0][i] = y[i];
        segsc[0] = sc;
        for (int n = n0; n < n1; ++n) {
            dp5_replay<CHART>(P, y, sc, k1, hs[n], pt, pp);
            for (int i = 0; i < 4; ++i) segy[n - n0 + 1][i] = y[i];
            segsc[n - n0 + 1] = sc;
        }
        int n = n1 - 1;
        if (n1 == N) {                      // last segment: seed at the disk event of step N-1
            double lam0[4], lam1[4], inten = 0.0;
            if (!event_seed<CHART>(F, segy[n - n0], segy[n - n0 + 1], hs[n], pt, pp, lam0, lam1, gc, gP, &inten))
                return false;
            for (int i = 0; i < 4; ++i) lam[i] = lam1[i];
            dp5_step_adjoint<CHART>(F.M, par, P, segy[n - n0], segsc[n - n0], hs[n], pt, pp, lam, gc, gP);
            for (int i = 0; i < 4; ++i) lam[i] += lam0[i];
            --n;
        }
        for (; n >= n0; --n)
            dp5_step_adjoint<CHART>(F.M, par, P, segy[n - n0], segsc[n - n0], hs[n], pt, pp, lam, gc, gP);
    }
    for (int q = 0; q < AP_N; ++q) {
        double g = gc[0]*Jc[0][q] + gc[1]*Jc[1][q];
        for (int i = 0; i < 4; ++i) g += lam[i]*Jy[i][q];
        if (q < 3) g += gP[q];
        grad[q] = g;
    }
    return true;
}

// ---------------------------------------------------------------------------
//  Reverse pass in parallel over checkpoint segments (the GPU's reverse pass).
//
//  The adjoint of a segment is linear in the adjoint at its end, so every segment's map is
//  formed on its own, once per basis direction of that end adjoint: four independent sweeps
//  (one lane each) give the transposed 4x4 state map and the 4x5 parameter rows of the
//  segment.  The last step of a ray, which holds the disk event, is the pixel's "event entry":
//  the seed propagated through that one step.  A pixel's gradient is then the event entry
//  carried back through its segments' small matrices, last to first (adjoint_compose): the
//  same quantity as adjoint_pixel's single sweep, summed in another order.
// ---------------------------------------------------------------------------
constexpr int kSegOut = 9;     // per sweep: adjoint at the segment start (4), p_t, p_phi (2), a, Q, Lambda (3)

// checkpoint segments of a ray of N >= 1 accepted steps: steps [16k, min(16k + 16, N - 1))
KT_HD int adjoint_segments(int N) { return (N - 1 + kCkptEvery - 1) / kCkptEvery; }

// Cost classes of the reverse pass's entries, so that the entries of one class run side by side:
// class 0 the full segments (16 steps swept), classes 1..16 a ray's last segment of 16 .. 1 steps,
// classes 17..32 the event entry, whose replay from the last checkpoint takes 16 .. 1 steps.
constexpr int kAdjClasses = 2 * kCkptEvery + 1;
KT_HD int adjoint_last_class(int N) {                // the last segment's class (N >= 2)
    const int m = (N - 1) - kCkptEvery * (adjoint_segments(N) - 1);
    return 1 + kCkptEvery - m;
}
KT_HD int adjoint_event_class(int N) {
    const int r = N - kCkptEvery * ((N - 1) / kCkptEvery);
    return 1 + 2 * kCkptEvery - r;
}

// conserved momenta of the ray at pixel (px, py), by the forward pass's own camera path
template <int CHART>
KT_HD void ray_momenta(const FrameDev& F, int px, int py, double& pt, double& pp) {
    double st[7];
    camera_ray(F, (double)px, (double)py, st);
    if (CHART == CHART_KS) bl_to_ks_covector(F, st);
    pt = st[5];
    pp = st[6];
}

// The replayed states of one segment.  LocalStates keeps them in the lane (the host build);
// the GPU's segment kernel passes a store in LDS that the four sweeps of a segment share (its
// first lane writes, all four read after sync()), which keeps them out of scratch.
struct LocalStates {
    double y[kCkptEvery + 1][4];
    SinCos<double> sc[kCkptEvery + 1];
    KT_HD void put(int i, const double v[4], const SinCos<double>& s) {
        for (int k = 0; k < 4; ++k) y[i][k] = v[k];
        sc[i] = s;
    }
    KT_HD void get(int i, double v[4], SinCos<double>& s) const {
        for (int k = 0; k < 4; ++k) v[k] = y[i][k];
        s = sc[i];
    }
    KT_HD void sync() {}
};

// replay of the accepted steps n0 .. n1 - 1 from the checkpoint before step n0 (ck: that
// checkpoint's 6 values); state i of the store = the state before step n0 + i, i = 0 .. n1 - n0
template <int CHART, class Store>
KT_HD void replay_steps(const Knds<double>& P, const double* hs, const double* ck, int n0, int n1, double pt,
                        double pp, Store& st) {
    double y[4] = {ck[0], ck[1], ck[2], ck[3]};
    SinCos<double> sc;
    sc.s = ck[4];
    sc.c = ck[5];
    double k1[5];
    rhs<CHART>(P, y[0], sc, y[2], y[3], pt, pp, k1);
    st.put(0, y, sc);
    for (int n = n0; n < n1; ++n) {
        dp5_replay<CHART>(P, y, sc, k1, hs[n], pt, pp);
        st.put(n - n0 + 1, y, sc);
    }
    st.sync();
}

// Segment k (steps [16k, min(16k + 16, N - 1))) of a ray of N accepted steps, swept back from
// the unit adjoint e_d at its end.  ck: the ray's checkpoints (6 per segment).
template <int CHART, class Store>
KT_HD void adjoint_segment(const FrameDev& F, double pt, double pp, int N, const double* hs, const double* ck,
                           int k, int d, Store& st, double out[kSegOut]) {
    const double par[3] = {F.a, F.Q, F.Lambda};
    const Knds<double>& P = F.P;
    const int n0 = k*kCkptEvery;
    const int n1 = (n0 + kCkptEvery < N - 1) ? n0 + kCkptEvery : N - 1;
    replay_steps<CHART>(P, hs, ck + 6*k, n0, n1, pt, pp, st);
    double lam[4] = {0.0, 0.0, 0.0, 0.0}, gc[2] = {0.0, 0.0}, gP[3] = {0.0, 0.0, 0.0};
    lam[d] = 1.0;
    for (int n = n1 - 1; n >= n0; --n) {
        double y[4];
        SinCos<double> sc;
        st.get(n - n0, y, sc);
        dp5_step_adjoint<CHART>(F.M, par, P, y, sc, hs[n], pt, pp, lam, gc, gP);
    }
    for (int i = 0; i < 4; ++i) out[i] = lam[i];
    out[4] = gc[0]; out[5] = gc[1];
    out[6] = gP[0]; out[7] = gP[1]; out[8] = gP[2];
}

// The event entry of a ray of N >= 1 accepted steps: the disk seed on step N - 1 carried back
// through that step.  out: the adjoint at the start of step N - 1 and the parameter sums.
// Returns false when the replayed last step has no disk root.
template <int CHART, class Store>
KT_HD bool adjoint_event(const FrameDev& F, double pt, double pp, int N, const double* hs, const double* ck,
                         Store& st, double out[kSegOut]) {
    const double par[3] = {F.a, F.Q, F.Lambda};
    const Knds<double>& P = F.P;
    const int k = (N - 1) / kCkptEvery, n0 = k*kCkptEvery;
    replay_steps<CHART>(P, hs, ck + 6*k, n0, N, pt, pp, st);
    const int n = N - 1;
    double y0[4], y1[4];
    SinCos<double> sc0, sc1;
    st.get(n - n0, y0, sc0);
    st.get(n - n0 + 1, y1, sc1);
    double lam0[4], lam1[4], gc[2] = {0.0, 0.0}, gP[3] = {0.0, 0.0, 0.0}, inten = 0.0;
    if (!event_seed<CHART>(F, y0, y1, hs[n], pt, pp, lam0, lam1, gc, gP, &inten)) return false;
    dp5_step_adjoint<CHART>(F.M, par, P, y0, sc0, hs[n], pt, pp, lam1, gc, gP);
    for (int i = 0; i < 4; ++i) out[i] = lam1[i] + lam0[i];
    out[4] = gc[0]; out[5] = gc[1];
    out[6] = gP[0]; out[7] = gP[1]; out[8] = gP[2];
    return true;
}

// A pixel's gradient from its event entry ev and its nseg segments' sweeps (segment k < nseg - 1,
// direction d at seg + (4k + d) kSegOut; the last segment's at last + d kSegOut): the end adjoint
// carried back segment by segment, then through the camera map.
template <int CHART>
KT_HD void adjoint_compose(const FrameDev& F, int px, int py, int nseg, const double* ev, const double* seg,
                           const double* last, double grad[AP_N]) {
    double al, be;
    pixel_angles(F, (double)px, (double)py, al, be);
    double y0[4], cc[2], Jy[4][AP_N], Jc[2][AP_N];
    init_jacobian<CHART>(F, al, be, y0, cc, Jy, Jc);
    double lam[4] = {ev[0], ev[1], ev[2], ev[3]}, gc[2] = {ev[4], ev[5]}, gP[3] = {ev[6], ev[7], ev[8]};
    for (int k = nseg - 1; k >= 0; --k) {
        const double* S = (k == nseg - 1) ? last : seg + 4*kSegOut*k;
        double nl[4] = {0.0, 0.0, 0.0, 0.0};
        for (int d = 0; d < 4; ++d) {
            const double* o = S + kSegOut*d;
            for (int i = 0; i < 4; ++i) nl[i] = fma(lam[d], o[i], nl[i]);
            gc[0] = fma(lam[d], o[4], gc[0]);
            gc[1] = fma(lam[d], o[5], gc[1]);
            for (int q = 0; q < 3; ++q) gP[q] = fma(lam[d], o[6 + q], gP[q]);
        }
        for (int i = 0; i < 4; ++i) lam[i] = nl[i];
    }
    for (int q = 0; q < AP_N; ++q) {
        double g = gc[0]*Jc[0][q] + gc[1]*Jc[1][q];
        for (int i = 0; i < 4; ++i) g += lam[i]*Jy[i][q];
        if (q < 3) g += gP[q];
        grad[q] = g;
    }
}

}  // namespace kt
