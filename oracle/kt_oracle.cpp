// ============================================================================
//  kt_oracle.cpp -- CPU FP64 oracle for the KNdS geodesic hot path.
//
//  TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
//  cpu_baseline leg of bench.py may load this library, and only as the
//  checker or as the timed CPU baseline.  The product (kerrtracecpp2_amd)
//  never links, loads or calls it.
//
//  This is a restatement of the reference CPU path in plain IEEE binary64
//  (built with g++ -O2, no -ffast-math, no FMA contraction), written so that
//  every floating-point expression is evaluated in the same order as the
//  reference; the per-pixel GeoPixel output is then bit-identical to the
//  reference binary built from source (checked by tests/test_oracle_golden.py
//  against fixtures made from that binary).
//
//  Reference anchors (file:line in the KerrTraceCpp2 tree):
//    metric ............ knds_metric.hpp:45-132, 150-230, 237-411
//    integrators ....... geodesic.hpp:40-437
//    camera ............ camera.hpp:32-145
//    ray bundles ....... ray_bundle.hpp:56-436
//    BL tracer ......... main.cpp:329-434
//    KS tracer ......... main.cpp:1303-1660
//    shading ........... main.cpp:140-171, 1663-1687, 1799-1866
//    Interstellar ...... main.cpp:1689-1797
//    frame driver ...... main.cpp:2014-2206 (GeoPixel fill 2098-2147)
//    camera spp ........ main.cpp:2232-2322
//
//  Parity status per function is listed in the header comment of
//  tests/test_oracle_golden.py and tests/test_oracle_invariants.py.
// ============================================================================
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace kto {

static constexpr double PI = 3.14159265358979323846;  // == M_PI

template <class T> static inline T clampv(T v, T lo, T hi) { return v < lo ? lo : v > hi ? hi : v; }

// ---------------------------------------------------------------------------
//  KNdS metric (knds_metric.hpp:28-412)
// ---------------------------------------------------------------------------
struct Spacetime {
    double M, a, Q, Lambda;

    double sigma(double r, double th) const {            // knds_metric.hpp:45-48
        double ct = std::cos(th);
        return r*r + a*a*ct*ct;
    }
    double delta_r(double r) const {                      // :51-53
        return (r*r + a*a) * (1.0 - Lambda*r*r/3.0) - 2.0*M*r + Q*Q;
    }
    double delta_th(double th) const {                    // :56-59
        double ct = std::cos(th);
        return 1.0 + Lambda*a*a*ct*ct/3.0;
    }
    double xi() const { return 1.0 + Lambda*a*a/3.0; }    // :62

    // g_{mu nu} in BL, components (tt, tphi, rr, thth, phiphi)   :71-88
    void lower_bl(double r, double th, double g[4][4]) const {
        const double sg = sigma(r, th), dr = delta_r(r), dt = delta_th(th);
        const double x2 = xi()*xi();
        const double s = std::sin(th), s2 = s*s;
        const double ra = r*r + a*a;
        const double den = sg*x2;
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
        g[0][0] = (-dr + dt*a*a*s2) / den;
        g[0][3] = g[3][0] = a*s2*(dr - dt*ra) / den;
        g[1][1] = sg/dr;
        g[2][2] = sg/dt;
        g[3][3] = s2*(dt*ra*ra - dr*a*a*s2) / den;
    }

    // g^{mu nu} in BL   :99-117
    void upper_bl(double r, double th, double g[4][4]) const {
        const double sg = sigma(r, th), dr = delta_r(r), dt = delta_th(th);
        const double x2 = xi()*xi();
        const double s = std::sin(th), s2 = s*s;
        const double ra = r*r + a*a;
        const double den = sg*dr*dt;
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
        g[0][0] = -x2*(dt*ra*ra - dr*a*a*s2) / den;
        g[0][3] = g[3][0] = a*x2*(dr - dt*ra) / den;
        g[1][1] = dr/sg;
        g[2][2] = dt/sg;
        if (s2 > 1e-14) g[3][3] = x2*(dr - dt*a*a*s2) / (den*s2);
    }

    // H = 1/2 g^{mu nu} p_mu p_nu   :122-132
    double ham(double r, double th, double pr, double pth, double pt, double pp) const {
        double g[4][4];
        upper_bl(r, th, g);
        return 0.5*(  g[0][0]*pt*pt
                    + 2.0*g[0][3]*pt*pp
                    + g[1][1]*pr*pr
                    + g[2][2]*pth*pth
                    + g[3][3]*pp*pp);
    }

    // ---- Kerr-Schild Cartesian, Lambda = 0 (:150-230) ----
    static double ks_radius(double X, double Y, double Z, double as) {      // :150-155
        const double R2 = X*X + Y*Y + Z*Z;
        const double a2 = as*as;
        const double b = R2 - a2;
        return std::sqrt(0.5*(b + std::sqrt(b*b + 4.0*a2*Z*Z)));
    }
    double ks_scalar(double r, double rho2) const { return (2.0*M*r - Q*Q) / rho2; }   // :158-160
    static void ks_null(double X, double Y, double Z, double r, double as, int sgn, double l[4]) {  // :168-178
        const double ra = r*r + as*as;
        const double sa = double(sgn) * as;
        l[0] = double(sgn);
        l[1] = (r*X + sa*Y) / ra;
        l[2] = (r*Y - sa*X) / ra;
        l[3] = Z / r;
    }
    void upper_ks(double X, double Y, double Z, double g[4][4]) const {     // :199-218 (ingoing)
        const double r = ks_radius(X, Y, Z, a);
        const double rho2 = r*r + a*a*Z*Z/(r*r);
        const double Hs = ks_scalar(r, rho2);
        double l[4];
        ks_null(X, Y, Z, r, a, +1, l);
        const double lu[4] = {-l[0], l[1], l[2], l[3]};
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
        g[0][0] = -1.0; g[1][1] = 1.0; g[2][2] = 1.0; g[3][3] = 1.0;
        for (int m = 0; m < 4; ++m)
            for (int n = 0; n < 4; ++n)
                g[m][n] -= Hs * lu[m] * lu[n];
    }
    double ham_ks(double X, double Y, double Z, const double p[4]) const {  // :221-230
        double g[4][4];
        upper_ks(X, Y, Z, g);
        double h = 0.0;
        for (int m = 0; m < 4; ++m)
            for (int n = 0; n < 4; ++n)
                h += g[m][n]*p[m]*p[n];
        return 0.5*h;
    }

    // ---- horizon / ISCO / Omega (:237-363) ----
    double horizon() const {
        if (std::abs(Lambda) < 1e-15) {
            const double disc = M*M - a*a - Q*Q;
            return M + std::sqrt(std::max(disc, 0.0));
        }
        const double scale = std::max(std::abs(M), 1.0);
        const double rmin = std::max(1e-6*scale, 1e-9);
        double rmax;
        if (Lambda > 0.0) rmax = std::max(20.0*scale, 1.25*std::sqrt(3.0/Lambda));
        else              rmax = 200.0*scale + 20.0*std::abs(a) + 20.0*std::abs(Q);
        if (!(rmax > rmin)) rmax = rmin * 10.0;

        auto bisect = [&](double lo, double hi) -> double {
            double flo = delta_r(lo), fhi = delta_r(hi);
            if (!std::isfinite(flo) || !std::isfinite(fhi)) return std::numeric_limits<double>::quiet_NaN();
            for (int i = 0; i < 160; ++i) {
                const double mid = 0.5*(lo + hi);
                const double fm = delta_r(mid);
                if (!std::isfinite(fm)) break;
                if ((flo > 0.0 && fm < 0.0) || (flo < 0.0 && fm > 0.0)) { hi = mid; fhi = fm; }
                else { lo = mid; flo = fm; }
            }
            (void)fhi;
            return 0.5*(lo + hi);
        };
        double roots[16];
        int nr = 0;
        auto add_root = [&](double rr) {
            if (!std::isfinite(rr) || rr <= 0.0) return;
            for (int i = 0; i < nr; ++i) {
                const double rel = std::max({1.0, std::abs(roots[i]), std::abs(rr)});
                if (std::abs(roots[i] - rr) <= 1e-8 * rel) return;
            }
            if (nr < 16) roots[nr++] = rr;
        };
        const int N = 20000;
        const double span = std::log(rmax / rmin);
        double r0 = rmin, f0 = delta_r(r0);
        for (int i = 1; i <= N; ++i) {
            const double t = double(i) / double(N);
            const double r1 = rmin * std::exp(span * t);
            const double f1 = delta_r(r1);
            if (std::isfinite(f0) && std::isfinite(f1)) {
                if (f0 == 0.0) add_root(r0);
                if (f1 == 0.0) add_root(r1);
                if ((f0 > 0.0 && f1 < 0.0) || (f0 < 0.0 && f1 > 0.0)) add_root(bisect(r0, r1));
            }
            r0 = r1; f0 = f1;
        }
        if (nr == 0) {
            const double disc = M*M - a*a - Q*Q;
            return M + std::sqrt(std::max(disc, 0.0));
        }
        std::sort(roots, roots + nr);
        if (Lambda > 0.0 && nr >= 2) return roots[nr - 2];
        return roots[nr - 1];
    }

    double omega_k(double r) const {                       // :354-363
        const double s = (a < 0.0) ? 1.0 : -1.0;
        const double Me = M - Q*Q/(2.0*r) + Lambda*a*r*r/3.0;
        const double sq = std::sqrt(std::max(Me, 0.0));
        const double den = r*std::sqrt(r) + s*a*sq;
        if (std::abs(den) < 1e-14) return 0.0;
        return s * sq / den;
    }

    double isco() const {                                  // :318-349
        if (std::abs(Lambda) < 1e-15 && std::abs(Q) < 1e-15) {
            if (std::abs(a) < 1e-10) return 6.0*M;
            double ap = a/M;
            double z1 = 1.0 + std::cbrt(1.0-ap*ap) * (std::cbrt(1.0+ap)+std::cbrt(1.0-ap));
            double z2 = std::sqrt(3.0*ap*ap + z1*z1);
            return M*(3.0 + z2 - std::sqrt((3.0-z1)*(3.0+z1+2.0*z2)));
        }
        double rh = horizon();
        double r = rh * 1.05;
        double dr = 1e-4*M;
        auto Lc = [&](double rr) -> double {
            double Om = omega_k(rr);
            double g[4][4];
            lower_bl(rr, PI/2.0, g);
            double N2 = -(g[0][0] + 2.0*g[0][3]*Om + g[3][3]*Om*Om);
            if (N2 <= 0.0) return -1e30;
            double ut = 1.0/std::sqrt(N2);
            return -(g[3][0] + g[3][3]*Om)*ut;
        };
        double Lp = Lc(r);
        for (; r < 30.0*M; r += dr) {
            double L = Lc(r + dr);
            if (L < Lp && Lp > 0.0) return r;
            Lp = L;
        }
        return r;
    }

    static void bl_to_ks(double r, double th, double ph, double as, double& X, double& Y, double& Z) {  // :368-386
        const double st = std::sin(th), ct = std::cos(th);
        const double sf = std::sin(ph), cf = std::cos(ph);
        X = st*(r*cf - as*sf);
        Y = st*(r*sf + as*cf);
        Z = r*ct;
    }
    static void ks_to_bl(double X, double Y, double Z, double as, double& r, double& th, double& ph) {  // :389-411
        r = ks_radius(X, Y, Z, as);
        double Zr = Z/r;
        th = std::acos(Zr < -1.0 ? -1.0 : Zr > 1.0 ? 1.0 : Zr);
        double st = std::sqrt(std::max(1.0 - Z*Z/(r*r), 0.0));
        double ra = r*r + as*as;
        if (st > 1e-10) {
            double cf = (X*r + Y*as) / (st * ra);
            double sf = (Y*r - X*as) / (st * ra);
            ph = std::atan2(sf, cf);
        } else {
            ph = 0.0;
        }
    }
};

// ---------------------------------------------------------------------------
//  Geodesic integrators (geodesic.hpp)
// ---------------------------------------------------------------------------
// Reference behaviours the oracle reproduces by default (same bit values as the KT_QUIRK_*
// macros of include/kt_hip.h); a cleared bit runs the corrected behaviour instead.
enum : uint32_t {
    QK_DP5_PHI_FROZEN = 1u << 0,      // DOPRI5 never advances phi            geodesic.hpp:300-310
    QK_ESC_POSITIONAL = 1u << 1,      // positional TraceResult fill          main.cpp:360,428-433
    QK_KS_INIT_NO_SHIFT = 1u << 2,    // KS covector without the shift terms  main.cpp:1373-1430
    QK_BUNDLE_HORIZON_ESC = 1u << 3,  // bundle horizon -> empty result       ray_bundle.hpp:426-428
    QK_BUNDLE_ROOT_SKIP = 1u << 4,    // failed bundle root skips the tests   ray_bundle.hpp:357-365
    QK_REFERENCE = 0x1Fu,
};

struct Ray { double r, th, ph, pr, pth, pt, pp; };        // geodesic.hpp:33-37

static inline double phi_rate(const Spacetime& g, double r, double th, double pt, double pp) {  // :40-45
    double u[4][4];
    g.upper_bl(r, th, u);
    return u[3][0]*pt + u[3][3]*pp;
}

static inline double unit_clamp(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }   // :49-53

static inline double hermite(double y0, double y1, double d0, double d1, double h, double al) {  // :55-66
    al = unit_clamp(al);
    const double a2 = al*al;
    const double a3 = a2*al;
    const double h00 =  2.0*a3 - 3.0*a2 + 1.0;
    const double h10 =        a3 - 2.0*a2 + al;
    const double h01 = -2.0*a3 + 3.0*a2;
    const double h11 =        a3 -       a2;
    return h00*y0 + h10*(h*d0) + h01*y1 + h11*(h*d1);
}

static inline bool crosses(double f0, double f1) { return ((f0 <= 0.0 && f1 >= 0.0) || (f0 >= 0.0 && f1 <= 0.0)); }  // :68-70

// first root of the Hermite cubic in [0,1]: 8 bins x 8 bisections   :97-140
static bool hermite_first_root(double y0, double y1, double d0, double d1, double h, double target,
                               double& out, int bins = 8, int iters = 8) {
    bins = std::max(2, bins);
    const double feps = 1e-12;
    auto f = [&](double al) { return hermite(y0, y1, d0, d1, h, al) - target; };
    double a0 = 0.0;
    double f0 = y0 - target;
    if (std::abs(f0) <= feps) { out = 0.0; return true; }
    for (int i = 1; i <= bins; ++i) {
        const double a1 = double(i) / double(bins);
        const double f1 = (i == bins) ? (y1 - target) : f(a1);
        if (std::abs(f1) <= feps) { out = a1; return true; }
        if (crosses(f0, f1)) {
            double lo = a0, hi = a1;
            for (int k = 0; k < iters; ++k) {
                const double mid = 0.5*(lo + hi);
                const double fm = f(mid);
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

// Hamiltonian RHS with central differences   :143-166
static inline void bl_rhs(const Spacetime& g, double r, double th, double pr, double pth, double pt, double pp,
                          double& dr, double& dth, double& dpr, double& dpth) {
    double u[4][4];
    g.upper_bl(r, th, u);
    dr  = u[1][1] * pr;
    dth = u[2][2] * pth;
    const double er = 1e-5 * std::max(std::abs(r), 0.01);
    const double et = 1e-6;
    dpr  = -(g.ham(r+er, th, pr, pth, pt, pp) - g.ham(r-er, th, pr, pth, pt, pp)) / (2.0*er);
    dpth = -(g.ham(r, th+et, pr, pth, pt, pp) - g.ham(r, th-et, pr, pth, pt, pp)) / (2.0*et);
}

// classic RK4 with phi on the same stage points   :169-194
static void rk4(const Spacetime& g, Ray& s, double h) {
    double dr1,dt1,dpr1,dpt1, dr2,dt2,dpr2,dpt2, dr3,dt3,dpr3,dpt3, dr4,dt4,dpr4,dpt4;
    const double r=s.r, th=s.th, pr=s.pr, pth=s.pth, pt=s.pt, pp=s.pp;
    bl_rhs(g, r,            th,            pr,             pth,             pt,pp, dr1,dt1,dpr1,dpt1);
    bl_rhs(g, r+.5*h*dr1,   th+.5*h*dt1,   pr+.5*h*dpr1,   pth+.5*h*dpt1,   pt,pp, dr2,dt2,dpr2,dpt2);
    bl_rhs(g, r+.5*h*dr2,   th+.5*h*dt2,   pr+.5*h*dpr2,   pth+.5*h*dpt2,   pt,pp, dr3,dt3,dpr3,dpt3);
    bl_rhs(g, r+   h*dr3,   th+   h*dt3,   pr+   h*dpr3,   pth+   h*dpt3,   pt,pp, dr4,dt4,dpr4,dpt4);
    const double f1 = phi_rate(g, r,          th,          pt, pp);
    const double f2 = phi_rate(g, r+.5*h*dr1, th+.5*h*dt1, pt, pp);
    const double f3 = phi_rate(g, r+.5*h*dr2, th+.5*h*dt2, pt, pp);
    const double f4 = phi_rate(g, r+   h*dr3, th+   h*dt3, pt, pp);
    s.r   += h/6.0*(dr1  + 2*dr2  + 2*dr3  + dr4);
    s.th  += h/6.0*(dt1  + 2*dt2  + 2*dt3  + dt4);
    s.ph  += h/6.0*(f1   + 2*f2   + 2*f3   + f4);
    s.pr  += h/6.0*(dpr1 + 2*dpr2 + 2*dpr3 + dpr4);
    s.pth += h/6.0*(dpt1 + 2*dpt2 + 2*dpt3 + dpt4);
}

// step doubling   :204-257
static bool rk4_doubling(const Spacetime& g, Ray& s, double& h, double tol) {
    const Ray s0 = s;
    Ray A = s0; rk4(g, A, h);
    Ray B = s0; rk4(g, B, 0.5*h); rk4(g, B, 0.5*h);
    const double err = std::sqrt(
        (A.r   - B.r)   * (A.r   - B.r)   +
        (A.th  - B.th)  * (A.th  - B.th)  +
        (A.pr  - B.pr)  * (A.pr  - B.pr)  +
        (A.pth - B.pth) * (A.pth - B.pth)) / 15.0;
    if (!std::isfinite(err)) {
        s = s0;
        h = (std::isfinite(h) && h > 1e-10) ? h * 0.5 : 1e-6;
        if (h < 1e-10) h = 1e-10;
        return false;
    }
    const bool ok = (err < tol || h < 1e-10);
    if (ok) {
        s = B;
        const double sc = (err > 1e-14) ? 0.9 * std::pow(tol/err, 0.2) : 4.0;
        double hn = h * sc;
        if (!std::isfinite(hn)) hn = h;
        h = hn;
        if (h > 100.0) h = 100.0;
        if (h < 1e-10) h = 1e-10;
    } else {
        s = s0;
        const double half = h * 0.5;
        double hn = h * 0.9 * std::pow(tol/err, 0.25);
        if (!std::isfinite(hn)) hn = half;
        h = hn;
        if (h > half) h = half;
        if (h < 1e-10) h = 1e-10;
    }
    return ok;
}

// Dormand-Prince 5(4) with FSAL; phi is not advanced (reference behaviour)   :261-414
struct K4 { double v[4]; };
static inline K4 nan4() { K4 k; k.v[0]=k.v[1]=k.v[2]=k.v[3]=std::numeric_limits<double>::quiet_NaN(); return k; }
static inline void rhs4(const Spacetime& g, const Ray& s, K4& k) {
    bl_rhs(g, s.r, s.th, s.pr, s.pth, s.pt, s.pp, k.v[0], k.v[1], k.v[2], k.v[3]);
}
static inline Ray stage(const Ray& s0, double h, const K4* ks, const double* w, int n) {   // :294-310
    Ray s = s0;
    for (int i = 0; i < n; ++i) {
        s.r   += h * w[i] * ks[i].v[0];
        s.th  += h * w[i] * ks[i].v[1];
        s.pr  += h * w[i] * ks[i].v[2];
        s.pth += h * w[i] * ks[i].v[3];
    }
    return s;
}
// adv_phi (quirk QK_DP5_PHI_FROZEN cleared): phi advanced with the 5th-order weights from
// dphi/dl at the six stage points, as RK4 does (geodesic.hpp:184-191); the reference's
// DOPRI5 leaves phi at its initial value (advance(), geodesic.hpp:300-310).
static bool dopri5(const Spacetime& g, Ray& s, double& h, K4& fsal, double tol, bool adv_phi = false) {
    K4 k1;
    if (!std::isnan(fsal.v[0])) k1 = fsal; else rhs4(g, s, k1);
    static const double c2[] = {1.0/5.0};
    static const double c3[] = {3.0/40.0, 9.0/40.0};
    static const double c4[] = {44.0/45.0, -56.0/15.0, 32.0/9.0};
    static const double c5[] = {19372.0/6561.0, -25360.0/2187.0, 64448.0/6561.0, -212.0/729.0};
    static const double c6[] = {9017.0/3168.0, -355.0/33.0, 46732.0/5247.0, 49.0/176.0, -5103.0/18656.0};
    static const double b5[] = {35.0/384.0, 0.0, 500.0/1113.0, 125.0/192.0, -2187.0/6784.0, 11.0/84.0};
    static const double b4[] = {5179.0/57600.0, 0.0, 7571.0/16695.0, 393.0/640.0, -92097.0/339200.0,
                                187.0/2100.0, 1.0/40.0};
    K4 k[7];
    k[0] = k1;
    Ray s2 = stage(s, h, k, c2, 1); rhs4(g, s2, k[1]);
    Ray s3 = stage(s, h, k, c3, 2); rhs4(g, s3, k[2]);
    Ray s4 = stage(s, h, k, c4, 3); rhs4(g, s4, k[3]);
    Ray s5 = stage(s, h, k, c5, 4); rhs4(g, s5, k[4]);
    Ray s6 = stage(s, h, k, c6, 5); rhs4(g, s6, k[5]);
    const Ray y5 = stage(s, h, k, b5, 6);
    rhs4(g, y5, k[6]);
    const Ray y4 = stage(s, h, k, b4, 7);
    const double dy[4] = {y5.r - y4.r, y5.th - y4.th, y5.pr - y4.pr, y5.pth - y4.pth};
    const double y0[4] = {s.r, s.th, s.pr, s.pth};
    double e2 = 0.0;
    for (int i = 0; i < 4; ++i) {
        double q = dy[i] / (tol * (1.0 + std::abs(y0[i])));
        e2 += q*q;
    }
    const double err = std::sqrt(e2 / 4.0);
    if (!std::isfinite(err)) {
        h = (std::isfinite(h) && h > 1e-10) ? h * 0.5 : 1e-6;
        if (h < 1e-10) h = 1e-10;
        fsal.v[0] = std::numeric_limits<double>::quiet_NaN();
        return false;
    }
    const bool ok = (err <= 1.0 || h < 1e-10);
    if (ok) {
        if (adv_phi) {
            const Ray* st[6] = {&s, &s2, &s3, &s4, &s5, &s6};
            double dph = 0.0;
            for (int i = 0; i < 6; ++i)
                if (b5[i] != 0.0) dph += b5[i] * phi_rate(g, st[i]->r, st[i]->th, s.pt, s.pp);
            const double ph1 = s.ph + h * dph;
            s = y5;
            s.ph = ph1;
        } else {
            s = y5;
        }
        fsal = k[6];
        const double fac = (err > 1e-12) ? 0.9 * std::pow(1.0/err, 0.2) : 4.0;
        double hn = h * fac;
        if (!std::isfinite(hn)) hn = h;
        if (hn > 100.0) hn = 100.0;
        if (hn < 1e-10) hn = 1e-10;
        h = hn;
    } else {
        double hn = h * 0.9 * std::pow(1.0/err, 0.25);
        if (!std::isfinite(hn)) hn = h * 0.5;
        h = std::max(hn, 1e-10);
        fsal.v[0] = std::numeric_limits<double>::quiet_NaN();
    }
    return ok;
}

enum Method { RK4D = 0, DP5 = 1 };
static inline bool take_step(const Spacetime& g, Ray& s, double& h, int method, K4& fsal, double tol,
                             uint32_t quirks = QK_REFERENCE) {  // :423-431
    if (method == DP5) return dopri5(g, s, h, fsal, tol, (quirks & QK_DP5_PHI_FROZEN) == 0);
    return rk4_doubling(g, s, h, tol);
}

// ---------------------------------------------------------------------------
//  Camera (camera.hpp:20-145)
// ---------------------------------------------------------------------------
struct Cam {
    double r_obs, th_obs, ph_obs, fov;   // radians
    int W, H;
    Cam(double r, double th_deg, double ph_deg, double fov_deg, int w, int h)
        : r_obs(r), th_obs(th_deg * PI / 180.0), ph_obs(ph_deg * PI / 180.0),
          fov(fov_deg * PI / 180.0), W(w), H(h) {}

    void screen_angles(double px, double py, double& al, double& be) const {   // :38-47
        const int span = (W > 1) ? (W - 1) : 1;
        al = fov * (px - 0.5*(W-1)) / span;
        be = fov * (0.5*(H-1) - py) / span;
    }

    Ray shoot(double al, double be, const Spacetime& g) const {   // :50-145
        const double r0 = r_obs, t0 = th_obs;
        double L[4][4];
        g.lower_bl(r0, t0, L);
        const double ca = std::cos(al), sa = std::sin(al);
        const double cb = std::cos(be), sb = std::sin(be);
        const double gtt = L[0][0], gtp = L[0][3], grr = L[1][1], gth = L[2][2], gpp = L[3][3];
        const double sq_rr = std::sqrt(std::abs(grr));
        const double sq_th = std::sqrt(std::abs(gth));
        const double dphi = gpp - (gtp*gtp)/gtt;
        const bool frame_ok = (gtt < -1e-14) && std::isfinite(dphi) && (dphi > 1e-14) &&
                              std::isfinite(sq_rr) && (sq_rr > 1e-14) &&
                              std::isfinite(sq_th) && (sq_th > 1e-14);
        double ut_ = 1.0;
        double ur_ = -ca*cb / std::max(sq_rr, 1e-14);
        double uth = -sb    / std::max(sq_th, 1e-14);
        double uph = -sa*cb / std::sqrt(std::max(std::abs(gpp), 1e-14));
        if (frame_ok) {
            const double u0 = 1.0 / std::sqrt(-gtt);
            const double ep = 1.0 / std::sqrt(dphi);
            const double et = -gtp/gtt * ep;
            const double nr = -ca * cb;
            const double nt = -sb;
            const double np = -sa * cb;
            ut_ = u0 + np * et;
            ur_ = nr / sq_rr;
            uth = nt / sq_th;
            uph = np * ep;
        }
        double pt  = gtt*ut_ + gtp*uph;
        double pr  = grr*ur_;
        double pth = gth*uth;
        double pp  = gtp*ut_ + gpp*uph;
        double U[4][4];
        g.upper_bl(r0, t0, U);
        const double A = U[0][0];
        const double B = 2.0 * U[0][3] * pp;
        const double C = U[1][1]*pr*pr + U[2][2]*pth*pth + U[3][3]*pp*pp;
        const double disc = B*B - 4.0*A*C;
        if (disc >= 0.0 && std::abs(A) > 1e-15) {
            const double sq = std::sqrt(disc);
            const double p1 = (-B - sq) / (2.0*A);
            const double p2 = (-B + sq) / (2.0*A);
            pt = (p1 < 0.0) ? p1 : p2;
            if (pt > 0.0) pt = std::min(p1, p2);
        }
        return {r0, t0, ph_obs, pr, pth, pt, pp};
    }
    Ray pixel(int px, int py, const Spacetime& g, double ox, double oy) const {
        double al, be;
        screen_angles((double)px + ox, (double)py + oy, al, be);
        return shoot(al, be, g);
    }
};

// ---------------------------------------------------------------------------
//  Per-ray tracers
// ---------------------------------------------------------------------------
enum Fate { ESC = 0, DISK = 1, HOLE = 2, ESC_B = 3 };
struct Result {                     // main.cpp:187-191 (TraceResult) + bundle magnification
    int fate; double r, g;
    double ph_disk = 0.0, th_esc = 0.0, ph_esc = 0.0;
    double magnif = 1.0;
};
struct Counters { int64_t accepted = 0, rejected = 0; };

struct Controls { int max_steps; double step_init; double tol; uint32_t quirks = QK_REFERENCE; };

static double disk_g(double r, double pt, double pp, const Spacetime& g) {   // main.cpp:329-341
    double Om = g.omega_k(r);
    double b = -pp / (-pt);
    double L[4][4]; g.lower_bl(r, PI/2.0, L);
    double d2 = -(L[0][0]+2.0*L[0][3]*Om+L[3][3]*Om*Om);
    if (d2 <= 0.0) return 1.0;
    const double dv = std::abs(1.0 - Om*b);
    if (dv < 1e-8) return 20.0;
    const double red = std::sqrt(d2) / dv;
    return std::isfinite(red) ? red : 1.0;
}

// Escape-type TraceResult {ESCAPED, r, 1, theta, phi}.  The reference fills it positionally
// (main.cpp:360,428-433): phi_disk <- theta, theta_esc <- phi, phi_esc <- 0; with the quirk
// cleared the angles go to theta_esc / phi_esc.
static Result esc_result(double r, double th, double ph, uint32_t quirks) {
    Result R{ESC, r, 1.0};
    if (quirks & QK_ESC_POSITIONAL) { R.ph_disk = th; R.th_esc = ph; R.ph_esc = 0.0; }
    else { R.ph_disk = 0.0; R.th_esc = th; R.ph_esc = ph; }
    return R;
}

// BL driver   main.cpp:343-434
static Result trace_bl(Ray s, const Spacetime& g, double rin, double rout, double resc, double rh,
                       int method, const Controls& c, Counters* cnt) {
    double h = std::max(c.step_init, 1e-10);
    const double rcut = rh * 1.03;
    K4 fsal = nan4();
    const int nmax = std::max(1, c.max_steps);
    for (int it = 0; it < nmax; ++it) {
        const Ray sp = s;
        double hu = h;
        int rej = 0;
        while (true) {
            hu = h;
            if (take_step(g, s, h, method, fsal, c.tol, c.quirks)) { if (cnt) cnt->accepted++; break; }
            if (cnt) cnt->rejected++;
            if (!std::isfinite(h) || ++rej > 64) return esc_result(s.r, s.th, s.ph, c.quirks);   // :360
        }
        double best = 2.0;
        int ev = 0;   // 0 none, 1 disk, 2 horizon, 3 escape
        double d_r = 0.0, d_g = 1.0, d_ph = 0.0;
        const double q0 = sp.th - PI/2.0, q1 = s.th - PI/2.0;
        const bool near_eq = crosses(q0, q1) || (std::min(std::abs(q0), std::abs(q1)) < 0.35);
        if (near_eq) {
            double dr0,dt0,dpr0,dpt0, dr1,dt1,dpr1,dpt1;
            bl_rhs(g, sp.r, sp.th, sp.pr, sp.pth, sp.pt, sp.pp, dr0,dt0,dpr0,dpt0);
            bl_rhs(g, s.r,  s.th,  s.pr,  s.pth,  s.pt,  s.pp,  dr1,dt1,dpr1,dpt1);
            double al = 0.0;
            if (hermite_first_root(sp.th, s.th, dt0, dt1, hu, PI/2.0, al, 8, 8)) {
                const double rhit = hermite(sp.r, s.r, dr0, dr1, hu, al);
                if (rhit >= rin && rhit <= rout) {
                    d_r = rhit;
                    d_g = clampv(disk_g(rhit, s.pt, s.pp, g), 0.0, 20.0);
                    d_ph = sp.ph + al * (s.ph - sp.ph);
                    best = al;
                    ev = 1;
                }
            }
        }
        const bool hz = ((sp.r > rcut) && (s.r <= rcut)) || (s.r <= rcut);
        if (hz) {
            const double dn = sp.r - s.r;
            double ah = (std::abs(dn) > 1e-12) ? ((sp.r - rcut) / dn) : 0.0;
            ah = clampv(ah, 0.0, 1.0);
            if (ah < best) { best = ah; ev = 2; }
        }
        const bool es = ((sp.r < resc) && (s.r >= resc)) || (s.r >= resc);
        if (es) {
            const double dn = s.r - sp.r;
            double ae = (std::abs(dn) > 1e-12) ? ((resc - sp.r) / dn) : 1.0;
            ae = clampv(ae, 0.0, 1.0);
            if (ae < best) { best = ae; ev = 3; }
        }
        if (ev == 1) { Result R{DISK, d_r, d_g}; R.ph_disk = d_ph; return R; }
        if (ev == 2) { Result R{HOLE, sp.r + best * (s.r - sp.r), 0.0}; return R; }
        if (ev == 3)    // {ESCAPED, r_escape, 1.0, th_esc, ph_esc} (main.cpp:428-430)
            return esc_result(resc, sp.th + best * (s.th - sp.th), sp.ph + best * (s.ph - sp.ph), c.quirks);
    }
    return esc_result(s.r, s.th, s.ph, c.quirks);      // :433
}

// ---- KS tracer (main.cpp:1303-1660) ----
struct KSRay { double X, Y, Z, pX, pY, pZ, pT; };

static bool solve3(const double Ain[3][3], const double bin[3], double x[3]) {   // :1314-1346
    double A[3][4] = {
        {Ain[0][0], Ain[0][1], Ain[0][2], bin[0]},
        {Ain[1][0], Ain[1][1], Ain[1][2], bin[1]},
        {Ain[2][0], Ain[2][1], Ain[2][2], bin[2]},
    };
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        double best = std::abs(A[piv][c]);
        for (int r = c + 1; r < 3; ++r) {
            const double v = std::abs(A[r][c]);
            if (v > best) { best = v; piv = r; }
        }
        if (best < 1e-14) return false;
        if (piv != c) for (int k = c; k < 4; ++k) std::swap(A[c][k], A[piv][k]);
        const double inv = 1.0 / A[c][c];
        for (int k = c; k < 4; ++k) A[c][k] *= inv;
        for (int r = 0; r < 3; ++r) {
            if (r == c) continue;
            const double f = A[r][c];
            for (int k = c; k < 4; ++k) A[r][k] -= f * A[c][k];
        }
    }
    x[0] = A[0][3]; x[1] = A[1][3]; x[2] = A[2][3];
    return true;
}

static void bl_ks_jacobian(double r, double th, double ph, double as, double J[3][3]) {   // :1348-1371
    const double er = 1e-6 * (std::abs(r) + 1.0);
    const double et = 1e-6;
    const double ep = 1e-6;
    auto f = [&](double rr, double tt, double pp) {
        double X, Y, Z;
        Spacetime::bl_to_ks(rr, tt, pp, as, X, Y, Z);
        return std::array<double, 3>{X, Y, Z};
    };
    const auto rp = f(r + er, th, ph), rm = f(r - er, th, ph);
    const auto tp = f(r, th + et, ph), tm = f(r, th - et, ph);
    const auto pp_ = f(r, th, ph + ep), pm = f(r, th, ph - ep);
    for (int i = 0; i < 3; ++i) {
        J[i][0] = (rp[i] - rm[i]) / (2.0 * er);
        J[i][1] = (tp[i] - tm[i]) / (2.0 * et);
        J[i][2] = (pp_[i] - pm[i]) / (2.0 * ep);
    }
}

static bool covec_bl_to_ks(double r, double th, double ph, double as, double pr, double pth, double pp,
                           double& pX, double& pY, double& pZ) {   // :1373-1389
    double J[3][3];
    bl_ks_jacobian(r, th, ph, as, J);
    double A[3][3];
    for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) A[j][i] = J[i][j];
    const double b[3] = {pr, pth, pp};
    double x[3];
    if (!solve3(A, b, x)) return false;
    pX = x[0]; pY = x[1]; pZ = x[2];
    return true;
}

static std::array<double, 3> covec_ks_to_bl(double r, double th, double ph, double as,
                                            double pX, double pY, double pZ) {   // :1391-1402
    double J[3][3];
    bl_ks_jacobian(r, th, ph, as, J);
    std::array<double, 3> out{};
    const double p[3] = {pX, pY, pZ};
    for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) out[j] += J[i][j] * p[i];
    return out;
}

// exact = false (quirk QK_KS_INIT_NO_SHIFT, the reference): the BL spatial covector is
// carried over as is and p_T re-solved from H = 0.  exact = true: the full covector map.
// With T = v - r and phi~ the ingoing Kerr coordinates (dv = dt + (r^2+a^2)/Delta dr,
// dphi~ = dphi + a/Delta dr), p_T = p_t and the radial component seen at fixed phi~ is
//   p_r + p_t - ((r^2+a^2) p_t + a p_phi) / Delta,
// which is what the reference's solve leaves out (its dT/dr and dphi~/dr terms).
static bool ks_init(const Ray& b, const Spacetime& g, KSRay& k, bool exact = false) {   // :1404-1430
    if (std::abs(g.Lambda) > 1e-15) return false;
    Spacetime::bl_to_ks(b.r, b.th, b.ph, g.a, k.X, k.Y, k.Z);
    if (exact) {
        const double ra = b.r*b.r + g.a*g.a;
        const double dl = ra - 2.0*g.M*b.r + g.Q*g.Q;
        const double pr_T = b.pr + b.pt - (ra*b.pt + g.a*b.pp) / dl;
        if (!covec_bl_to_ks(b.r, b.th, b.ph, g.a, pr_T, b.pth, b.pp, k.pX, k.pY, k.pZ)) return false;
        k.pT = b.pt;
        return std::isfinite(k.pT);
    }
    if (!covec_bl_to_ks(b.r, b.th, b.ph, g.a, b.pr, b.pth, b.pp, k.pX, k.pY, k.pZ)) return false;
    double u[4][4];
    g.upper_ks(k.X, k.Y, k.Z, u);
    const double A = u[0][0];
    const double B = 2.0 * (u[0][1] * k.pX + u[0][2] * k.pY + u[0][3] * k.pZ);
    const double C =
        u[1][1] * k.pX * k.pX + u[2][2] * k.pY * k.pY + u[3][3] * k.pZ * k.pZ +
        2.0 * u[1][2] * k.pX * k.pY + 2.0 * u[1][3] * k.pX * k.pZ + 2.0 * u[2][3] * k.pY * k.pZ;
    const double disc = B * B - 4.0 * A * C;
    if (disc < 0.0 || std::abs(A) < 1e-15) return false;
    const double sq = std::sqrt(disc);
    const double p1 = (-B - sq) / (2.0 * A);
    const double p2 = (-B + sq) / (2.0 * A);
    k.pT = (p1 < 0.0) ? p1 : p2;
    if (k.pT > 0.0) k.pT = std::min(p1, p2);
    return std::isfinite(k.pT);
}

static void ks_rhs(const Spacetime& g, const KSRay& s, double& dX, double& dY, double& dZ,
                   double& dpX, double& dpY, double& dpZ) {   // :1432-1454
    double u[4][4];
    g.upper_ks(s.X, s.Y, s.Z, u);
    dX = u[1][0] * s.pT + u[1][1] * s.pX + u[1][2] * s.pY + u[1][3] * s.pZ;
    dY = u[2][0] * s.pT + u[2][1] * s.pX + u[2][2] * s.pY + u[2][3] * s.pZ;
    dZ = u[3][0] * s.pT + u[3][1] * s.pX + u[3][2] * s.pY + u[3][3] * s.pZ;
    auto Hx = [&](double X, double Y, double Z) {
        double p[4] = {s.pT, s.pX, s.pY, s.pZ};
        return g.ham_ks(X, Y, Z, p);
    };
    const double eX = 1e-5 * (std::abs(s.X) + 0.1);
    const double eY = 1e-5 * (std::abs(s.Y) + 0.1);
    const double eZ = 1e-5 * (std::abs(s.Z) + 0.1);
    dpX = -(Hx(s.X + eX, s.Y, s.Z) - Hx(s.X - eX, s.Y, s.Z)) / (2.0 * eX);
    dpY = -(Hx(s.X, s.Y + eY, s.Z) - Hx(s.X, s.Y - eY, s.Z)) / (2.0 * eY);
    dpZ = -(Hx(s.X, s.Y, s.Z + eZ) - Hx(s.X, s.Y, s.Z - eZ)) / (2.0 * eZ);
}

static void ks_rk4(const Spacetime& g, KSRay& s, double h) {   // :1456-1486
    double a[6], b[6], c[6], d[6];
    const KSRay s0 = s;
    ks_rhs(g, s0, a[0], a[1], a[2], a[3], a[4], a[5]);
    KSRay s2{s0.X + 0.5 * h * a[0], s0.Y + 0.5 * h * a[1], s0.Z + 0.5 * h * a[2],
             s0.pX + 0.5 * h * a[3], s0.pY + 0.5 * h * a[4], s0.pZ + 0.5 * h * a[5], s0.pT};
    ks_rhs(g, s2, b[0], b[1], b[2], b[3], b[4], b[5]);
    KSRay s3{s0.X + 0.5 * h * b[0], s0.Y + 0.5 * h * b[1], s0.Z + 0.5 * h * b[2],
             s0.pX + 0.5 * h * b[3], s0.pY + 0.5 * h * b[4], s0.pZ + 0.5 * h * b[5], s0.pT};
    ks_rhs(g, s3, c[0], c[1], c[2], c[3], c[4], c[5]);
    KSRay s4{s0.X + h * c[0], s0.Y + h * c[1], s0.Z + h * c[2],
             s0.pX + h * c[3], s0.pY + h * c[4], s0.pZ + h * c[5], s0.pT};
    ks_rhs(g, s4, d[0], d[1], d[2], d[3], d[4], d[5]);
    s.X  += h / 6.0 * (a[0] + 2.0 * b[0] + 2.0 * c[0] + d[0]);
    s.Y  += h / 6.0 * (a[1] + 2.0 * b[1] + 2.0 * c[1] + d[1]);
    s.Z  += h / 6.0 * (a[2] + 2.0 * b[2] + 2.0 * c[2] + d[2]);
    s.pX += h / 6.0 * (a[3] + 2.0 * b[3] + 2.0 * c[3] + d[3]);
    s.pY += h / 6.0 * (a[4] + 2.0 * b[4] + 2.0 * c[4] + d[4]);
    s.pZ += h / 6.0 * (a[5] + 2.0 * b[5] + 2.0 * c[5] + d[5]);
}

static bool ks_doubling(const Spacetime& g, KSRay& s, double& h, double tol) {   // :1488-1528
    const KSRay s0 = s;
    KSRay A = s0; ks_rk4(g, A, h);
    KSRay B = s0; ks_rk4(g, B, 0.5 * h); ks_rk4(g, B, 0.5 * h);
    const double err = std::sqrt(
        (A.X  - B.X)  * (A.X  - B.X)  +
        (A.Y  - B.Y)  * (A.Y  - B.Y)  +
        (A.Z  - B.Z)  * (A.Z  - B.Z)  +
        (A.pX - B.pX) * (A.pX - B.pX) +
        (A.pY - B.pY) * (A.pY - B.pY) +
        (A.pZ - B.pZ) * (A.pZ - B.pZ)) / 15.0;
    if (!std::isfinite(err)) {
        s = s0;
        h = (std::isfinite(h) && h > 1e-10) ? h * 0.5 : 1e-6;
        if (h < 1e-10) h = 1e-10;
        return false;
    }
    const bool ok = (err < tol || h < 1e-10);
    if (ok) {
        s = B;
        const double sc = (err > 1e-14) ? 0.9 * std::pow(tol / err, 0.2) : 4.0;
        double hn = h * sc;
        if (!std::isfinite(hn)) hn = h;
        h = clampv(hn, 1e-10, 100.0);
    } else {
        s = s0;
        const double half = h * 0.5;
        double hn = h * 0.9 * std::pow(tol / err, 0.25);
        if (!std::isfinite(hn)) hn = half;
        h = hn;
        if (h > half) h = half;
        if (h < 1e-10) h = 1e-10;
    }
    return ok;
}

static Result trace_ks(const Ray& b, const Spacetime& g, double rin, double rout, double resc, double rh,
                       int method, const Controls& c, Counters* cnt) {   // :1530-1660
    KSRay s{};
    if (!ks_init(b, g, s, (c.quirks & QK_KS_INIT_NO_SHIFT) == 0)) return trace_bl(b, g, rin, rout, resc, rh, method, c, cnt);
    const double rcut = rh * 1.03;
    double h = std::max(c.step_init, 1e-10);
    double rprev = Spacetime::ks_radius(s.X, s.Y, s.Z, g.a);
    const int nmax = std::max(1, c.max_steps);
    for (int it = 0; it < nmax; ++it) {
        const KSRay sp = s;
        double hu = h;
        int rej = 0;
        while (true) {
            hu = h;
            if (ks_doubling(g, s, h, c.tol)) { if (cnt) cnt->accepted++; break; }
            if (cnt) cnt->rejected++;
            if (!std::isfinite(h) || ++rej > 64) {
                double rr, tt, pp;
                Spacetime::ks_to_bl(s.X, s.Y, s.Z, g.a, rr, tt, pp);
                return esc_result(rr, tt, pp, c.quirks);   // :1572
            }
        }
        const double rnow = Spacetime::ks_radius(s.X, s.Y, s.Z, g.a);
        double best = 2.0;
        int ev = 0;
        double d_r = 0.0, d_g = 1.0, d_ph = 0.0;
        const bool near_eq = crosses(sp.Z, s.Z) || (std::min(std::abs(sp.Z), std::abs(s.Z)) < 0.35);
        if (near_eq) {
            double dX0=0.0, dY0=0.0, dZ0=0.0, dpX0=0.0, dpY0=0.0, dpZ0=0.0;
            double dX1=0.0, dY1=0.0, dZ1=0.0, dpX1=0.0, dpY1=0.0, dpZ1=0.0;
            ks_rhs(g, sp, dX0, dY0, dZ0, dpX0, dpY0, dpZ0);
            ks_rhs(g, s,  dX1, dY1, dZ1, dpX1, dpY1, dpZ1);
            double al = 0.0;
            bool hit = hermite_first_root(sp.Z, s.Z, dZ0, dZ1, hu, 0.0, al, 8, 8);
            if (!hit && crosses(sp.Z, s.Z)) {
                const double dn = sp.Z - s.Z;
                al = (std::abs(dn) > 1e-12) ? (sp.Z / dn) : 0.5;
                al = clampv(al, 0.0, 1.0);
                hit = true;
            }
            if (hit) {
                const double Xh  = hermite(sp.X,  s.X,  dX0,  dX1,  hu, al);
                const double Yh  = hermite(sp.Y,  s.Y,  dY0,  dY1,  hu, al);
                const double Zh  = hermite(sp.Z,  s.Z,  dZ0,  dZ1,  hu, al);
                const double pXh = hermite(sp.pX, s.pX, dpX0, dpX1, hu, al);
                const double pYh = hermite(sp.pY, s.pY, dpY0, dpY1, hu, al);
                const double pZh = hermite(sp.pZ, s.pZ, dpZ0, dpZ1, hu, al);
                const double pTh = sp.pT;
                double rh_, th_, ph_;
                Spacetime::ks_to_bl(Xh, Yh, Zh, g.a, rh_, th_, ph_);
                if (rh_ >= rin && rh_ <= rout) {
                    const auto pb = covec_ks_to_bl(rh_, th_, ph_, g.a, pXh, pYh, pZh);
                    d_r = rh_;
                    d_g = clampv(disk_g(rh_, pTh, pb[2], g), 0.0, 20.0);
                    d_ph = ph_;
                    best = al;
                    ev = 1;
                }
            }
        }
        const bool hz = ((rprev > rcut) && (rnow <= rcut)) || (rnow <= rcut);
        if (hz) {
            const double dn = rprev - rnow;
            double ah = (std::abs(dn) > 1e-12) ? ((rprev - rcut) / dn) : 0.0;
            ah = clampv(ah, 0.0, 1.0);
            if (ah < best) { best = ah; ev = 2; }
        }
        const bool es = ((rprev < resc) && (rnow >= resc)) || (rnow >= resc);
        if (es) {
            const double dn = rnow - rprev;
            double ae = (std::abs(dn) > 1e-12) ? ((resc - rprev) / dn) : 1.0;
            ae = clampv(ae, 0.0, 1.0);
            if (ae < best) { best = ae; ev = 3; }
        }
        if (ev == 1) { Result R{DISK, d_r, d_g}; R.ph_disk = d_ph; return R; }
        if (ev == 2) { Result R{HOLE, rprev + best * (rnow - rprev), 0.0}; return R; }
        if (ev == 3) {
            const double Xe = sp.X + best * (s.X - sp.X);
            const double Ye = sp.Y + best * (s.Y - sp.Y);
            const double Ze = sp.Z + best * (s.Z - sp.Z);
            double re, te, pe;
            Spacetime::ks_to_bl(Xe, Ye, Ze, g.a, re, te, pe);
            return esc_result(re, te, pe, c.quirks);   // :1650
        }
        rprev = rnow;
    }
    double re, te, pe;
    Spacetime::ks_to_bl(s.X, s.Y, s.Z, g.a, re, te, pe);
    return esc_result(re, te, pe, c.quirks);   // :1659
}

// ---- Jacobi ray bundles (ray_bundle.hpp:56-436) ----
static void hess(const Spacetime& g, double r, double th, double pr, double pth, double pt, double pp,
                 double Hh[4][4]) {   // :56-108
    const double e[4] = {
        1e-5 * (std::abs(r)   + 0.1),
        1e-5,
        1e-5 * (std::abs(pr)  + 1e-3),
        1e-5 * (std::abs(pth) + 1e-3)
    };
    auto ev = [&](double d0, double d1, double d2, double d3) {
        return g.ham(r+d0, th+d1, pr+d2, pth+d3, pt, pp);
    };
    const double H0 = ev(0,0,0,0);
    double v[4][2];
    v[0][0]=ev( e[0],0,0,0); v[0][1]=ev(-e[0],0,0,0);
    v[1][0]=ev(0, e[1],0,0); v[1][1]=ev(0,-e[1],0,0);
    v[2][0]=ev(0,0, e[2],0); v[2][1]=ev(0,0,-e[2],0);
    v[3][0]=ev(0,0,0, e[3]); v[3][1]=ev(0,0,0,-e[3]);
    for (int i = 0; i < 4; ++i) Hh[i][i] = (v[i][0]-2.0*H0+v[i][1])/(e[i]*e[i]);
    auto mixed = [&](int i, int j) -> double {
        double di[4]={}, dj[4]={};
        di[i]=e[i]; dj[j]=e[j];
        double pp_ = ev(di[0]+dj[0],di[1]+dj[1],di[2]+dj[2],di[3]+dj[3]);
        double pm  = ev(di[0]-dj[0],di[1]-dj[1],di[2]-dj[2],di[3]-dj[3]);
        double mp  = ev(-di[0]+dj[0],-di[1]+dj[1],-di[2]+dj[2],-di[3]+dj[3]);
        double mm  = ev(-di[0]-dj[0],-di[1]-dj[1],-di[2]-dj[2],-di[3]-dj[3]);
        return (pp_ - pm - mp + mm) / (4.0*e[i]*e[j]);
    };
    for (int i = 0; i < 4; ++i)
        for (int j = i+1; j < 4; ++j)
            Hh[i][j] = Hh[j][i] = mixed(i, j);
}

struct Bundle { Ray geo; double W[4][2]; };

static void bundle_rhs(const Spacetime& g, const Ray& s, const double W[4][2], double dz[4], double dW[4][2]) {  // :137-164
    double dr, dth, dpr, dpth;
    bl_rhs(g, s.r, s.th, s.pr, s.pth, s.pt, s.pp, dr, dth, dpr, dpth);
    dz[0]=dr; dz[1]=dth; dz[2]=dpr; dz[3]=dpth;
    double Hh[4][4];
    hess(g, s.r, s.th, s.pr, s.pth, s.pt, s.pp, Hh);
    double Mx[4][4];
    for (int j = 0; j < 4; j++) Mx[0][j] =  Hh[2][j];
    for (int j = 0; j < 4; j++) Mx[1][j] =  Hh[3][j];
    for (int j = 0; j < 4; j++) Mx[2][j] = -Hh[0][j];
    for (int j = 0; j < 4; j++) Mx[3][j] = -Hh[1][j];
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 2; k++) {
            dW[i][k] = 0.0;
            for (int j = 0; j < 4; j++) dW[i][k] += Mx[i][j]*W[j][k];
        }
}

static void bundle_rk4(const Spacetime& g, Bundle& bs, double h) {   // :167-214
    const Ray s0 = bs.geo;
    double z1[4], W1[4][2], z2[4], W2[4][2], z3[4], W3[4][2], z4[4], W4[4][2];
    auto shift = [](const Ray& s, const double dz[4], double f, const double W[4][2], const double dW[4][2],
                    Ray& ns, double nW[4][2]) {
        ns = s;
        ns.r  += f*dz[0]; ns.th  += f*dz[1];
        ns.pr += f*dz[2]; ns.pth += f*dz[3];
        for (int i = 0; i < 4; i++) for (int k = 0; k < 2; k++) nW[i][k] = W[i][k]+f*dW[i][k];
    };
    bundle_rhs(g, s0, bs.W, z1, W1);
    Ray s2; double Wb[4][2];
    shift(s0, z1, 0.5*h, bs.W, W1, s2, Wb);
    bundle_rhs(g, s2, Wb, z2, W2);
    Ray s3; double Wc[4][2];
    shift(s0, z2, 0.5*h, bs.W, W2, s3, Wc);
    bundle_rhs(g, s3, Wc, z3, W3);
    Ray s4; double Wd[4][2];
    shift(s0, z3, h, bs.W, W3, s4, Wd);
    bundle_rhs(g, s4, Wd, z4, W4);
    const double f1 = phi_rate(g, s0.r, s0.th, s0.pt, s0.pp);
    const double f2 = phi_rate(g, s2.r, s2.th, s0.pt, s0.pp);
    const double f3 = phi_rate(g, s3.r, s3.th, s0.pt, s0.pp);
    const double f4 = phi_rate(g, s4.r, s4.th, s0.pt, s0.pp);
    bs.geo.r   += h/6.0*(z1[0]+2*z2[0]+2*z3[0]+z4[0]);
    bs.geo.th  += h/6.0*(z1[1]+2*z2[1]+2*z3[1]+z4[1]);
    bs.geo.pr  += h/6.0*(z1[2]+2*z2[2]+2*z3[2]+z4[2]);
    bs.geo.pth += h/6.0*(z1[3]+2*z2[3]+2*z3[3]+z4[3]);
    bs.geo.ph  += h/6.0*(f1 + 2*f2 + 2*f3 + f4);
    for (int i = 0; i < 4; i++) for (int k = 0; k < 2; k++)
        bs.W[i][k] += h/6.0*(W1[i][k]+2*W2[i][k]+2*W3[i][k]+W4[i][k]);
}

static bool bundle_doubling(const Spacetime& g, Bundle& bs, double& h, double tol) {   // :217-263
    const Bundle s0 = bs;
    Bundle A = s0; bundle_rk4(g, A, h);
    Bundle B = s0; bundle_rk4(g, B, 0.5*h); bundle_rk4(g, B, 0.5*h);
    const double err = std::sqrt(
        (A.geo.r   - B.geo.r)   * (A.geo.r   - B.geo.r)   +
        (A.geo.th  - B.geo.th)  * (A.geo.th  - B.geo.th)  +
        (A.geo.pr  - B.geo.pr)  * (A.geo.pr  - B.geo.pr)  +
        (A.geo.pth - B.geo.pth) * (A.geo.pth - B.geo.pth)) / 15.0;
    if (!std::isfinite(err)) {
        bs = s0;
        h = (std::isfinite(h) && h > 1e-10) ? h * 0.5 : 1e-6;
        if (h < 1e-10) h = 1e-10;
        return false;
    }
    const bool ok = (err < tol || h < 1e-10);
    if (ok) {
        bs = B;
        const double sc = (err > 1e-14) ? 0.9 * std::pow(tol/err, 0.2) : 4.0;
        double hn = h * sc;
        if (!std::isfinite(hn)) hn = h;
        h = hn;
        if (h > 100.0) h = 100.0;
        if (h < 1e-10) h = 1e-10;
    } else {
        bs = s0;
        const double half = h * 0.5;
        double hn = h * 0.9 * std::pow(tol/err, 0.25);
        if (!std::isfinite(hn)) hn = half;
        h = hn;
        if (h > half) h = half;
        if (h < 1e-10) h = 1e-10;
    }
    return ok;
}

static Result trace_bundle(int px, int py, const Cam& cam, const Spacetime& g, double rin, double rout,
                           double resc, double rh, const Controls& c, double ox, double oy, Counters* cnt) {
    // ray_bundle.hpp:301-436; result mapped to the GeoPixel fill of main.cpp:2102-2111
    double al, be;
    cam.screen_angles((double)px + ox, (double)py + oy, al, be);
    Bundle bs;
    bs.geo = cam.shoot(al, be, g);
    {   // init_bundle  :278-298
        const double eps = cam.fov / cam.W * 0.5;
        const Ray ap = cam.shoot(al+eps, be, g), am = cam.shoot(al-eps, be, g);
        const Ray bp = cam.shoot(al, be+eps, g), bm = cam.shoot(al, be-eps, g);
        bs.W[0][0] = (ap.r   - am.r)   / (2.0*eps);
        bs.W[1][0] = (ap.th  - am.th)  / (2.0*eps);
        bs.W[2][0] = (ap.pr  - am.pr)  / (2.0*eps);
        bs.W[3][0] = (ap.pth - am.pth) / (2.0*eps);
        bs.W[0][1] = (bp.r   - bm.r)   / (2.0*eps);
        bs.W[1][1] = (bp.th  - bm.th)  / (2.0*eps);
        bs.W[2][1] = (bp.pr  - bm.pr)  / (2.0*eps);
        bs.W[3][1] = (bp.pth - bm.pth) / (2.0*eps);
    }
    const double rcut = rh * 1.03;
    double h = std::max(c.step_init, 1e-10);
    const int nmax = std::max(1, c.max_steps);
    Result none{ESC, 0.0, 1.0};   // BundleResult{} : no hit, r 0, g 1, magnif 1, angles 0
    for (int it = 0; it < nmax; ++it) {
        const Bundle bp = bs;
        double hu = h;
        int rej = 0;
        while (true) {
            hu = h;
            if (bundle_doubling(g, bs, h, c.tol)) { if (cnt) cnt->accepted++; break; }
            if (cnt) cnt->rejected++;
            if (!std::isfinite(h) || ++rej > 64) return none;
        }
        const double r = bs.geo.r;
        double best = 2.0;
        int ev = 0;
        double d_r = 0.0, d_g = 1.0, d_det = 1.0, d_ph = 0.0;
        const double q0 = bp.geo.th - PI/2.0, q1 = bs.geo.th - PI/2.0;
        const bool near_eq = crosses(q0, q1) || (std::min(std::abs(q0), std::abs(q1)) < 0.35);
        if (near_eq) {
            double dr0,dt0,dpr0,dpt0, dr1,dt1,dpr1,dpt1;
            bl_rhs(g, bp.geo.r, bp.geo.th, bp.geo.pr, bp.geo.pth, bp.geo.pt, bp.geo.pp, dr0,dt0,dpr0,dpt0);
            bl_rhs(g, bs.geo.r, bs.geo.th, bs.geo.pr, bs.geo.pth, bs.geo.pt, bs.geo.pp, dr1,dt1,dpr1,dpt1);
            double a = 0.0;
            const bool found = hermite_first_root(bp.geo.th, bs.geo.th, dt0, dt1, hu, PI/2.0, a, 8, 8);
            // the reference skips the horizon and escape tests of this step (:360-364)
            if (!found && (c.quirks & QK_BUNDLE_ROOT_SKIP)) continue;
            const double rhit = found ? hermite(bp.geo.r, bs.geo.r, dr0, dr1, hu, a) : 0.0;
            if (found && rhit >= rin && rhit <= rout) {
                const double Om = g.omega_k(rhit);
                const double b = -bs.geo.pp / (-bs.geo.pt);
                double L[4][4];
                g.lower_bl(rhit, PI/2.0, L);
                const double d2 = -(L[0][0]+2.0*L[0][3]*Om+L[3][3]*Om*Om);
                const double dv = 1.0 - Om*b;
                double red = (d2 > 0.0 && std::abs(dv) > 1e-10) ? std::sqrt(d2)/dv : 1.0;
                red = red < 0.0 ? 0.0 : red > 20.0 ? 20.0 : red;
                const double J00 = bp.W[0][0] + a*(bs.W[0][0] - bp.W[0][0]);
                const double J01 = bp.W[0][1] + a*(bs.W[0][1] - bp.W[0][1]);
                const double J10 = bp.W[1][0] + a*(bs.W[1][0] - bp.W[1][0]);
                const double J11 = bp.W[1][1] + a*(bs.W[1][1] - bp.W[1][1]);
                double det = std::abs(J00*J11 - J01*J10);
                det = det < 1e-12 ? 1e-12 : det;
                d_r = rhit; d_g = red; d_det = det;
                d_ph = bp.geo.ph + a * (bs.geo.ph - bp.geo.ph);
                best = a;
                ev = 1;
            }
        }
        const bool hz = ((bp.geo.r > rcut) && (r <= rcut)) || (r <= rcut);
        if (hz) {
            const double dn = bp.geo.r - r;
            double ah = (std::abs(dn) > 1e-12) ? ((bp.geo.r - rcut) / dn) : 0.0;
            ah = ah < 0.0 ? 0.0 : ah > 1.0 ? 1.0 : ah;
            if (ah < best) { best = ah; ev = 2; }
        }
        const bool es = ((bp.geo.r < resc) && (r >= resc)) || (r >= resc);
        if (es) {
            const double dn = r - bp.geo.r;
            double ae = (std::abs(dn) > 1e-12) ? ((resc - bp.geo.r) / dn) : 1.0;
            ae = ae < 0.0 ? 0.0 : ae > 1.0 ? 1.0 : ae;
            if (ae < best) { best = ae; ev = 3; }
        }
        if (ev == 1) { Result R{DISK, d_r, d_g}; R.magnif = d_det; R.ph_disk = d_ph; return R; }
        if (ev == 2) {   // the reference returns the empty BundleResult (:426-428): sky at the pole
            if (c.quirks & QK_BUNDLE_HORIZON_ESC) return none;
            return Result{HOLE, 0.0, 0.0};
        }
        if (ev == 3) {
            Result R{ESC, resc, 1.0};
            R.th_esc = bp.geo.th + best * (bs.geo.th - bp.geo.th);
            R.ph_esc = bp.geo.ph + best * (bs.geo.ph - bp.geo.ph);
            return R;
        }
    }
    Result R{ESC, bs.geo.r, 1.0};
    return R;
}

// ---------------------------------------------------------------------------
//  Shading (main.cpp:140-171, 1663-1687, 1799-1866)
// ---------------------------------------------------------------------------
struct RGB8 { uint8_t r, g, b; };

static RGB8 black_body(double T) {   // :1663-1674
    T = clampv(T, 800.0, 4e4);
    double t = std::log10(T/800.0)/std::log10(4e4/800.0);
    double R, G, B;
    if (t < 0.25)      { R = 1; G = t/0.25*0.4; B = 0; }
    else if (t < 0.5)  { double f = (t-0.25)/0.25; R = 1; G = 0.4+f*0.4; B = f*0.3; }
    else if (t < 0.75) { double f = (t-0.5)/0.25;  R = 1; G = 0.8+f*0.2; B = 0.3+f*0.5; }
    else               { double f = (t-0.75)/0.25; R = 1-f*0.2; G = 1; B = 0.8+f*0.2; }
    return {(uint8_t)(clampv(R,0.0,1.0)*255), (uint8_t)(clampv(G,0.0,1.0)*255), (uint8_t)(clampv(B,0.0,1.0)*255)};
}
static double pt_emiss(double r, double rI) {   // :1676-1680
    if (r <= rI) return 0.0;
    double x = std::sqrt(rI/r);
    return (1.0-x)/(r*r*r);
}
static double reinhard(double x, double exposure, double gamma) {   // :1683-1687
    x = x * exposure;
    x = x / (1.0 + x);
    return std::pow(clampv(x, 0.0, 1.0), 1.0/gamma);
}
struct Colour { double exposure, gamma, temp_scale; };

static RGB8 disk_rgb(double r, double red, double mag, double M, double rI, const Colour& cp) {   // :1799-1814
    double T = 6500.0 * cp.temp_scale * std::sqrt(6.0*M/r) * clampv(red, 0.2, 5.0);
    double pn = pt_emiss(r, rI);
    double rpk = 3.0*rI;
    double ppk = pt_emiss(rpk, rI);
    double I = (ppk > 0.0) ? pn/ppk : 0.0;
    const double rc = clampv(red, 0.1, 10.0);
    const double lift = 1.0 + 0.85 * clampv(1.0 - rc, 0.0, 1.0);
    I *= std::pow(rc, 4.0) * lift;
    I *= clampv(1.0/mag, 0.05, 5.0);
    auto c = black_body(T);
    return {(uint8_t)(reinhard(c.r/255.0*I, cp.exposure, cp.gamma)*255),
            (uint8_t)(reinhard(c.g/255.0*I, cp.exposure, cp.gamma)*255),
            (uint8_t)(reinhard(c.b/255.0*I, cp.exposure, cp.gamma)*255)};
}

// Interstellar palette (main.cpp:1689-1797).  HSV -> RGB8 with round-half-up bytes (:1690-1708).
static RGB8 hsv8(double hue, double sat, double v) {
    hue -= std::floor(hue);
    const double x = hue * 6.0;
    const int sector = (int)x;
    const double fr = x - sector;
    const double lo = v * (1.0 - sat), fall = v * (1.0 - sat * fr), rise = v * (1.0 - sat * (1.0 - fr));
    double c[3];
    switch (sector % 6) {
        case 0:  c[0] = v;    c[1] = rise; c[2] = lo;   break;
        case 1:  c[0] = fall; c[1] = v;    c[2] = lo;   break;
        case 2:  c[0] = lo;   c[1] = v;    c[2] = rise; break;
        case 3:  c[0] = lo;   c[1] = fall; c[2] = v;    break;
        case 4:  c[0] = rise; c[1] = lo;   c[2] = v;    break;
        default: c[0] = v;    c[1] = lo;   c[2] = fall; break;
    }
    return {(uint8_t)(c[0]*255 + 0.5), (uint8_t)(c[1]*255 + 0.5), (uint8_t)(c[2]*255 + 0.5)};
}
// fract(43758.5453 * sin(id * a + b)) (:1711-1715)
static double cell_fract(int id, double a, double b) {
    const double v = std::sin((double)id * a + b) * 43758.5453;
    return v - std::floor(v);
}
struct Cells { int rings, sectors; double sigma, hue_offset, r_in, r_out; };

// (:1722-1797) Gaussian-weighted mean colour of the ring x sector cells around (r, phi):
// weight exp(-d^2 / 2 sigma^2) in cell units, sector distance wrapped, rings clamped at
// the edges, cells with a third hash below 0.3 left out; then the blackbody palette's
// emissivity / redshift / magnification factor and tone map.
static RGB8 disk_rgb_cells(double r, double phi, double red, double mag, double rI, const Cells& k,
                           const Colour& cp) {
    const int nr = k.rings > 0 ? k.rings : 1, ns = k.sectors > 0 ? k.sectors : 1;
    const double sg = k.sigma > 0 ? k.sigma : 0.5;
    const double two_pi = 2.0*PI;
    const double u = clampv((r - k.r_in) / (k.r_out - k.r_in), 0.0, 1.0) * nr;
    const double v = (phi - std::floor(phi / two_pi) * two_pi) / two_pi * ns;
    const double q = 1.0 / (2.0 * sg * sg);
    const int reach = (int)std::ceil(3.0 * sg) + 1;
    double sum[3] = {0.0, 0.0, 0.0}, wsum = 0.0;
    for (int i = -reach; i <= nr + reach; ++i) {
        const double du = u - (double)i;
        if (du * du * q > 16.0) continue;
        const int ring = std::min(std::max(i, 0), nr - 1);
        for (int j = -reach; j <= ns + reach; ++j) {
            double dv = v - (double)j;
            dv = dv - std::floor((dv + ns * 0.5) / ns) * ns;
            const double w = std::exp(-(du * du + dv * dv) * q);
            if (w < 1e-4) continue;
            const int id = ring * ns + ((j % ns) + ns) % ns;
            const double c1 = cell_fract(id, 127.1, 311.7), c2 = cell_fract(id, 269.5, 183.3);
            if (cell_fract(id, 419.2, 371.9) < 0.30) continue;
            const RGB8 c = hsv8(std::fmod(c1 * 0.14 + k.hue_offset, 1.0), 0.70 + c2 * 0.30, 0.50 + c1 * 0.50);
            sum[0] += w * c.r; sum[1] += w * c.g; sum[2] += w * c.b;
            wsum += w;
        }
    }
    if (wsum < 1e-6) return {0, 0, 0};
    const double pk = pt_emiss(3.0*rI, rI);
    double I = (pk > 0.0) ? pt_emiss(r, rI)/pk : 0.0;
    const double rc = clampv(red, 0.1, 10.0);
    I *= std::pow(rc, 4.0) * (1.0 + 0.85 * clampv(1.0 - rc, 0.0, 1.0));
    I *= clampv(1.0/mag, 0.05, 5.0);
    const double nrm = wsum * 255.0;
    return {(uint8_t)(reinhard(sum[0] / nrm * I, cp.exposure, cp.gamma)*255),
            (uint8_t)(reinhard(sum[1] / nrm * I, cp.exposure, cp.gamma)*255),
            (uint8_t)(reinhard(sum[2] / nrm * I, cp.exposure, cp.gamma)*255)};
}

static RGB8 sky(const uint8_t* px, int w, int h, double th, double ph) {   // :150-170
    if (!px || w <= 0 || h <= 0) return {0,0,0};
    double pw = std::fmod(ph, 2.0*PI);
    if (pw < 0) pw += 2.0*PI;
    double uf = pw / (2.0*PI) * (w-1);
    double vf = th / PI * (h-1);
    int u0 = (int)uf, u1 = (w > 1) ? ((u0 + 1) % w) : 0;
    int v0 = (int)vf, v1 = std::min(v0+1, h-1);
    double fu = uf-u0, fv = vf-v0;
    auto at = [&](int u, int v) -> const uint8_t* { return px + 3*(clampv(v,0,h-1)*w + clampv(u,0,w-1)); };
    auto lp = [](double a, double b, double t) { return a + t*(b-a); };
    double R = lp(lp(at(u0,v0)[0],at(u1,v0)[0],fu), lp(at(u0,v1)[0],at(u1,v1)[0],fu), fv);
    double G = lp(lp(at(u0,v0)[1],at(u1,v0)[1],fu), lp(at(u0,v1)[1],at(u1,v1)[1],fu), fv);
    double B = lp(lp(at(u0,v0)[2],at(u1,v0)[2],fu), lp(at(u0,v1)[2],at(u1,v1)[2],fu), fv);
    return {(uint8_t)R, (uint8_t)G, (uint8_t)B};
}

}  // namespace kto

// ============================================================================
//  C ABI used by the tests (ctypes)
// ============================================================================
using namespace kto;

extern "C" {

// 28-byte record, identical to main.cpp:84-94
struct kto_geopixel {
    uint8_t outcome; uint8_t pad[3];
    float r, redshift, magnif, phi_disk, theta_esc, phi_esc;
};
static_assert(sizeof(kto_geopixel) == 28, "GeoPixel layout");

struct kto_frame {
    int32_t width, height;
    int32_t chart;        // 0 = BL, 1 = KS (requested; KS with Lambda != 0 runs BL, main.cpp:2034-2041)
    int32_t integrator;   // 0 = RK4 step doubling, 1 = DOPRI5
    int32_t bundles;      // 1 = Jacobi ray bundles (BL only, main.cpp:2042-2049)
    int32_t max_steps;
    double  step_init, tol;
    double  M, a, Q, Lambda;
    double  theta_deg, phi_deg, r_obs, fov_deg, disk_out;
    double  offset_x, offset_y;
    uint32_t quirks;      // QK_* mask (QK_REFERENCE: every reference behaviour)
    int32_t pad;
};

struct kto_frame_info {     // per-frame scalars (KGeoMeta subset + effective chart)
    double r_horizon, r_isco, r_disk_in, r_disk_out, r_escape;
    int32_t eff_chart, eff_bundles;
};

int kto_version(void) { return 1; }

void kto_frame_info_get(const kto_frame* f, kto_frame_info* out) {
    Spacetime g{f->M, f->a, f->Q, f->Lambda};
    out->r_horizon = g.horizon();
    out->r_isco = g.isco();
    out->r_disk_in = out->r_isco;
    out->r_disk_out = f->disk_out;
    out->r_escape = f->r_obs * 1.05;
    int chart = f->chart;
    if (chart == 1 && std::abs(f->Lambda) > 1e-15) chart = 0;
    if (f->bundles && chart == 1) chart = 0;
    out->eff_chart = chart;
    out->eff_bundles = f->bundles ? 1 : 0;
}

static void fill_pixel(kto_geopixel& p, const Result& R, bool bundle) {   // main.cpp:2102-2147
    if (bundle) {
        const bool hit = (R.fate == DISK);
        p.outcome = hit ? 1 : (R.fate == HOLE ? 2 : 0);   // HOLE only with QK_BUNDLE_HORIZON_ESC cleared
        p.r = hit ? (float)R.r : 0.0f;
        p.redshift = (float)R.g;
        p.magnif = (float)R.magnif;
        p.phi_disk = (float)R.ph_disk;
        p.theta_esc = (float)R.th_esc;
        p.phi_esc = (float)R.ph_esc;
    } else {
        p.outcome = (uint8_t)R.fate;
        p.r = (float)R.r;
        p.redshift = (float)R.g;
        p.magnif = 1.0f;
        p.phi_disk = (float)R.ph_disk;
        p.theta_esc = (float)R.th_esc;
        p.phi_esc = (float)R.ph_esc;
    }
    p.pad[0] = p.pad[1] = p.pad[2] = 0;
}

// Trace an explicit list of pixels.  n_threads <= 0: OpenMP default.
// Optional per-ray outputs (may be NULL): accepted / rejected step counts and
// the full-precision (double) result fields r, g, theta_esc, phi_esc.
int kto_trace_pixels(const kto_frame* f, const int32_t* px, const int32_t* py, int64_t n,
                     kto_geopixel* out, int64_t* accepted, int64_t* rejected, double* r64,
                     double* g64, double* th64, double* ph64, int n_threads) {
    if (!f || !px || !py || !out || n < 0) return -1;
    Spacetime g{f->M, f->a, f->Q, f->Lambda};
    kto_frame_info fi;
    kto_frame_info_get(f, &fi);
    const Cam cam(f->r_obs, f->theta_deg, f->phi_deg, f->fov_deg, f->width, f->height);
    const Controls c{std::max(1, f->max_steps), std::max(1e-10, f->step_init), std::max(1e-10, f->tol), f->quirks};
    const double rh = fi.r_horizon;   // hoisted: r_horizon() is a pure function of (M,a,Q,Lambda)
#ifdef _OPENMP
    const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
#endif
    for (int64_t i = 0; i < n; ++i) {
        Counters cnt;
        Result R;
        if (fi.eff_bundles) {
            R = trace_bundle(px[i], py[i], cam, g, fi.r_disk_in, fi.r_disk_out, fi.r_escape, rh, c,
                             f->offset_x, f->offset_y, &cnt);
        } else {
            Ray s = cam.pixel(px[i], py[i], g, f->offset_x, f->offset_y);
            if (fi.eff_chart == 1)
                R = trace_ks(s, g, fi.r_disk_in, fi.r_disk_out, fi.r_escape, rh, f->integrator, c, &cnt);
            else
                R = trace_bl(s, g, fi.r_disk_in, fi.r_disk_out, fi.r_escape, rh, f->integrator, c, &cnt);
        }
        fill_pixel(out[i], R, fi.eff_bundles != 0);
        if (accepted) accepted[i] = cnt.accepted;
        if (rejected) rejected[i] = cnt.rejected;
        if (r64) r64[i] = (fi.eff_bundles && R.fate != DISK) ? 0.0 : R.r;
        if (g64) g64[i] = R.g;
        if (th64) th64[i] = R.th_esc;
        if (ph64) ph64[i] = R.ph_esc;
    }
#ifndef _OPENMP
    (void)n_threads;
#endif
    return 0;
}

// Whole frame, row-major GeoPixel (trace_geodesics, main.cpp:2014-2206)
int kto_trace_frame(const kto_frame* f, kto_geopixel* out, int n_threads) {
    if (!f || !out || f->width <= 0 || f->height <= 0) return -1;
    const int64_t n = (int64_t)f->width * f->height;
    std::vector<int32_t> px(n), py(n);
    for (int64_t i = 0; i < n; ++i) { px[i] = (int32_t)(i % f->width); py[i] = (int32_t)(i / f->width); }
    return kto_trace_pixels(f, px.data(), py.data(), n, out, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, n_threads);
}

// Phase 2: GeoPixel -> RGB8 (colorize_buffer, blackbody palette)
int kto_colorize(const kto_geopixel* geo, int64_t n, double M, double r_isco, double exposure, double gamma,
                 double temp_scale, const uint8_t* bg, int bg_w, int bg_h, uint8_t* rgb_out) {
    if (!geo || !rgb_out || n < 0) return -1;
    const Colour cp{exposure, gamma, temp_scale};
    for (int64_t i = 0; i < n; ++i) {
        RGB8 c{0, 0, 0};
        const kto_geopixel& p = geo[i];
        if (p.outcome == 1) c = disk_rgb(p.r, p.redshift, p.magnif, M, r_isco, cp);
        else if ((p.outcome == 0 || p.outcome == 3) && bg) c = sky(bg, bg_w, bg_h, p.theta_esc, p.phi_esc);
        rgb_out[3*i+0] = c.r; rgb_out[3*i+1] = c.g; rgb_out[3*i+2] = c.b;
    }
    return 0;
}

// Phase 2 with the Interstellar palette (colorize_buffer, main.cpp:1839-1845): disk pixels
// take the cell colours over [r_disk_in, r_disk_out]; sky and black pixels as above.
int kto_colorize_interstellar(const kto_geopixel* geo, int64_t n, double M, double r_isco, double r_disk_in,
                              double r_disk_out, int rings, int sectors, double sigma, double hue_offset,
                              double exposure, double gamma, const uint8_t* bg, int bg_w, int bg_h, uint8_t* rgb_out) {
    if (!geo || !rgb_out || n < 0) return -1;
    (void)M;
    const Colour cp{exposure, gamma, 1.0};
    const Cells k{rings, sectors, sigma, hue_offset, r_disk_in, r_disk_out};
    for (int64_t i = 0; i < n; ++i) {
        RGB8 c{0, 0, 0};
        const kto_geopixel& p = geo[i];
        if (p.outcome == 1) c = disk_rgb_cells(p.r, p.phi_disk, p.redshift, p.magnif, r_isco, k, cp);
        else if ((p.outcome == 0 || p.outcome == 3) && bg) c = sky(bg, bg_w, bg_h, p.theta_esc, p.phi_esc);
        rgb_out[3*i+0] = c.r; rgb_out[3*i+1] = c.g; rgb_out[3*i+2] = c.b;
    }
    return 0;
}

// camera-spp offsets (main.cpp:2241-2284)
int kto_spp_offsets(int spp, double* xy_out) {
    std::vector<std::pair<double, double>> o;
    if (spp <= 1) { o.emplace_back(0.0, 0.0); }
    else if (spp == 2) { o.emplace_back(-0.35, 0.0); o.emplace_back(+0.35, 0.0); }
    else {
        if ((spp % 2) == 1) o.emplace_back(0.0, 0.0);
        const std::pair<double, double> seed[] = {{0.35, 0.00}, {0.00, 0.35}, {0.25, 0.25}, {0.45, 0.20},
                                                  {0.20, 0.45}, {0.40, 0.00}, {0.00, 0.40}};
        for (const auto& p : seed) {
            if ((int)o.size() + 2 > spp) break;
            o.emplace_back(+p.first, +p.second);
            o.emplace_back(-p.first, -p.second);
        }
        if ((int)o.size() < spp) {
            constexpr double golden = 2.39996322972865332;
            int k = 0;
            while ((int)o.size() < spp) {
                const double rr = 0.49 * std::sqrt((double)(k + 1) / (double)(spp + 1));
                const double aa = golden * (double)(k + 1);
                const double x = rr * std::cos(aa), y = rr * std::sin(aa);
                o.emplace_back(x, y);
                if ((int)o.size() < spp) o.emplace_back(-x, -y);
                ++k;
            }
        }
        o.resize((size_t)spp);
    }
    const int n = std::max(1, spp);
    for (int i = 0; i < n; ++i) { xy_out[2*i] = o[i].first; xy_out[2*i+1] = o[i].second; }
    return n;
}

// multi-sample resolve in linear space (main.cpp:2286-2321)
int kto_spp_resolve(const uint8_t* passes, int n_pass, int64_t n_pix, double gamma, uint8_t* out) {
    const double ge = std::max(gamma, 1e-6);
    for (int64_t i = 0; i < n_pix; ++i) {
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (int s = 0; s < n_pass; ++s) {
                const double x = clampv((double)passes[((int64_t)s * n_pix + i) * 3 + c] / 255.0, 0.0, 1.0);
                acc += std::pow(x, ge);
            }
            const double l = clampv(acc * (1.0 / (double)n_pass), 0.0, 1.0);
            out[i*3 + c] = (uint8_t)std::lround(255.0 * std::pow(l, 1.0 / ge));
        }
    }
    return 0;
}

// ---- function-level entry points (known-answer and invariant tests) ----
void kto_metric_upper_bl(const double* mp, double r, double th, double* out16) {
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    double u[4][4]; g.upper_bl(r, th, u);
    std::memcpy(out16, u, sizeof(u));
}
void kto_metric_lower_bl(const double* mp, double r, double th, double* out16) {
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    double u[4][4]; g.lower_bl(r, th, u);
    std::memcpy(out16, u, sizeof(u));
}
void kto_metric_upper_ks(const double* mp, double X, double Y, double Z, double* out16) {
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    double u[4][4]; g.upper_ks(X, Y, Z, u);
    std::memcpy(out16, u, sizeof(u));
}
double kto_hamiltonian_bl(const double* mp, const double* s6) {   // r, th, pr, pth, pt, pp
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    return g.ham(s6[0], s6[1], s6[2], s6[3], s6[4], s6[5]);
}
void kto_rhs_bl(const double* mp, const double* s6, double* out5) {   // dr, dth, dpr, dpth, dphi
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    bl_rhs(g, s6[0], s6[1], s6[2], s6[3], s6[4], s6[5], out5[0], out5[1], out5[2], out5[3]);
    out5[4] = phi_rate(g, s6[0], s6[1], s6[4], s6[5]);
}
double kto_r_horizon(const double* mp) { Spacetime g{mp[0], mp[1], mp[2], mp[3]}; return g.horizon(); }
double kto_r_isco(const double* mp) { Spacetime g{mp[0], mp[1], mp[2], mp[3]}; return g.isco(); }
double kto_omega_k(const double* mp, double r) { Spacetime g{mp[0], mp[1], mp[2], mp[3]}; return g.omega_k(r); }
double kto_disk_redshift(const double* mp, double r, double pt, double pp) {
    Spacetime g{mp[0], mp[1], mp[2], mp[3]}; return disk_g(r, pt, pp, g);
}
// camera: cam6 = r_obs, theta_deg, phi_deg, fov_deg, W, H; out7 = r, th, ph, pr, pth, pt, pp
void kto_camera_pixel(const double* mp, const double* cam6, double px, double py, double* out7) {
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    const Cam c(cam6[0], cam6[1], cam6[2], cam6[3], (int)cam6[4], (int)cam6[5]);
    double al, be;
    c.screen_angles(px, py, al, be);
    const Ray s = c.shoot(al, be, g);
    out7[0] = s.r; out7[1] = s.th; out7[2] = s.ph; out7[3] = s.pr; out7[4] = s.pth; out7[5] = s.pt; out7[6] = s.pp;
}
void kto_camera_angles(const double* mp, const double* cam6, double al, double be, double* out7) {
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    const Cam c(cam6[0], cam6[1], cam6[2], cam6[3], (int)cam6[4], (int)cam6[5]);
    const Ray s = c.shoot(al, be, g);
    out7[0] = s.r; out7[1] = s.th; out7[2] = s.ph; out7[3] = s.pr; out7[4] = s.pth; out7[5] = s.pt; out7[6] = s.pp;
}
// one adaptive attempt; state7 in/out, h in/out, fsal4 in/out (NaN = cold). returns 1 if accepted
int kto_adaptive_step(const double* mp, int method, double* state7, double* h, double* fsal4, double tol) {
    Spacetime g{mp[0], mp[1], mp[2], mp[3]};
    Ray s{state7[0], state7[1], state7[2], state7[3], state7[4], state7[5], state7[6]};
    K4 f; for (int i = 0; i < 4; ++i) f.v[i] = fsal4 ? fsal4[i] : std::numeric_limits<double>::quiet_NaN();
    const bool ok = take_step(g, s, *h, method, f, tol);
    state7[0] = s.r; state7[1] = s.th; state7[2] = s.ph; state7[3] = s.pr; state7[4] = s.pth;
    state7[5] = s.pt; state7[6] = s.pp;
    if (fsal4) for (int i = 0; i < 4; ++i) fsal4[i] = f.v[i];
    return ok ? 1 : 0;
}
void kto_bl_to_ks(double r, double th, double ph, double a, double* xyz) {
    Spacetime::bl_to_ks(r, th, ph, a, xyz[0], xyz[1], xyz[2]);
}
void kto_ks_to_bl(double X, double Y, double Z, double a, double* rtp) {
    Spacetime::ks_to_bl(X, Y, Z, a, rtp[0], rtp[1], rtp[2]);
}
double kto_r_ks(double X, double Y, double Z, double a) { return Spacetime::ks_radius(X, Y, Z, a); }
int kto_hermite_first_root(double y0, double y1, double d0, double d1, double h, double target, double* alpha) {
    return hermite_first_root(y0, y1, d0, d1, h, target, *alpha, 8, 8) ? 1 : 0;
}
double kto_hermite(double y0, double y1, double d0, double d1, double h, double al) {
    return hermite(y0, y1, d0, d1, h, al);
}
void kto_blackbody(double T, uint8_t* rgb) { RGB8 c = black_body(T); rgb[0] = c.r; rgb[1] = c.g; rgb[2] = c.b; }

}  // extern "C"
