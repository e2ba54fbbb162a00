// Host build of the product's per-ray tracer (kerrtracecpp2_amd/csrc/kt_trace.hpp),
// for CPU-only tests of the kernel logic against the oracle.  Test
// infrastructure: not part of the product path, never loaded by it.
#include <cmath>
#include <type_traits>
#include <cstring>
#include <vector>

#include "../../include/kt_hip.h"
#include "../../kerrtracecpp2_amd/csrc/kt_adjoint.hpp"
#include "../../kerrtracecpp2_amd/csrc/kt_bundle.hpp"
#include "../../kerrtracecpp2_amd/csrc/kt_setup.hpp"
#include "../../kerrtracecpp2_amd/csrc/kt_shade.hpp"
#include "../../kerrtracecpp2_amd/csrc/kt_trace.hpp"

using namespace kt;

static void fill_frame(const kt_frame& f, const kt_opts& o, FrameDev& F) {
    std::memset(&F, 0, sizeof(F));
    F.M = f.M; F.a = f.a; F.Q = f.Q; F.Lambda = f.Lambda;
    F.r_obs = f.r_obs;
    F.th_obs = f.theta_deg * kPi / 180.0; F.ph_obs = f.phi_deg * kPi / 180.0; F.fov = f.fov_deg * kPi / 180.0;
    F.disk_out = f.disk_out; F.off_x = f.offset_x; F.off_y = f.offset_y;
    F.W = f.width; F.H = f.height;
    int ch = o.chart;
    if (ch == KT_CHART_KS && (o.quirks & KT_QUIRK_KS_LAMBDA_TO_BL) && std::fabs(f.Lambda) > 1e-15) ch = KT_CHART_BL;
    F.chart = ch;
    F.ks_ref_init = (ch == KT_CHART_KS && (o.quirks & KT_QUIRK_KS_INIT_NO_SHIFT) && std::fabs(f.Lambda) <= 1e-15);
    frame_knds(F);
    frame_camera_consts(F);
    F.r_h = horizon_seq(F);
    F.isco_closed = isco_bardeen(F.M, F.a);
    F.r_isco = isco_seq(F, F.r_h);
    F.r_in = F.r_isco; F.r_out = F.disk_out; F.r_esc = F.r_obs * 1.05; F.r_cut = F.r_h * 1.03;
}

template <int CHART, int METHOD, int PREC, bool PHI>
static void run(const FrameDev& F, const RunOpts& O, const int32_t* px, const int32_t* py, int64_t n,
                kt_geopixel* out, int64_t* acc, int64_t* rhs32, int64_t* rhs64) {
    const Knds<float> Pf = knds_cast<float>(F.P);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        RayState S;
        RayOut o;
        ray_init<CHART, PREC>(F, O, px[i], py[i], S);
        for (;;) {
            const bool use64 = (PREC == PREC_FP64) ? true : wants_fp64(F, O, S);
            if (ray_advance<CHART, METHOD, PREC, PHI>(F, F.P, Pf, O, use64, S, o)) break;
        }
        uint32_t w[7];
        pack_geopixel(o, w);
        std::memcpy(&out[i], w, 28);
        if (acc) acc[i] = S.n_acc;
        if (rhs32) rhs32[i] = S.rhs32;
        if (rhs64) rhs64[i] = S.rhs64;
    }
}

template <int CHART>
static void run_chart(int method, int prec, bool phi, const FrameDev& F, const RunOpts& O, const int32_t* px,
                      const int32_t* py, int64_t n, kt_geopixel* out, int64_t* acc, int64_t* r32, int64_t* r64) {
    if (method == METHOD_RK4) {
        if (prec) run<CHART, METHOD_RK4, PREC_MIXED, true>(F, O, px, py, n, out, acc, r32, r64);
        else run<CHART, METHOD_RK4, PREC_FP64, true>(F, O, px, py, n, out, acc, r32, r64);
    } else if (phi) {
        if (prec) run<CHART, METHOD_DP5, PREC_MIXED, true>(F, O, px, py, n, out, acc, r32, r64);
        else run<CHART, METHOD_DP5, PREC_FP64, true>(F, O, px, py, n, out, acc, r32, r64);
    } else {
        if (prec) run<CHART, METHOD_DP5, PREC_MIXED, false>(F, O, px, py, n, out, acc, r32, r64);
        else run<CHART, METHOD_DP5, PREC_FP64, false>(F, O, px, py, n, out, acc, r32, r64);
    }
}

extern "C" int ktc_trace_pixels(const kt_frame* f, const kt_opts* o, const int32_t* px, const int32_t* py,
                                int64_t n, kt_geopixel* out, int64_t* acc, int64_t* rhs32, int64_t* rhs64,
                                double* info5) {
    FrameDev F;
    fill_frame(*f, *o, F);
    RunOpts O;
    O.max_steps = o->max_steps < 1 ? 1 : o->max_steps;
    O.quirks = o->quirks;
    O.step_init = o->step_init < 1e-10 ? 1e-10 : o->step_init;
    O.rtol = o->rtol; O.atol = o->atol;
    O.isl_noise = o->mixed_noise > 0.0 ? o->mixed_noise : kMixedNoiseDefault;
    O.isl_r = o->mixed_r_island > 0.0 ? o->mixed_r_island : kMixedRIslandDefault;
    O.isl_hold = o->mixed_hold > 0 ? o->mixed_hold : kMixedHoldDefault;
    if (info5) { info5[0] = F.r_h; info5[1] = F.r_isco; info5[2] = F.r_esc; info5[3] = F.chart; info5[4] = F.ks_ref_init; }
    const bool phi = !(o->integrator == KT_INTEGRATOR_DOPRI5 && (o->quirks & KT_QUIRK_DOPRI5_PHI_FROZEN));
    if (F.chart == CHART_KS) run_chart<CHART_KS>(o->integrator, o->precision, phi, F, O, px, py, n, out, acc, rhs32, rhs64);
    else run_chart<CHART_BL>(o->integrator, o->precision, phi, F, O, px, py, n, out, acc, rhs32, rhs64);
    return 0;
}

extern "C" int ktc_shade(const kt_geopixel* geo, int64_t n, double M, double r_isco, double exposure, double gamma,
                         double temp_scale, const uint8_t* bg, int bw, int bh, uint32_t* rgba) {
    ShadeParams p;
    std::memset(&p, 0, sizeof(p));
    p.M = M; p.r_isco = r_isco; p.exposure = exposure; p.gamma = gamma; p.temp_scale = temp_scale;
    for (int64_t i = 0; i < n; ++i)
        rgba[i] = shade_pixel(geo[i].outcome, geo[i].r, geo[i].redshift, geo[i].magnif, geo[i].phi_disk,
                              geo[i].theta_esc, geo[i].phi_esc, p, bg, bw, bh);
    return 0;
}

// Interstellar palette on the host (same code as the shade kernel)
extern "C" int ktc_shade_interstellar(const kt_geopixel* geo, int64_t n, double M, double r_isco, double r_in,
                                      double r_out, int rings, int sectors, double sigma, double hue, double exposure,
                                      double gamma, const uint8_t* bg, int bw, int bh, uint32_t* rgba) {
    ShadeParams p;
    std::memset(&p, 0, sizeof(p));
    p.M = M; p.r_isco = r_isco; p.exposure = exposure; p.gamma = gamma; p.temp_scale = 1.0;
    p.palette = 1; p.rings = rings; p.sectors = sectors; p.sigma = sigma; p.hue_offset = hue;
    p.r_in = r_in; p.r_out = r_out;
    for (int64_t i = 0; i < n; ++i)
        rgba[i] = shade_pixel(geo[i].outcome, geo[i].r, geo[i].redshift, geo[i].magnif, geo[i].phi_disk,
                              geo[i].theta_esc, geo[i].phi_esc, p, bg, bw, bh);
    return 0;
}

extern "C" int ktc_isco_exact(double M, double a, double Q, double Lambda, double* seq_out, double* seg_out) {
    FrameDev F;
    std::memset(&F, 0, sizeof(F));
    F.M = M; F.a = a; F.Q = Q; F.Lambda = Lambda;
    const double rh = horizon_seq(F);
    *seq_out = isco_seq(F, rh);
    ScanSeg seg[kMaxSeg];
    const int ns = isco_segments(rh * 1.05, 1e-4 * M, 30.0 * M, seg);
    if (ns < 0) { *seg_out = -1.0; return -1; }
    // same search as the device kernel, sequentially
    for (int64_t k = 0;; ++k) {
        const double rk = isco_abscissa(seg, ns, k);
        if (!(rk < 30.0 * M)) { *seg_out = rk; return ns; }
        if (isco_L(F, isco_abscissa(seg, ns, k + 1)) < isco_L(F, rk) && isco_L(F, rk) > 0.0) { *seg_out = rk; return ns; }
    }
}

extern "C" int ktc_trace_bundle_pixels(const kt_frame* f, const kt_opts* o, const int32_t* px, const int32_t* py,
                                       int64_t n, kt_geopixel* out, int64_t* acc) {
    kt_opts ob = *o;
    ob.bundles = 1;
    FrameDev F;
    fill_frame(*f, ob, F);
    F.chart = CHART_BL;
    RunOpts O;
    O.max_steps = o->max_steps < 1 ? 1 : o->max_steps;
    O.quirks = o->quirks;
    O.step_init = o->step_init < 1e-10 ? 1e-10 : o->step_init;
    O.rtol = o->rtol; O.atol = o->atol;
    O.isl_noise = kMixedNoiseDefault; O.isl_r = kMixedRIslandDefault;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < n; ++i) {
        BundleState B;
        RayOut ro;
        double mag = 1.0;
        bundle_init(F, O, px[i], py[i], B);
        if (O.quirks & Q_BUNDLE_ROOT_SKIP)
            while (bundle_step<false>(F, O, B, ro, mag) != STEP_DONE) {}
        else
            while (bundle_step<true>(F, O, B, ro, mag) != STEP_DONE) {}
        uint32_t w[7];
        pack_geopixel_m(ro, mag, w);
        std::memcpy(&out[i], w, 28);
        if (acc) acc[i] = B.g.n_acc;
    }
    return 0;
}

// the inline binary64 sin/cos of the kernels, on n arguments
extern "C" int ktc_sincos(const double* x, int64_t n, double* s, double* c) {
    for (int64_t i = 0; i < n; ++i) kt_sincos(x[i], &s[i], &c[i]);
    return 0;
}

// ---------------------------------------------------------------------------
//  adjoint on the host: forward with records, then the kernels' reverse pass
// ---------------------------------------------------------------------------
// the GPU's segment-parallel reverse pass, run serially: four sweeps per checkpoint segment,
// the event entry, then the composition
template <int CHART>
static bool adjoint_pixel_segmented(const FrameDev& F, int px, int py, int N, const double* hs, const double* ck,
                                    double g[AP_N]) {
    double pt, pp;
    ray_momenta<CHART>(F, px, py, pt, pp);
    const int nseg = adjoint_segments(N);
    std::vector<double> seg((size_t)4 * kSegOut * (nseg > 0 ? nseg : 1));
    LocalStates st;
    for (int k = 0; k < nseg; ++k)
        for (int d = 0; d < 4; ++d)
            adjoint_segment<CHART>(F, pt, pp, N, hs, ck, k, d, st, &seg[(size_t)(4*k + d)*kSegOut]);
    double ev[kSegOut];
    if (!adjoint_event<CHART>(F, pt, pp, N, hs, ck, st, ev)) return false;
    adjoint_compose<CHART>(F, px, py, nseg, ev, seg.data(), seg.data() + 4*kSegOut*(nseg > 0 ? nseg - 1 : 0), g);
    return true;
}

template <int CHART, bool SEG>
static void adjoint_run(const FrameDev& F, const RunOpts& O, const int32_t* px, const int32_t* py, int64_t n,
                        double* inten, double* grad, int32_t* nacc) {
    const Knds<float> Pf = knds_cast<float>(F.P);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        RayState S;
        RayOut o;
        std::vector<double> hs, ck;
        ray_init<CHART, PREC_FP64>(F, O, px[i], py[i], S);
        auto push_ck = [&]() { ck.insert(ck.end(), {S.y[0], S.y[1], S.y[2], S.y[3], S.sc.s, S.sc.c}); };
        push_ck();
        for (;;) {
            const double h = S.h;
            const int n0 = S.n_acc;
            const bool done = ray_advance<CHART, METHOD_DP5, PREC_FP64, false>(F, F.P, Pf, O, true, S, o);
            if (S.n_acc > n0) {
                hs.push_back(h);
                if (!done && S.n_acc % kCkptEvery == 0) push_ck();
            }
            if (done) break;
        }
        const double I = (o.fate == 1) ? disk_intensity_t<double>(o.r, o.g, F.r_isco) : 0.0;
        if (inten) inten[i] = I;
        if (nacc) nacc[i] = S.n_acc;
        double g[AP_N] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (grad && I > 0.0 && S.n_acc >= 1) {
            const bool ok = SEG ? adjoint_pixel_segmented<CHART>(F, px[i], py[i], S.n_acc, hs.data(), ck.data(), g)
                                : adjoint_pixel<CHART>(F, px[i], py[i], S.n_acc, hs.data(), ck.data(), g);
            if (!ok) for (int q = 0; q < AP_N; ++q) g[q] = std::nan("");
        }
        if (grad) for (int q = 0; q < AP_N; ++q) grad[i*AP_N + q] = g[q];
    }
}

// intensity (and dI/dp when grad != NULL) of n pixels, binary64 DOPRI5
static int adjoint_pixels_impl(const kt_frame* f, const kt_opts* o, const int32_t* px, const int32_t* py, int64_t n,
                               double* inten, double* grad, int32_t* nacc, double* info5, bool seg) {
    FrameDev F;
    fill_frame(*f, *o, F);
    RunOpts O;
    O.max_steps = o->max_steps < 1 ? 1 : o->max_steps;
    O.quirks = o->quirks;
    O.step_init = o->step_init < 1e-10 ? 1e-10 : o->step_init;
    O.rtol = o->rtol; O.atol = o->atol;
    O.isl_noise = kMixedNoiseDefault;
    O.isl_r = kMixedRIslandDefault;
    if (info5) { info5[0] = F.r_h; info5[1] = F.r_isco; info5[2] = F.r_esc; info5[3] = F.chart; info5[4] = F.ks_ref_init; }
    if (F.chart == CHART_KS) {
        if (seg) adjoint_run<CHART_KS, true>(F, O, px, py, n, inten, grad, nacc);
        else adjoint_run<CHART_KS, false>(F, O, px, py, n, inten, grad, nacc);
    } else {
        if (seg) adjoint_run<CHART_BL, true>(F, O, px, py, n, inten, grad, nacc);
        else adjoint_run<CHART_BL, false>(F, O, px, py, n, inten, grad, nacc);
    }
    return 0;
}

// intensity (and dI/dp when grad != NULL) of n pixels: one sequential reverse sweep per pixel
extern "C" int ktc_adjoint_pixels(const kt_frame* f, const kt_opts* o, const int32_t* px, const int32_t* py, int64_t n,
                                  double* inten, double* grad, int32_t* nacc, double* info5) {
    return adjoint_pixels_impl(f, o, px, py, n, inten, grad, nacc, info5, false);
}

// the same through the GPU's decomposition: segment sweeps per basis direction, composed
extern "C" int ktc_adjoint_pixels_seg(const kt_frame* f, const kt_opts* o, const int32_t* px, const int32_t* py,
                                      int64_t n, double* inten, double* grad, int32_t* nacc, double* info5) {
    return adjoint_pixels_impl(f, o, px, py, n, inten, grad, nacc, info5, true);
}

// templated camera map (T = double) against the kernels' camera_ray + covector: 7 values per pixel each
extern "C" int ktc_camera_pair(const kt_frame* f, const kt_opts* o, const int32_t* px, const int32_t* py, int64_t n,
                               double* st_kernel, double* st_template) {
    FrameDev F;
    fill_frame(*f, *o, F);
    for (int64_t i = 0; i < n; ++i) {
        double st[7], al, be;
        camera_ray(F, (double)px[i], (double)py[i], st);
        if (F.chart == CHART_KS) bl_to_ks_covector(F, st);
        for (int k = 0; k < 7; ++k) st_kernel[7*i + k] = st[k];
        pixel_angles(F, (double)px[i], (double)py[i], al, be);
        ObsT<double> C;
        observer_t<double>(F.M, F.a, F.Q, F.Lambda, F.r_obs, F.th_obs, C);
        double s2[7];
        camera_t<double>(C, F.r_obs, F.th_obs, al, be, s2);
        s2[2] = st[2];
        if (F.chart == CHART_KS) bl_to_ks_t<double>(knds_t<double>(F.M, F.a, F.Q, F.Lambda), F.ks_ref_init != 0, s2);
        for (int k = 0; k < 7; ++k) st_template[7*i + k] = s2[k];
    }
    return 0;
}

// rhs Jacobian (dual numbers) at one point: J[4][9]
extern "C" int ktc_rhs_jacobian(int chart, double M, const double* par3, const double* y4, double pt, double pp,
                                double* J) {
    double Jl[4][9];
    if (chart == CHART_KS) rhs_jacobian<CHART_KS>(M, par3, y4, pt, pp, Jl);
    else rhs_jacobian<CHART_BL>(M, par3, y4, pt, pp, Jl);
    for (int i = 0; i < 4; ++i) for (int k = 0; k < 9; ++k) J[9*i + k] = Jl[i][k];
    return 0;
}

// vector-Jacobian product of the right-hand side (reverse mode), g[9] = J^T mu
extern "C" int ktc_rhs_vjp(int chart, double M, const double* par3, const double* y4, double pt, double pp,
                           const double* mu4, double* g9) {
    if (chart == CHART_KS) rhs_vjp<CHART_KS>(M, par3, y4, pt, pp, mu4, g9);
    else rhs_vjp<CHART_BL>(M, par3, y4, pt, pp, mu4, g9);
    return 0;
}

// rhs value at one point with sin/cos from the argument (for finite differences)
extern "C" int ktc_rhs(int chart, double M, const double* par3, const double* y4, double pt, double pp, double* f4) {
    const Knds<double> P = knds_t<double>(M, par3[0], par3[1], par3[2]);
    SinCos<double> sc;
    kt_sincos(y4[1], &sc.s, &sc.c);
    double o5[5];
    if (chart == CHART_KS) rhs<CHART_KS>(P, y4[0], sc, y4[2], y4[3], pt, pp, o5);
    else rhs<CHART_BL>(P, y4[0], sc, y4[2], y4[3], pt, pp, o5);
    for (int i = 0; i < 4; ++i) f4[i] = o5[i];
    return 0;
}

// the right-hand side with (sin, cos) given as independent inputs, and the reverse-mode
// adjoints of (sin, cos) from rhs_vjp_sc (finite differences in s and c separately pin them)
extern "C" int ktc_rhs_sc(int chart, double M, const double* par3, const double* y4, double s, double c, double pt,
                          double pp, double* f4) {
    const Knds<double> P = knds_t<double>(M, par3[0], par3[1], par3[2]);
    const SinCos<double> sc = {s, c};
    double o5[5];
    if (chart == CHART_KS) rhs<CHART_KS>(P, y4[0], sc, y4[2], y4[3], pt, pp, o5);
    else rhs<CHART_BL>(P, y4[0], sc, y4[2], y4[3], pt, pp, o5);
    for (int i = 0; i < 4; ++i) f4[i] = o5[i];
    return 0;
}

extern "C" int ktc_rhs_vjp_sc(int chart, double M, const double* par3, const double* y4, double s, double c,
                              double pt, double pp, const double* mu4, double* g2) {
    const Knds<double> P = knds_t<double>(M, par3[0], par3[1], par3[2]);
    const SinCos<double> sc = {s, c};
    double g9[9];
    if (chart == CHART_KS) rhs_vjp_sc<CHART_KS>(P, par3[1], par3[2], y4, sc, pt, pp, mu4, g9, g2);
    else rhs_vjp_sc<CHART_BL>(P, par3[1], par3[2], y4, sc, pt, pp, mu4, g9, g2);
    return 0;
}

// One DOPRI5 step of the forward pass's arithmetic (dp5_replay) from y with sin/cos of y[1],
// and its adjoint (dp5_step_adjoint): lam4 at the step end -> out9 = adjoints of y (4),
// p_t, p_phi, a, Q, Lambda.
extern "C" int ktc_dp5_step(int chart, double M, const double* par3, const double* y4, double pt, double pp,
                            double h, double* y_out) {
    const Knds<double> P = knds_t<double>(M, par3[0], par3[1], par3[2]);
    double y[4] = {y4[0], y4[1], y4[2], y4[3]}, k1[5];
    SinCos<double> sc;
    kt_sincos(y[1], &sc.s, &sc.c);
    if (chart == CHART_KS) {
        rhs<CHART_KS>(P, y[0], sc, y[2], y[3], pt, pp, k1);
        dp5_replay<CHART_KS>(P, y, sc, k1, h, pt, pp);
    } else {
        rhs<CHART_BL>(P, y[0], sc, y[2], y[3], pt, pp, k1);
        dp5_replay<CHART_BL>(P, y, sc, k1, h, pt, pp);
    }
    for (int i = 0; i < 4; ++i) y_out[i] = y[i];
    return 0;
}

extern "C" int ktc_dp5_step_adjoint(int chart, double M, const double* par3, const double* y4, double pt, double pp,
                                    double h, const double* lam4, double* out9) {
    const Knds<double> P = knds_t<double>(M, par3[0], par3[1], par3[2]);
    SinCos<double> sc;
    kt_sincos(y4[1], &sc.s, &sc.c);
    double lam[4] = {lam4[0], lam4[1], lam4[2], lam4[3]}, gc[2] = {0.0, 0.0}, gP[3] = {0.0, 0.0, 0.0};
    if (chart == CHART_KS) dp5_step_adjoint<CHART_KS>(M, par3, P, y4, sc, h, pt, pp, lam, gc, gP);
    else dp5_step_adjoint<CHART_BL>(M, par3, P, y4, sc, h, pt, pp, lam, gc, gP);
    for (int i = 0; i < 4; ++i) out9[i] = lam[i];
    out9[4] = gc[0]; out9[5] = gc[1];
    out9[6] = gP[0]; out9[7] = gP[1]; out9[8] = gP[2];
    return 0;
}

// Fixed-step pieces of the adjoint (no step-size control): replay N steps of sizes hs from
// (y0, p_t, p_phi) with parameters par = (a, Q, Lambda) and return the intensity of the disk
// event in the last step; with lam_out != NULL also the reverse pass: adjoints of y0 (4),
// (p_t, p_phi) (2) and (a, Q, Lambda) (3).
extern "C" int ktc_replay_intensity(const kt_frame* f, const kt_opts* o, const double* par, const double* y0,
                                    double pt, double pp, const double* hs, int N, double* inten, double* lam_out) {
    FrameDev F;
    kt_frame f2 = *f;
    f2.a = par[0]; f2.Q = par[1]; f2.Lambda = par[2];
    fill_frame(f2, *o, F);
    const int chart = F.chart;
    auto go = [&](auto chart_tag) -> int {
        constexpr int CH = decltype(chart_tag)::value;
        std::vector<double> ys((size_t)(N + 1) * 4);
        std::vector<SinCos<double>> scs((size_t)N + 1);
        double y[4] = {y0[0], y0[1], y0[2], y0[3]};
        SinCos<double> sc;
        kt_sincos(y[1], &sc.s, &sc.c);
        double k1[5];
        rhs<CH>(F.P, y[0], sc, y[2], y[3], pt, pp, k1);
        for (int i = 0; i < 4; ++i) ys[i] = y[i];
        scs[0] = sc;
        for (int n = 0; n < N; ++n) {
            dp5_replay<CH>(F.P, y, sc, k1, hs[n], pt, pp);
            for (int i = 0; i < 4; ++i) ys[(size_t)(n + 1)*4 + i] = y[i];
            scs[n + 1] = sc;
        }
        double lam0[4], lam1[4], gc[2] = {0, 0}, gP[3] = {0, 0, 0};
        if (!event_seed<CH>(F, &ys[(size_t)(N - 1)*4], &ys[(size_t)N*4], hs[N - 1], pt, pp, lam0, lam1, gc, gP, inten))
            return 1;
        if (!lam_out) return 0;
        double lam[4] = {lam1[0], lam1[1], lam1[2], lam1[3]};
        const double prm[3] = {F.a, F.Q, F.Lambda};
        for (int n = N - 1; n >= 0; --n) {
            dp5_step_adjoint<CH>(F.M, prm, F.P, &ys[(size_t)n*4], scs[n], hs[n], pt, pp, lam, gc, gP);
            if (n == N - 1) for (int i = 0; i < 4; ++i) lam[i] += lam0[i];
        }
        for (int i = 0; i < 4; ++i) lam_out[i] = lam[i];
        lam_out[4] = gc[0]; lam_out[5] = gc[1];
        lam_out[6] = gP[0]; lam_out[7] = gP[1]; lam_out[8] = gP[2];
        return 0;
    };
    if (chart == CHART_KS) return go(std::integral_constant<int, CHART_KS>());
    return go(std::integral_constant<int, CHART_BL>());
}

// the accepted step sizes and the initial state of one pixel's binary64 DOPRI5 ray
extern "C" int ktc_ray_steps(const kt_frame* f, const kt_opts* o, int px, int py, double* hs, int cap, int* n_out,
                             double* y0c, int* fate) {
    FrameDev F;
    fill_frame(*f, *o, F);
    RunOpts O;
    O.max_steps = o->max_steps; O.quirks = o->quirks; O.step_init = o->step_init; O.rtol = o->rtol; O.atol = o->atol;
    O.isl_noise = kMixedNoiseDefault; O.isl_r = kMixedRIslandDefault;
    const Knds<float> Pf = knds_cast<float>(F.P);
    RayState S;
    RayOut out;
    auto go = [&](auto chart_tag) {
        constexpr int CH = decltype(chart_tag)::value;
        ray_init<CH, PREC_FP64>(F, O, px, py, S);
        y0c[0] = S.y[0]; y0c[1] = S.y[1]; y0c[2] = S.y[2]; y0c[3] = S.y[3]; y0c[4] = S.pt; y0c[5] = S.pp;
        int n = 0;
        for (;;) {
            const double h = S.h;
            const int n0 = S.n_acc;
            const bool done = ray_advance<CH, METHOD_DP5, PREC_FP64, false>(F, F.P, Pf, O, true, S, out);
            if (S.n_acc > n0 && n < cap) hs[n++] = h;
            if (done) break;
        }
        *n_out = n;
        *fate = out.fate;
    };
    if (F.chart == CHART_KS) go(std::integral_constant<int, CHART_KS>());
    else go(std::integral_constant<int, CHART_BL>());
    return 0;
}

// Event pre-tests against the search they guard.  Per case: th0, th1, dth0, dth1, r0, r1, dr0, dr1,
// h, r_in, r_out.  out (4 per case): hermite_may_touch on theta; hermite_first_root found a
// crossing of pi/2; hermite_may_reach on r; the crossing's Hermite r lies in [r_in, r_out].
extern "C" int ktc_event_filters(const double* in, int64_t n, int32_t* out) {
    const double halfpi = kPi / 2.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* c = in + 11*i;
        const double th0 = c[0], th1 = c[1], dth0 = c[2], dth1 = c[3];
        const double r0 = c[4], r1 = c[5], dr0 = c[6], dr1 = c[7], h = c[8], r_in = c[9], r_out = c[10];
        double al = 0.0;
        const bool found = hermite_first_root(th0, th1, dth0, dth1, h, halfpi, al);
        const double rh = found ? hermite(r0, r1, dr0, dr1, h, al) : 0.0;
        out[4*i + 0] = hermite_may_touch(th0, th1, dth0, dth1, h, halfpi) ? 1 : 0;
        out[4*i + 1] = found ? 1 : 0;
        out[4*i + 2] = hermite_may_reach(r0, r1, dr0, dr1, h, r_in, r_out) ? 1 : 0;
        out[4*i + 3] = (found && rh >= r_in && rh <= r_out) ? 1 : 0;
    }
    return 0;
}

// The mixed path's constraint projection (kt_trace.hpp renormalise) on a BL or KS state.
// in: 6 per case (r, theta, p_r, p_theta, p_t, p_phi); out: 2 per case (p_r, p_theta after).
extern "C" int ktc_renormalise(int chart, double M, const double* par3, const double* in, int64_t n, double* out) {
    kt_frame f{};
    f.M = M; f.a = par3[0]; f.Q = par3[1]; f.Lambda = par3[2];
    f.r_obs = 40.0; f.theta_deg = 80.0; f.fov_deg = 30.0; f.disk_out = 12.0; f.width = 8; f.height = 8;
    kt_opts o{};
    o.chart = chart; o.integrator = KT_INTEGRATOR_DOPRI5; o.rtol = 1e-9; o.atol = 1e-9; o.max_steps = 1000;
    o.step_init = 1.0; o.quirks = KT_QUIRKS_REFERENCE;
    FrameDev F;
    fill_frame(f, o, F);
    for (int64_t i = 0; i < n; ++i) {
        const double* c = in + 6*i;
        RayState S;
        S.y[0] = c[0]; S.y[1] = c[1]; S.y[2] = c[2]; S.y[3] = c[3];
        S.pt = c[4]; S.pp = c[5];
        kt_sincos(S.y[1], &S.sc.s, &S.sc.c);
        if (chart == KT_CHART_KS) renormalise<CHART_KS>(F.P, S); else renormalise<CHART_BL>(F.P, S);
        out[2*i] = S.y[2];
        out[2*i + 1] = S.y[3];
    }
    return 0;
}
