// Host build of the product's per-ray tracer (kerrtracecpp2_amd/csrc/kt_trace.hpp),
// for CPU-only tests of the kernel logic against the oracle.  Test
// infrastructure: not part of the product path, never loaded by it.
#include <cstring>
#include <vector>

#include "../../include/kt_hip.h"
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
    if (info5) { info5[0] = F.r_h; info5[1] = F.r_isco; info5[2] = F.r_esc; info5[3] = F.chart; info5[4] = F.ks_ref_init; }
    const bool phi = !(o->integrator == KT_INTEGRATOR_DOPRI5 && (o->quirks & KT_QUIRK_DOPRI5_PHI_FROZEN));
    if (F.chart == CHART_KS) run_chart<CHART_KS>(o->integrator, o->precision, phi, F, O, px, py, n, out, acc, rhs32, rhs64);
    else run_chart<CHART_BL>(o->integrator, o->precision, phi, F, O, px, py, n, out, acc, rhs32, rhs64);
    return 0;
}

extern "C" int ktc_shade(const kt_geopixel* geo, int64_t n, double M, double r_isco, double exposure, double gamma,
                         double temp_scale, const uint8_t* bg, int bw, int bh, uint32_t* rgba) {
    ShadeParams p{M, r_isco, exposure, gamma, temp_scale};
    for (int64_t i = 0; i < n; ++i)
        rgba[i] = shade_pixel(geo[i].outcome, geo[i].r, geo[i].redshift, geo[i].magnif, geo[i].theta_esc,
                              geo[i].phi_esc, p, bg, bw, bh);
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
        while (bundle_step(F, O, B, ro, mag) != STEP_DONE) {}
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
