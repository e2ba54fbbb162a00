// ============================================================================
//  dropin_check.cpp -- the HIP backend as the reference's host code would call it.
//
//  1. The USE_HIP branch that INTEGRATION.md section 2 gives for render_image()
//     (main.cpp:2209-2478) is compiled here verbatim: build.py copies the code block
//     out of INTEGRATION.md into _build/snippet.inc, which is included inside a
//     render_image() of the reference's signature.  The types around it are the
//     reference caller's context as far as the block touches it (main.cpp:62-94,
//     140-142, 174-197, 1886-1893; geodesic.hpp:417).  KNdSMetric's horizon and ISCO
//     return NaN: the backend recomputes both on the device and must not use them.
//  2. kthip::hip_render(), the cuda_render() call shape (gpu/cuda/tracer.cuh:23-26).
//  3. kthip::hip_render_ex() with the reference's controls and a background.
//
//  Usage: dropin_check OUT_DIR   -> OUT_DIR/{snippet.rgb, snippet.geo, cuda_shape.abgr,
//                                   render_ex.abgr}  for the 320x180 BL RK4 frame
//                                   (a=0.5, theta 80, fov 45, r_obs 40, disk to 12 M)
// ============================================================================
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hip_backend.hpp"

// ---- the reference caller's context ----------------------------------------
enum class CoordinateChart { KS, BL };
enum class RaySolverMode { STANDARD, SEMI_ANALYTIC, ELLIPTIC_CLOSED };
enum class Integrator { RK4_DOUBLING, DOPRI5 };
struct RGB { uint8_t r, g, b; };
struct ColorParams { double exposure = 1.0, gamma = 2.2, temp_scale = 1.0; };
struct GeoPixel {
    uint8_t outcome; uint8_t _pad[3];
    float r, redshift, magnif, phi_disk, theta_esc, phi_esc;
};
static_assert(sizeof(GeoPixel) == 28, "GeoPixel");
struct IntegratorControls { int max_steps = 500000; double step_init = 1.0; double tol = 1e-7; };
struct FrameParams {
    double a = 0.998, theta = 80.0, phi = 0.0, r_obs = 40.0, disk_out = 12.0, fov = 30.0;
    bool wormhole = false;
};
struct BackgroundImage { int w = 0, h = 0; std::vector<uint8_t> px; };
struct KNdSMetric {
    KNdSMetric(double, double, double, double) {}
    double r_horizon() const { return NAN; }
    double r_isco() const { return NAN; }
};

static std::vector<RGB> render_image(int W, int H, const FrameParams& fp, const BackgroundImage& bg, bool use_bundles,
                                     RaySolverMode solver_mode, CoordinateChart chart, Integrator intg,
                                     const IntegratorControls& ctl, double M_bh, double Q_bh, double Lam,
                                     bool gpu_fp64, const ColorParams& cp, std::vector<GeoPixel>* geo_out,
                                     double pixel_offset_x, double pixel_offset_y) {
#include "snippet.inc"
    throw std::runtime_error("the USE_HIP branch did not take this frame");
}

static void save(const std::string& path, const void* p, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, n, f) != n) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: %s OUT_DIR\n", argv[0]); return 2; }
    const std::string out = argv[1];
    const int W = 320, H = 180;
    try {
        BackgroundImage bg;
        bg.w = 512; bg.h = 256;
        bg.px.resize((size_t)bg.w * bg.h * 3);
        if (kt_make_checker_background(1234, bg.w, bg.h, 32, 16, bg.px.data()) != KT_OK) throw std::runtime_error("sky");
        FrameParams fp;
        fp.a = 0.5; fp.theta = 80.0; fp.phi = 0.0; fp.r_obs = 40.0; fp.disk_out = 12.0; fp.fov = 45.0;
        IntegratorControls ctl;
        ColorParams cp;
        std::vector<GeoPixel> geo;
        const std::vector<RGB> img = render_image(W, H, fp, bg, false, RaySolverMode::STANDARD, CoordinateChart::BL,
                                                  Integrator::RK4_DOUBLING, ctl, 1.0, 0.0, 0.0, true, cp, &geo, 0.0, 0.0);
        save(out + "/snippet.rgb", img.data(), img.size() * sizeof(RGB));
        save(out + "/snippet.geo", geo.data(), geo.size() * sizeof(GeoPixel));

        const double d2r = 3.14159265358979323846 / 180.0;
        const kthip::KNdSParams_HIP kp{1.0, 0.5, 0.0, 0.0, NAN, NAN, 12.0};
        const kthip::CameraParams_HIP cam{40.0, 80.0 * d2r, 0.0, 45.0 * d2r, W, H, 0};
        const std::vector<uint32_t> px = kthip::hip_render(kp, cam, true);
        save(out + "/cuda_shape.abgr", px.data(), px.size() * 4);

        kthip::RenderOptions ro;
        ro.bg_rgb = bg.px.data(); ro.bg_w = bg.w; ro.bg_h = bg.h;
        const kthip::RenderResult r = kthip::hip_render_ex(kp, cam, true, ro);
        save(out + "/render_ex.abgr", r.rgba.data(), r.rgba.size() * 4);
        std::printf("dropin_check: %d x %d, %llu rays, %llu accepted steps, r_horizon %.12f r_isco %.12f\n", W, H,
                    (unsigned long long)r.stats.rays, (unsigned long long)r.stats.steps_accepted, r.info.r_horizon,
                    r.info.r_isco);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dropin_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
