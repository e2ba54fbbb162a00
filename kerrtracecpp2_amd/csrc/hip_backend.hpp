// ============================================================================
//  hip_backend.hpp -- C++ host wrapper over the C ABI (include/kt_hip.h): the
//  drop-in for the reference's CUDA backend call.
//
//  Reference interface (KerrTraceCpp2 tree):
//    struct KNdSParams_CUDA {M,a,Q,Lambda,r_horizon,r_isco,r_disk_out}  gpu/cuda/tracer.cuh:11-14
//    struct CameraParams_CUDA {r_obs,theta_obs,phi_obs,fov_h,width,height,chart}   :16-20
//    std::vector<uint32_t> cuda_render(kp, cp, bool require_fp64)          :23-26
//    errors: std::runtime_error (CUDA_CHECK, gpu/cuda/tracer.cu:19-23)
//  hip_render() has the same parameter meaning (angles in radians, chart 0 BL /
//  1 KS), returns the same packed ABGR pixels and throws std::runtime_error on
//  failure.  hip_render_ex() adds what render_image() passes to the Metal
//  backend (integrator, tolerance, max steps, initial step, offsets, bundles,
//  background) and returns the GeoPixel buffer as well.
// ============================================================================
#pragma once
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kt_hip.h"

namespace kthip {

struct KNdSParams_HIP {
    double M, a, Q, Lambda;
    double r_horizon, r_isco, r_disk_out;     // r_horizon / r_isco are recomputed on the device
};

struct CameraParams_HIP {
    double r_obs, theta_obs, phi_obs, fov_h;  // radians, as main.cpp:2435 passes them
    int width, height;
    int chart;                                // 0 = BL, 1 = KS
};

struct RenderOptions {
    int integrator = KT_INTEGRATOR_RK4;
    double tol = 1e-7;                // reference single tolerance (rtol = atol = tol)
    double atol = -1.0;               // < 0: same as tol
    int max_steps = 500000;
    double step_init = 1.0;
    bool bundles = false;
    bool mixed_precision = false;     // default precision of hip_render_ex; --gpu-fp64 forces FP64
    unsigned quirks = KT_QUIRKS_REFERENCE;
    double pixel_offset_x = 0.0, pixel_offset_y = 0.0;
    double exposure = 1.0, gamma = 2.2, temp_scale = 1.0;
    int palette = KT_PALETTE_BLACKBODY;   // ColorParams (main.cpp:70-81)
    int disk_rings = 7, disk_sectors = 14;
    double disk_sigma = 0.5, disk_hue_offset = 0.0;
    const uint8_t* bg_rgb = nullptr;  // equirect RGB8 background
    int bg_w = 0, bg_h = 0;
};

class Context {
public:
    explicit Context(int device = 0) {
        const int rc = kt_hip_create(device, &ctx_);
        if (rc != KT_OK) throw std::runtime_error("kt_hip_create failed (status " + std::to_string(rc) + ")");
    }
    ~Context() { kt_hip_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    kt_hip_ctx* get() const { return ctx_; }
    void check(int rc, const char* what) const {
        if (rc != KT_OK)
            throw std::runtime_error(std::string(what) + ": " + kt_hip_last_error(ctx_) + " (status " +
                                     std::to_string(rc) + ")");
    }
    std::string device_name() {
        char name[256];
        int ncu = 0, f64 = 0;
        check(kt_hip_device_info(ctx_, name, sizeof(name), &ncu, &f64), "kt_hip_device_info");
        return std::string(name) + ", " + std::to_string(ncu) + " CUs, FP64 native=" + (f64 ? "yes" : "no");
    }
    bool has_fp64() {
        int ncu = 0, f64 = 0;
        check(kt_hip_device_info(ctx_, nullptr, 0, &ncu, &f64), "kt_hip_device_info");
        return f64 != 0;
    }

private:
    kt_hip_ctx* ctx_ = nullptr;
};

inline Context& default_context() {
    static Context ctx(0);
    return ctx;
}

inline kt_frame make_frame(const KNdSParams_HIP& kp, const CameraParams_HIP& cp, double ox = 0.0, double oy = 0.0) {
    const double r2d = 180.0 / 3.14159265358979323846;
    kt_frame f;
    f.M = kp.M; f.a = kp.a; f.Q = kp.Q; f.Lambda = kp.Lambda;
    f.r_obs = cp.r_obs;
    f.theta_deg = cp.theta_obs * r2d;
    f.phi_deg = cp.phi_obs * r2d;
    f.fov_deg = cp.fov_h * r2d;
    f.disk_out = kp.r_disk_out;
    f.offset_x = ox; f.offset_y = oy;
    f.width = cp.width; f.height = cp.height;
    return f;
}

inline kt_opts make_opts(int chart, const RenderOptions& ro, bool fp64) {
    kt_opts o;
    kt_hip_default_opts(&o);
    o.chart = chart;
    o.integrator = ro.integrator;
    o.precision = fp64 ? KT_PRECISION_FP64 : (ro.mixed_precision ? KT_PRECISION_MIXED : KT_PRECISION_FP64);
    o.bundles = ro.bundles ? 1 : 0;
    o.max_steps = ro.max_steps < 1 ? 1 : ro.max_steps;
    o.quirks = ro.quirks;
    o.step_init = ro.step_init < 1e-10 ? 1e-10 : ro.step_init;
    o.rtol = ro.tol < 1e-10 ? 1e-10 : ro.tol;
    o.atol = ro.atol > 0.0 ? ro.atol : o.rtol;
    o.exposure = ro.exposure; o.gamma = ro.gamma; o.temp_scale = ro.temp_scale;
    o.palette = ro.palette; o.disk_rings = ro.disk_rings; o.disk_sectors = ro.disk_sectors;
    o.disk_sigma = ro.disk_sigma; o.disk_hue_offset = ro.disk_hue_offset;
    return o;
}

// Same contract as cuda_render(): FP64 kernel, reference defaults (RK4, tol 1e-7), packed ABGR.
inline std::vector<uint32_t> hip_render(const KNdSParams_HIP& kp, const CameraParams_HIP& cp, bool require_fp64,
                                        Context& ctx = default_context()) {
    if (require_fp64 && !ctx.has_fp64())
        throw std::runtime_error("HIP FP64 strict mode requested, but the device does not expose native FP64.");
    RenderOptions ro;
    const kt_frame f = make_frame(kp, cp);
    const kt_opts o = make_opts(cp.chart, ro, true);
    std::vector<uint32_t> px((size_t)cp.width * cp.height);
    ctx.check(kt_hip_set_background(ctx.get(), nullptr), "kt_hip_set_background");
    ctx.check(kt_hip_render(ctx.get(), &f, 1, &o, px.data(), nullptr, nullptr, nullptr), "kt_hip_render");
    return px;
}

struct RenderResult {
    std::vector<uint32_t> rgba;
    std::vector<kt_geopixel> geo;
    kt_frame_info info;
    kt_stats stats;
};

// Extended render: reference integrator controls, bundles, offsets, background, GeoPixels.
inline RenderResult hip_render_ex(const KNdSParams_HIP& kp, const CameraParams_HIP& cp, bool gpu_fp64,
                                  const RenderOptions& ro, Context& ctx = default_context()) {
    if (gpu_fp64 && !ctx.has_fp64())
        throw std::runtime_error("HIP FP64 strict mode requested, but the device does not expose native FP64.");
    const kt_frame f = make_frame(kp, cp, ro.pixel_offset_x, ro.pixel_offset_y);
    const kt_opts o = make_opts(cp.chart, ro, gpu_fp64);
    kt_background bg{ro.bg_rgb, ro.bg_w, ro.bg_h};
    ctx.check(kt_hip_set_background(ctx.get(), ro.bg_rgb ? &bg : nullptr), "kt_hip_set_background");
    RenderResult r;
    r.rgba.resize((size_t)cp.width * cp.height);
    r.geo.resize((size_t)cp.width * cp.height);
    ctx.check(kt_hip_render(ctx.get(), &f, 1, &o, r.rgba.data(), r.geo.data(), &r.info, &r.stats), "kt_hip_render");
    return r;
}

}  // namespace kthip
