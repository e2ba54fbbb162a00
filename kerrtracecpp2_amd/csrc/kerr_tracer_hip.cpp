// ============================================================================
//  kerr_tracer_hip -- command-line renderer over libkt_hip.so.
//
//  Flags, defaults, resolution presets, output naming and the two-phase
//  (--geo-only / --color-only) and animation modes follow the reference CLI
//  (main.cpp:2496-2947, 2951-3062); the backend tag is "gpu-hip".  Differences:
//    * every supported combination runs on the GPU (bundles, DOPRI5, KS with
//      any Lambda), there is no CPU fallback;
//    * --camera-spp N renders its N jittered passes as N frames of one
//      persistent launch and resolves them in linear light on the device
//      (kt_hip_render_spp; main.cpp:2232-2322);
//    * --anim renders its frames in batches of --anim-batch frames per launch;
//    * --geo-file also works for GPU renders (the reference GPU backends cannot
//      write .kgeo);
//    * --precision mixed|fp64 chooses the kernel; --gpu-fp64 forces fp64;
//    * --atol sets the DOPRI5 absolute tolerance separately (default: --tol);
//    * --quirks fixed turns off the reference-compatibility behaviours.
//  Solver modes other than "standard", the wormhole metric, linear
//  intersection and the Metal/debug switches are not provided by this backend.
//  Both disk palettes are (--disk-blackbody, --disk-interstellar with
//  --disk-rings/--disk-sectors/--disk-sigma/--disk-hue-offset).
// ============================================================================
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "hip_backend.hpp"
#include "kt_image.hpp"

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

std::string timestamp() {
    std::time_t t = std::time(nullptr);
    char ts[32];
    std::strftime(ts, sizeof(ts), "%Y%m%d-%H%M%S", std::localtime(&t));
    return ts;
}

void mkdir_p(const std::string& dir) {
    std::string cur;
    for (size_t i = 0; i < dir.size(); ++i) {
        cur.push_back(dir[i]);
        if (dir[i] == '/' || i + 1 == dir.size()) ::mkdir(cur.c_str(), 0755);
    }
}

bool file_exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

struct Frame {
    double a, theta, phi, r_obs, disk_out;
};

double smooth_step(double t) { return t * t * (3.0 - 2.0 * t); }
double lerp_angle(double a, double b, double t) {
    double d = b - a;
    while (d > 180.0) d -= 360.0;
    while (d < -180.0) d += 360.0;
    return a + d * t;
}

[[noreturn]] void die(const std::string& msg) {
    std::cerr << "Error: " << msg << "\n";
    std::exit(1);
}

}  // namespace

int main(int argc, char** argv) {
    bool use_bundles = false, preview = false, hd_preview = false, res_720p = false, res_2k = false, res_4k = false;
    int custom_w = 0, custom_h = 0;
    int integrator = KT_INTEGRATOR_RK4;
    int max_steps = 500000;
    double step_init = 1.0, tol = 1e-7, atol = -1.0;
    int camera_spp = 1;
    bool gpu_fp64 = false;
    std::string precision = "fp64";
    int chart = KT_CHART_KS;
    unsigned quirks = KT_QUIRKS_REFERENCE;
    std::string bg_path = "assets/backgrounds/sfondo5.jpg";
    std::string out_dir = std::getenv("KT_OUT_DIR") ? std::getenv("KT_OUT_DIR") : "out";
    int device = 0;
    double arg_a = 0.998, arg_disk_out = 12.0, arg_theta = 80.0, arg_phi = 0.0, arg_r_obs = -1.0;
    double arg_Q = 0.0, arg_Lam = 0.0, arg_fov = 30.0;
    double exposure = 1.0, gamma = 2.2, temp_scale = 1.0;
    int palette = KT_PALETTE_BLACKBODY, disk_rings = 7, disk_sectors = 14;
    double disk_sigma = 0.5, disk_hue = 0.0;
    bool geo_only = false, color_only = false;
    std::string geo_file;
    bool anim_mode = false, anim_ease = false, anim_keep_frames = false, anim_resume = false, anim_no_encode = false;
    int anim_frames = 60, anim_fps = 30, anim_crf = 18, anim_batch = 8;
    double anim_orbits = 0.0;
    std::string anim_output, anim_frames_dir;
    const double NaN = std::numeric_limits<double>::quiet_NaN();
    double th0 = NaN, th1 = NaN, ph0 = NaN, ph1 = NaN, r0 = NaN, r1 = NaN, a0 = NaN, a1 = NaN, d0 = NaN, d1 = NaN;
    std::vector<std::string> unsupported;

    for (int i = 1; i < argc; ++i) {
        const std::string arg(argv[i]);
        const bool has1 = i + 1 < argc, has2 = i + 2 < argc;
        auto num = [&]() { return std::stod(argv[++i]); };
        auto inum = [&]() { return std::stoi(argv[++i]); };
        if (arg == "--bundles") use_bundles = true;
        else if (arg == "--dopri5") integrator = KT_INTEGRATOR_DOPRI5;
        else if (arg == "--rk4") integrator = KT_INTEGRATOR_RK4;
        else if (arg == "--max-steps" && has1) max_steps = std::max(1, inum());
        else if ((arg == "--step-init" || arg == "--step-size") && has1) step_init = num();
        else if ((arg == "--integrator-tol" || arg == "--tol") && has1) tol = num();
        else if (arg == "--atol" && has1) atol = num();
        else if ((arg == "--camera-spp" || arg == "--spp") && has1) camera_spp = std::max(1, inum());
        else if (arg == "--gpu-fp64" || arg == "--gpu-fp64-on" || arg == "--metal-fp64" || arg == "--cuda-fp64" ||
                 arg == "--hip-fp64") gpu_fp64 = true;
        else if (arg == "--no-gpu-fp64" || arg == "--no-cuda-fp64" || arg == "--no-hip-fp64") gpu_fp64 = false;
        else if (arg == "--precision" && has1) precision = argv[++i];
        else if (arg == "--quirks" && has1) {
            const std::string q = argv[++i];
            quirks = (q == "fixed" || q == "none") ? 0u : KT_QUIRKS_REFERENCE;
        } else if (arg == "--bl") chart = KT_CHART_BL;
        else if (arg == "--ks") chart = KT_CHART_KS;
        else if (arg == "--chart" && has1) {
            const std::string c = argv[++i];
            if (c == "bl" || c == "BL") chart = KT_CHART_BL;
            else if (c == "ks" || c == "KS") chart = KT_CHART_KS;
        } else if (arg == "--preview") preview = true;
        else if (arg == "--hd") hd_preview = true;
        else if (arg == "--720p") res_720p = true;
        else if (arg == "--2k") res_2k = true;
        else if (arg == "--4k") res_4k = true;
        else if (arg == "--custom-res" && has2) { custom_w = inum(); custom_h = inum(); }
        else if (arg == "--bg" && has1) bg_path = argv[++i];
        else if (arg == "--a" && has1) arg_a = num();
        else if (arg == "--disk-out" && has1) arg_disk_out = num();
        else if (arg == "--theta" && has1) arg_theta = num();
        else if (arg == "--phi" && has1) arg_phi = num();
        else if (arg == "--r-obs" && has1) arg_r_obs = num();
        else if (arg == "--charge" && has1) arg_Q = num();
        else if (arg == "--lambda" && has1) arg_Lam = num();
        else if (arg == "--fov" && has1) arg_fov = num();
        else if (arg == "--exposure" && has1) exposure = num();
        else if (arg == "--gamma" && has1) gamma = num();
        else if (arg == "--temp-scale" && has1) temp_scale = num();
        else if (arg == "--disk-blackbody") palette = KT_PALETTE_BLACKBODY;
        else if (arg == "--disk-interstellar") palette = KT_PALETTE_INTERSTELLAR;
        else if (arg == "--disk-rings" && has1) disk_rings = inum();
        else if (arg == "--disk-sectors" && has1) disk_sectors = inum();
        else if (arg == "--disk-sigma" && has1) disk_sigma = num();
        else if (arg == "--disk-hue-offset" && has1) disk_hue = num();
        else if (arg == "--geo-only") geo_only = true;
        else if (arg == "--geo-file" && has1) geo_file = argv[++i];
        else if (arg == "--color-only" && has1) { color_only = true; geo_file = argv[++i]; }
        else if (arg == "--out-dir" && has1) out_dir = argv[++i];
        else if (arg == "--device" && has1) device = inum();
        else if (arg == "--anim") anim_mode = true;
        else if (arg == "--frames" && has1) anim_frames = inum();
        else if (arg == "--fps" && has1) anim_fps = inum();
        else if (arg == "--crf" && has1) anim_crf = inum();
        else if (arg == "--ease") anim_ease = true;
        else if (arg == "--keep-frames") anim_keep_frames = true;
        else if (arg == "--resume") anim_resume = true;
        else if (arg == "--no-encode") anim_no_encode = true;
        else if (arg == "--orbits" && has1) anim_orbits = num();
        else if (arg == "--output" && has1) anim_output = argv[++i];
        else if (arg == "--frames-dir" && has1) anim_frames_dir = argv[++i];
        else if (arg == "--anim-batch" && has1) anim_batch = std::max(1, inum());
        else if (arg == "--theta-start" && has1) th0 = num();
        else if (arg == "--theta-end" && has1) th1 = num();
        else if (arg == "--phi-start" && has1) ph0 = num();
        else if (arg == "--phi-end" && has1) ph1 = num();
        else if (arg == "--r-start" && has1) r0 = num();
        else if (arg == "--r-end" && has1) r1 = num();
        else if (arg == "--a-start" && has1) a0 = num();
        else if (arg == "--a-end" && has1) a1 = num();
        else if (arg == "--disk-out-start" && has1) d0 = num();
        else if (arg == "--disk-out-end" && has1) d1 = num();
        else if (arg == "--semi-analytic" || arg == "--elliptic" || arg == "--elliptic-closed" ||
                 arg == "--wormhole" || arg == "--intersection-linear" ||
                 arg == "--linear-intersection" || arg == "--debug-elliptic") unsupported.push_back(arg);
        else if ((arg == "--solver-mode" || arg == "--intersection") && has1) {
            const std::string v = argv[++i];
            if (v != "standard" && v != "hermite" && v != "cubic" && v != "event-hermite")
                unsupported.push_back(arg + " " + v);
        } else if ((arg == "--metal-kernel" || arg == "--wh-throat" || arg == "--wh-lensing" || arg == "--wh-tunnel" ||
                    arg == "--bg-b") && has1) ++i;  // accepted and ignored
        else if (arg == "--intersection-hermite" || arg == "--hermite-intersection" || arg == "--anti-fireflies" ||
                 arg == "--no-anti-fireflies" || arg == "--elliptic-fallback-black") {}
        else std::cerr << "Warning: unknown argument '" << arg << "'\n";
    }
    if (!unsupported.empty()) {
        std::string s;
        for (auto& u : unsupported) s += (s.empty() ? "" : ", ") + u;
        die("not provided by the HIP backend: " + s + " (it implements the standard solver, Kerr-Newman-de Sitter "
            "and Hermite disk intersection)");
    }
    if (precision != "fp64" && precision != "mixed") die("--precision must be fp64 or mixed");
    if (use_bundles && precision == "mixed" && !gpu_fp64) die("--bundles runs in binary64 only (--precision fp64)");
    max_steps = std::max(1, max_steps);
    step_init = std::max(1e-10, step_init);
    tol = std::max(1e-10, tol);

    kthip::RenderOptions ro;
    ro.integrator = integrator;
    ro.tol = tol;
    ro.atol = atol;
    ro.max_steps = max_steps;
    ro.step_init = step_init;
    ro.bundles = use_bundles;
    ro.mixed_precision = (precision == "mixed");
    ro.quirks = quirks;
    ro.exposure = exposure; ro.gamma = gamma; ro.temp_scale = temp_scale;
    ro.palette = palette; ro.disk_rings = disk_rings; ro.disk_sectors = disk_sectors;
    ro.disk_sigma = disk_sigma; ro.disk_hue_offset = disk_hue;
    {   // the checks every render entry point applies, before a device is opened
        char msg[256] = {0};
        const kt_opts chk = kthip::make_opts(chart, ro, gpu_fp64 || precision == "fp64");
        if (kt_check_opts(&chk, msg, sizeof msg) != KT_OK) die(std::string("invalid options: ") + msg);
    }

    std::vector<uint8_t> bg_rgb;
    int bg_w = 0, bg_h = 0;
    if (!bg_path.empty()) {
        std::string err;
        if (kthip::load_rgb_image(bg_path, bg_rgb, bg_w, bg_h, &err))
            std::cout << "Background: " << bg_path << " (" << bg_w << "x" << bg_h << ")\n";
        else
            std::cerr << "Warning: cannot load background '" << bg_path << "' (" << err << ")\n";
    }
    if (!bg_rgb.empty()) { ro.bg_rgb = bg_rgb.data(); ro.bg_w = bg_w; ro.bg_h = bg_h; }

    const double M_bh = 1.0;
    const int W = custom_w ? custom_w : res_4k ? 3840 : res_2k ? 2560 : res_720p ? 1280 : hd_preview ? 854 : preview ? 480 : 1920;
    const int H = custom_h ? custom_h : res_4k ? 2160 : res_2k ? 1440 : res_720p ? 720 : hd_preview ? 480 : preview ? 270 : 1080;
    if (W <= 0 || H <= 0) die("resolution must be positive");
    const double base_r_obs = arg_r_obs > 0 ? arg_r_obs : 40.0;
    const char* res_tag = res_4k ? "4k" : res_2k ? "2k" : custom_w ? "custom" : res_720p ? "720p" : hd_preview ? "hd"
                        : preview ? "preview" : use_bundles ? "bundles" : "trace";

    try {
        kthip::Context ctx(device);
        std::cout << "Device: " << ctx.device_name() << "\n";

        // ---------------------------------------------------------- color-only
        if (color_only) {
            std::vector<kt_geopixel> geo;
            kthip::KGeoMeta meta;
            std::string err;
            if (!kthip::load_kgeo(geo_file, geo, meta, &err)) die("cannot load " + geo_file + " (" + err + ")");
            std::cout << "Loaded: " << geo_file << "  " << meta.W << "x" << meta.H << "  a=" << meta.a_bh
                      << "  r_isco=" << meta.r_isco << "\n"
                      << "ColorParams: exposure=" << exposure << " gamma=" << gamma << " temp_scale=" << temp_scale << "\n";
            kt_background bg{ro.bg_rgb, ro.bg_w, ro.bg_h};
            ctx.check(kt_hip_set_background(ctx.get(), ro.bg_rgb ? &bg : nullptr), "kt_hip_set_background");
            const kt_opts o = kthip::make_opts(KT_CHART_BL, ro, true);
            std::vector<uint32_t> px(geo.size());
            ctx.check(kt_hip_colorize(ctx.get(), geo.data(), (int64_t)geo.size(), meta.M_bh, meta.r_isco, meta.r_disk_in,
                                      meta.r_disk_out, &o, px.data()),
                      "kt_hip_colorize");
            std::vector<uint8_t> rgb(px.size() * 3);
            kthip::abgr_to_rgb(px.data(), px.size(), rgb.data());
            std::string base = geo_file;
            if (base.size() > 5 && base.substr(base.size() - 5) == ".kgeo") base = base.substr(0, base.size() - 5);
            const std::string outfile = base + "_recolor_" + timestamp() + ".png";
            if (!kthip::write_png(outfile, rgb.data(), (int)meta.W, (int)meta.H)) die("cannot write " + outfile);
            std::cout << "Saved: " << outfile << "\n";
            return 0;
        }

        // Naming (main.cpp:2760-2853): the effective chart is BL for bundles and for KS with Lambda != 0
        // under the reference quirks; HIP runs DOPRI5 in both charts.
        int eff_chart = chart;
        if (use_bundles) eff_chart = KT_CHART_BL;
        if ((quirks & KT_QUIRK_KS_LAMBDA_TO_BL) && chart == KT_CHART_KS && std::fabs(arg_Lam) > 1e-15) eff_chart = KT_CHART_BL;
        const char* chart_tag = eff_chart == KT_CHART_KS ? "ks" : "bl";
        const std::string mode_tag = use_bundles ? std::string(chart_tag) + "-ray-bundle"
                                                 : std::string(chart_tag) + "-standard-" +
                                                       (integrator == KT_INTEGRATOR_DOPRI5 ? "dopri5" : "rk4");
        const std::string mode_backend_tag = mode_tag + "_gpu-hip" + (gpu_fp64 ? "_gpufp64req" : "") +
                                             (camera_spp > 1 ? "_spp" + std::to_string(camera_spp) : "");
        const bool fp64 = gpu_fp64 || precision == "fp64";

        auto to_kt_frames = [&](const std::vector<Frame>& fr) {
            std::vector<kt_frame> kf;
            for (const Frame& f : fr) {
                kthip::KNdSParams_HIP kp{M_bh, f.a, arg_Q, arg_Lam, 0.0, 0.0, f.disk_out};
                kthip::CameraParams_HIP cp{f.r_obs, 0.0, 0.0, 0.0, W, H, chart};
                kt_frame k = kthip::make_frame(kp, cp);
                k.theta_deg = f.theta; k.phi_deg = f.phi; k.fov_deg = arg_fov;  // degrees as given
                kf.push_back(k);
            }
            return kf;
        };
        auto render_frames = [&](const std::vector<Frame>& fr, std::vector<std::vector<uint8_t>>& images,
                                 std::vector<kt_geopixel>* geo_out, kt_stats* st) {
            std::vector<kt_frame> kf = to_kt_frames(fr);
            if (gpu_fp64 && !ctx.has_fp64())
                throw std::runtime_error("HIP FP64 strict mode requested, but the device does not expose native FP64.");
            const kt_opts o = kthip::make_opts(chart, ro, fp64);
            kt_background bg{ro.bg_rgb, ro.bg_w, ro.bg_h};
            ctx.check(kt_hip_set_background(ctx.get(), ro.bg_rgb ? &bg : nullptr), "kt_hip_set_background");
            const size_t npix = (size_t)W * H;
            images.assign(fr.size(), std::vector<uint8_t>(npix * 3));
            if (camera_spp > 1) {   // all passes of all frames in one launch, resolved on the device
                std::vector<uint8_t> rgb(npix * 3 * kf.size());
                ctx.check(kt_hip_render_spp(ctx.get(), kf.data(), (int)kf.size(), camera_spp, &o, rgb.data(), st),
                          "kt_hip_render_spp");
                for (size_t f = 0; f < fr.size(); ++f)
                    std::copy(rgb.begin() + f * npix * 3, rgb.begin() + (f + 1) * npix * 3, images[f].begin());
                return;
            }
            std::vector<uint32_t> px(npix * kf.size());
            if (geo_out) geo_out->resize(npix * kf.size());
            ctx.check(kt_hip_render(ctx.get(), kf.data(), (int)kf.size(), &o, px.data(),
                                    geo_out ? geo_out->data() : nullptr, nullptr, st),
                      "kt_hip_render");
            for (size_t f = 0; f < fr.size(); ++f) kthip::abgr_to_rgb(px.data() + f * npix, npix, images[f].data());
        };

        auto print_stats = [&](const kt_stats& st) {
            std::cout << "GPU: trace " << std::fixed << std::setprecision(3) << st.trace_ms << " ms, total "
                      << st.total_ms << " ms, " << st.rays << " rays, " << st.steps_accepted << " accepted / "
                      << st.steps_rejected << " rejected steps\n" << std::defaultfloat;
        };

        // -------------------------------------------------------- single frame
        if (!anim_mode) {
            Frame f{arg_a, arg_theta, arg_phi, base_r_obs * M_bh, arg_disk_out * M_bh};
            kt_frame_info info{};
            {
                kthip::KNdSParams_HIP kp{M_bh, f.a, arg_Q, arg_Lam, 0, 0, f.disk_out};
                kthip::CameraParams_HIP cp{f.r_obs, 0, 0, 0, W, H, chart};
                kt_frame k = kthip::make_frame(kp, cp);
                k.theta_deg = f.theta; k.phi_deg = f.phi; k.fov_deg = arg_fov;
                const kt_opts o = kthip::make_opts(chart, ro, fp64);
                ctx.check(kt_hip_frame_info(ctx.get(), &k, 1, &o, &info), "kt_hip_frame_info");
            }
            std::cout << "KNdS  M=" << M_bh << " a=" << f.a << " Q=" << arg_Q << " Λ=" << arg_Lam << "\n"
                      << "  r₊=" << info.r_horizon << "  r_ISCO=" << info.r_isco << "\n"
                      << "Mode: " << (use_bundles ? "ray bundles" : "single ray") << "  chart="
                      << (chart == KT_CHART_KS ? "KS" : "BL") << "  "
                      << (integrator == KT_INTEGRATOR_DOPRI5 ? "DOPRI5" : "RK4-doubling") << "  standard  backend=gpu-hip"
                      << "  precision=" << (fp64 ? "fp64" : "mixed") << "  max-steps=" << max_steps
                      << "  step-init=" << step_init << "  tol=" << tol << "  camera-spp=" << camera_spp
                      << "  gpu-fp64=" << (gpu_fp64 ? "on" : "off") << "  " << W << "x" << H << "\n"
                      << "ColorParams: exp=" << exposure << " γ=" << gamma << " T×=" << temp_scale << "\n";
            kthip::KGeoMeta meta{(uint32_t)W, (uint32_t)H, M_bh, f.a, arg_Q, arg_Lam, info.r_isco, info.r_isco,
                                 f.disk_out, f.theta, f.phi, f.r_obs};
            mkdir_p(out_dir);
            const double t0 = now_s();
            std::vector<std::vector<uint8_t>> images;
            std::vector<kt_geopixel> geo;
            kt_stats st{};
            const bool want_geo = geo_only || (!geo_file.empty() && camera_spp == 1);
            if (geo_only && camera_spp > 1) camera_spp = 1;  // phase 1 traces the centre pass only
            if (!geo_file.empty() && camera_spp > 1 && !geo_only)
                std::cerr << "Info: camera-spp>1 uses jittered multi-pass sampling and disables .kgeo export for this frame.\n";
            render_frames({f}, images, want_geo ? &geo : nullptr, &st);
            const double elapsed = now_s() - t0;
            if (geo_only) {
                std::cout << "Trace: " << elapsed << "s  (" << std::fixed << std::setprecision(1)
                          << W * H / elapsed / 1e3 << " kpix/s)\n" << std::defaultfloat;
                print_stats(st);
                const std::string path = geo_file.empty()
                    ? out_dir + "/" + res_tag + "_" + std::to_string(W) + "x" + std::to_string(H) + "_" +
                          mode_backend_tag + "_" + timestamp() + ".kgeo"
                    : geo_file;
                if (!kthip::save_kgeo(path, geo, meta)) die("cannot write " + path);
                std::cout << "Geo saved: " << path << "\n";
                return 0;
            }
            std::cout << "Time: " << elapsed << "s  (" << std::fixed << std::setprecision(1)
                      << W * H / elapsed / 1e3 << " kpix/s)\n" << std::defaultfloat;
            print_stats(st);
            const std::string outfile = out_dir + "/" + res_tag + "_" + std::to_string(W) + "x" + std::to_string(H) +
                                        "_" + mode_backend_tag + "_" + timestamp() + ".png";
            if (!kthip::write_png(outfile, images[0].data(), W, H)) die("cannot write " + outfile);
            std::cout << "Saved: " << outfile << "\n";
            if (want_geo && !geo_only) {
                if (!kthip::save_kgeo(geo_file, geo, meta)) die("cannot write " + geo_file);
                std::cout << "Geo saved: " << geo_file << "\n";
            }
            return 0;
        }

        // ----------------------------------------------------------- animation
        auto pick = [](double v, double fb) { return std::isnan(v) ? fb : v; };
        const double theta_s = pick(th0, arg_theta), theta_e = pick(th1, arg_theta);
        const double phi_s = pick(ph0, arg_phi), phi_e = pick(ph1, arg_phi);
        const double r_s = pick(r0, base_r_obs), r_e = pick(r1, base_r_obs);
        const double a_s = pick(a0, arg_a), a_e = pick(a1, arg_a);
        const double do_s = pick(d0, arg_disk_out), do_e = pick(d1, arg_disk_out);
        std::ostringstream tag;
        tag << res_tag;
        if (anim_orbits != 0.0) tag << "_orbit" << anim_orbits << "x";
        else if (phi_s != phi_e) tag << "_phi" << (int)phi_s << "-" << (int)phi_e;
        if (theta_s != theta_e) tag << "_th" << (int)theta_s << "-" << (int)theta_e;
        if (r_s != r_e) tag << "_r" << (int)r_s << "-" << (int)r_e;
        if (a_s != a_e) tag << "_a" << a_s << "-" << a_e;
        tag << "_" << mode_backend_tag << "_" << anim_frames << "f" << anim_fps << "fps";
        const std::string frames_dir = anim_frames_dir.empty() ? out_dir + "/anim_" + tag.str() : anim_frames_dir;
        const std::string output_file = anim_output.empty() ? out_dir + "/" + tag.str() + ".mp4" : anim_output;
        mkdir_p(frames_dir);
        std::cout << "Animation: " << anim_frames << " frames @ " << anim_fps << " fps, " << anim_batch
                  << " frames per launch\n  output: " << output_file << "\n";

        std::vector<std::pair<int, Frame>> todo;
        int skipped = 0;
        for (int k = 0; k < anim_frames; ++k) {
            char fname[1024];
            std::snprintf(fname, sizeof(fname), "%s/frame_%05d.png", frames_dir.c_str(), k);
            if (anim_resume && file_exists(fname)) { ++skipped; continue; }
            const double phase = anim_frames == 1 ? 0.0 : (double)k / (anim_frames - 1);
            const double t = anim_ease ? smooth_step(phase) : phase;
            Frame f;
            f.a = a_s + (a_e - a_s) * t;
            f.theta = theta_s + (theta_e - theta_s) * t;
            f.r_obs = (r_s + (r_e - r_s) * t) * M_bh;
            f.disk_out = (do_s + (do_e - do_s) * t) * M_bh;
            f.phi = anim_orbits != 0.0 ? phi_s + 360.0 * anim_orbits * phase : lerp_angle(phi_s, phi_e, t);
            todo.emplace_back(k, f);
        }
        const double t_total = now_s();
        int rendered = 0;
        // Without camera-spp every batch goes through one plan: its render is captured in a HIP
        // graph on the first batch (after autotuning the launch on it) and replayed for the others.
        kt_hip_plan* plan = nullptr;
        struct PlanGuard { kt_hip_plan*& p; ~PlanGuard() { if (p) kt_hip_plan_destroy(p); } } plan_guard{plan};
        if (camera_spp <= 1 && !todo.empty()) {
            if (gpu_fp64 && !ctx.has_fp64())
                throw std::runtime_error("HIP FP64 strict mode requested, but the device does not expose native FP64.");
            kt_background bg{ro.bg_rgb, ro.bg_w, ro.bg_h};
            ctx.check(kt_hip_set_background(ctx.get(), ro.bg_rgb ? &bg : nullptr), "kt_hip_set_background");
            const kt_opts o = kthip::make_opts(chart, ro, fp64);
            const int nb = (int)std::min(todo.size(), (size_t)anim_batch);
            ctx.check(kt_hip_plan_create(ctx.get(), &o, nb, (int64_t)nb * W * H, &plan), "kt_hip_plan_create");
            std::cout << "Plan: " << nb << " frames per graph launch\n";
        }
        for (size_t b = 0; b < todo.size(); b += (size_t)anim_batch) {
            const size_t e = std::min(todo.size(), b + (size_t)anim_batch);
            std::vector<Frame> fr;
            for (size_t j = b; j < e; ++j) fr.push_back(todo[j].second);
            std::vector<std::vector<uint8_t>> images;
            kt_stats st{};
            const double t0 = now_s();
            if (plan) {
                const std::vector<kt_frame> kf = to_kt_frames(fr);
                const size_t npix = (size_t)W * H;
                std::vector<uint32_t> px(npix * kf.size());
                ctx.check(kt_hip_plan_render(plan, kf.data(), (int)kf.size(), px.data(), nullptr, nullptr, &st),
                          "kt_hip_plan_render");
                images.assign(fr.size(), std::vector<uint8_t>(npix * 3));
                for (size_t f = 0; f < fr.size(); ++f) kthip::abgr_to_rgb(px.data() + f * npix, npix, images[f].data());
            } else {
                render_frames(fr, images, nullptr, &st);
            }
            for (size_t j = b; j < e; ++j) {
                char fname[1024];
                std::snprintf(fname, sizeof(fname), "%s/frame_%05d.png", frames_dir.c_str(), todo[j].first);
                if (!kthip::write_png(fname, images[j - b].data(), W, H)) die(std::string("cannot write ") + fname);
                ++rendered;
            }
            std::cout << "Frames " << todo[b].first + 1 << "-" << todo[e - 1].first + 1 << "/" << anim_frames << "  ("
                      << std::fixed << std::setprecision(2) << now_s() - t0 << " s, trace " << st.trace_ms << " ms)\n"
                      << std::defaultfloat << std::flush;
        }
        std::cout << "\nFrames: " << rendered << " rendered";
        if (skipped) std::cout << ", " << skipped << " skipped";
        std::cout << "  (" << now_s() - t_total << "s)\n";
        if (!anim_no_encode) {
            char cmd[4096];
            std::snprintf(cmd, sizeof(cmd),
                          "ffmpeg -y -framerate %d -i \"%s/frame_%%05d.png\" -c:v libx264 -pix_fmt yuv420p -crf %d "
                          "-movflags +faststart \"%s\" 2>&1",
                          anim_fps, frames_dir.c_str(), anim_crf, output_file.c_str());
            std::cout << "Encoding...\n";
            if (std::system(cmd) == 0) {
                std::cout << "Video saved: " << output_file << "\n";
                if (!anim_keep_frames) {  // remove only the frames this run names, then the folder if empty
                    for (int k = 0; k < anim_frames; ++k) {
                        char fname[1024];
                        std::snprintf(fname, sizeof(fname), "%s/frame_%05d.png", frames_dir.c_str(), k);
                        std::remove(fname);
                    }
                    ::rmdir(frames_dir.c_str());
                }
            } else {
                std::cerr << "ffmpeg failed. Frames in: " << frames_dir << "\n";
            }
        }
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
}
