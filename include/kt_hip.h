/*
 * kt_hip.h -- C ABI of the MI355X (gfx950) HIP backend for the KNdS geodesic tracer.
 *
 * Plain C: POD structs, pointers and sizes; no C++ or torch types cross this boundary.
 * Every entry point returns KT_OK (0) or a negative KT_E* code; the message of the last
 * failure on a context is available from kt_hip_last_error().
 *
 * Reference interface being replaced (KerrTraceCpp2 tree):
 *   cuda_render(KNdSParams_CUDA, CameraParams_CUDA, bool require_fp64)
 *       -> std::vector<uint32_t> ABGR                         gpu/cuda/tracer.cuh:11-26
 *   called from render_image()'s USE_CUDA branch                main.cpp:2413-2466
 *   GeoPixel (28 B) / .kgeo Phase-1 output                      main.cpp:84-137
 *   colorize_buffer() Phase-2 shading                           main.cpp:1817-1866
 */
#ifndef KT_HIP_H
#define KT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KT_HIP_ABI_VERSION 1

/* status codes */
#define KT_OK            0
#define KT_E_ARG        -1   /* invalid argument (null pointer, bad size, bad enum) */
#define KT_E_DEVICE     -2   /* no gfx950 device / HIP runtime error */
#define KT_E_NOFP64     -3   /* --gpu-fp64 strict mode requested but the device lacks FP64 */
#define KT_E_ALLOC      -4   /* device allocation failed */
#define KT_E_LAUNCH     -5   /* kernel launch or execution failed */
#define KT_E_UNSUPPORTED -6  /* option combination not provided by this backend */

/* enums (int32 in the structs) */
#define KT_CHART_BL 0          /* Boyer-Lindquist (r, theta, phi)                                */
#define KT_CHART_KS 1          /* ingoing Kerr-Schild; any Lambda (reference: Lambda == 0 only)  */
#define KT_INTEGRATOR_RK4 0    /* adaptive RK4 step doubling  (geodesic.hpp:204-257)             */
#define KT_INTEGRATOR_DOPRI5 1 /* Dormand-Prince 5(4) with FSAL (geodesic.hpp:312-414)           */
#define KT_PRECISION_FP64 0    /* the integration in binary64; only the step-size factor err^(-q)
                                  takes its log2/exp2 in binary32 (the accept test is binary64)  */
#define KT_PRECISION_MIXED 1   /* binary32 stage derivatives; binary64 state, accept test, constraint
                                  projection, and the steps near the horizon, the pole and turning
                                  points.  Jacobi bundles: binary64 only (KT_E_UNSUPPORTED)       */
#define KT_PALETTE_BLACKBODY 0     /* disk_colour (main.cpp:1799-1814)                              */
#define KT_PALETTE_INTERSTELLAR 1  /* disk_colour_interstellar (main.cpp:1689-1797)                 */

/* reference behaviours that a drop-in must reproduce (all set by default, KT_QUIRKS_REFERENCE) */
#define KT_QUIRK_DOPRI5_PHI_FROZEN   (1u << 0) /* DOPRI5 never advances phi (geodesic.hpp:300-310)        */
#define KT_QUIRK_ESCAPE_POSITIONAL   (1u << 1) /* escape fills {phi_disk,theta_esc,phi_esc} =
                                                  {theta,phi,0} (main.cpp:360,428-433,1572,1650,1659)     */
#define KT_QUIRK_KS_INIT_NO_SHIFT    (1u << 2) /* KS initial covector keeps the BL spatial components
                                                  (main.cpp:1373-1430)                                    */
#define KT_QUIRK_BUNDLE_HORIZON_ESC  (1u << 3) /* bundle horizon/reject-cap -> outcome 0, angles 0
                                                  (ray_bundle.hpp:426-428)                                */
#define KT_QUIRK_BUNDLE_ROOT_SKIP    (1u << 4) /* bundle step whose Hermite root search fails skips the
                                                  horizon/escape tests (ray_bundle.hpp:357-365)           */
#define KT_QUIRK_KS_LAMBDA_TO_BL     (1u << 5) /* KS with Lambda != 0 runs in BL (main.cpp:2034-2041);
                                                  NOT in KT_QUIRKS_REFERENCE: this backend has a KNdS KS  */
#define KT_QUIRKS_REFERENCE 0x1Fu

/* Per-pixel Phase-1 record, byte-identical to GeoPixel (main.cpp:84-94). */
typedef struct kt_geopixel {
    uint8_t outcome;   /* 0 escaped, 1 disk, 2 horizon */
    uint8_t pad[3];
    float r, redshift, magnif, phi_disk, theta_esc, phi_esc;
} kt_geopixel;

/* One frame: physics + camera.  Angles in DEGREES, as on the reference CLI (main.cpp:2548-2684). */
typedef struct kt_frame {
    double M, a, Q, Lambda;
    double r_obs, theta_deg, phi_deg, fov_deg;
    double disk_out;             /* outer disk radius; inner radius is r_isco (main.cpp:2080-2081) */
    double offset_x, offset_y;   /* sub-pixel camera offsets (camera.hpp:38-47)                    */
    int32_t width, height;
} kt_frame;

/* Integration / precision / shading options shared by all frames of a call. */
typedef struct kt_opts {
    int32_t chart;          /* KT_CHART_*                                                   */
    int32_t integrator;     /* KT_INTEGRATOR_*                                              */
    int32_t precision;      /* KT_PRECISION_*; --gpu-fp64 selects FP64                        */
    int32_t bundles;        /* 1: Jacobi ray bundles (ray_bundle.hpp), BL chart              */
    int32_t max_steps;      /* accepted-step cap (main.cpp:194)                              */
    uint32_t quirks;        /* KT_QUIRK_* mask                                              */
    double step_init;       /* initial affine step (main.cpp:195)                            */
    double rtol, atol;      /* RK4: err < rtol (absolute, geodesic.hpp:234);
                               DOPRI5: scale = atol + rtol*|y0| (reference: atol = rtol = tol);
                               atol below 1.1754943508222875e-38 (FLT_MIN) is raised to it   */
    double exposure, gamma, temp_scale;   /* ColorParams (main.cpp:72-81)                   */
    int32_t tuning_blocks_per_cu;         /* persistent grid = blocks_per_cu x CUs (256-thread
                                             blocks); 0: autotuned / built-in default         */
    int32_t tuning_occupancy;             /* waves per SIMD the trace kernel's register budget
                                             is sized for: 0 autotuned / built-in default
                                             (geodesics 3, bundles 1), 2 or 3 (geodesics),
                                             1 or 2 (bundles)                                   */
    int32_t tuning_chunk;                 /* rays per work-queue grab; 0: 64                   */
    int32_t tuning_priority_steps;        /* wave issue priority rises one level per this many
                                             accepted steps of the wave's longest-running ray
                                             (levels 0-3); 0: 48; negative: off               */
    double mixed_noise;     /* MIXED: binary32 stage-noise coefficient of the island test;
                               0 selects the built-in default                               */
    double mixed_r_island;  /* MIXED: binary64 steps below r < mixed_r_island * r_horizon;
                               0 selects the built-in default                               */
    int32_t palette;        /* KT_PALETTE_*: disk colouring (ColorParams::palette, main.cpp:70-81) */
    int32_t disk_rings;     /* Interstellar palette: rings x sectors cells (main.cpp:76-77)   */
    int32_t disk_sectors;
    int32_t mixed_hold;     /* MIXED: > 0: a rejected binary32 attempt is retried at the same step
                               size in binary64 and the ray stays in binary64 for this many
                               accepted steps (fewer attempts, far fewer binary32 steps);
                               0 or < 0: no retry (the built-in policy)                     */
    double disk_sigma;      /* Interstellar Gaussian blend width in cells (main.cpp:78).  With
                               s = ceil(3 sigma) + 1, (rings + 2s + 1)(sectors + 2s + 1) must
                               stay within 65536 cells per pixel, else KT_E_ARG            */
    double disk_hue_offset; /* Interstellar hue offset (main.cpp:79)                          */
} kt_opts;

/* Background: RGB8 equirectangular image (BackgroundImage, main.cpp:140-171). */
typedef struct kt_background {
    const uint8_t* rgb;     /* host pointer, w*h*3 bytes, may be NULL (black sky) */
    int32_t width, height;
} kt_background;

/* Per-frame scalars as computed on the device (KGeoMeta fields + horizon). */
typedef struct kt_frame_info {
    double r_horizon, r_isco, r_disk_in, r_disk_out, r_escape;
    int32_t eff_chart, eff_integrator;
} kt_frame_info;

/* Counters gathered on the device during a render. */
typedef struct kt_stats {
    double trace_ms;          /* kernel time of the geodesic launch (HIP events)           */
    double total_ms;          /* setup + trace + shade, device time                        */
    uint64_t rays;            /* primary rays traced                                       */
    uint64_t steps_accepted;  /* accepted integrator steps                                  */
    uint64_t steps_rejected;  /* rejected attempts                                          */
    uint64_t rhs_fp32;        /* right-hand-side evaluations in binary32                    */
    uint64_t rhs_fp64;        /* right-hand-side evaluations in binary64                    */
    uint64_t events;          /* Hermite disk searches run                                  */
} kt_stats;

/* Configuration chosen by the autotuner (kt_hip_autotune): the launch (grid, queue chunk,
 * register budget), the wave issue-priority steps and, for mixed precision, the island
 * threshold and radius.  Mixed candidates are kept only if, on the probe, they agree with a
 * binary64 render of it at the parity bars (fate >= 99.9 % and |dg| <= 1e-4 off the band of
 * fate edges, steep redshift and disk-edge images) and keep binary32 for at least 50 % of the
 * right-hand sides (or the built-in policy's share where that is lower).  Renders whose
 * corresponding kt_opts fields are 0 use the result. */
typedef struct kt_tune {
    int32_t blocks_per_cu;    /* persistent grid = blocks_per_cu x CUs                        */
    int32_t chunk;            /* rays per work-queue grab                                      */
    int32_t occupancy;        /* waves per SIMD of the register budget                         */
    int32_t candidates;       /* configurations timed                                          */
    double best_ms;           /* trace time of the chosen configuration on the probe           */
    double default_ms;        /* trace time of the built-in configuration on the probe         */
    int64_t probe_rays;       /* rays of the probe workload                                    */
    int32_t priority_steps;   /* accepted steps per wave issue-priority level (< 0: off)       */
    int32_t guarded;          /* MIXED: candidates refused by the accuracy / binary32 guard    */
    double mixed_noise;       /* MIXED: island threshold chosen (0 for binary64 and bundles)   */
    double mixed_r_island;    /* MIXED: island radius chosen, in horizon radii                 */
    double binary32_share;    /* MIXED: binary32 share of right-hand sides on the probe        */
    double max_dg;            /* MIXED: max |dg| against a binary64 render of the probe, off the
                                 parity band; fate agreement there was >= 99.9 %              */
} kt_tune;

typedef struct kt_hip_ctx kt_hip_ctx;
typedef struct kt_hip_plan kt_hip_plan;

/* Library / context life cycle. */
int  kt_hip_abi_version(void);
void kt_hip_default_opts(kt_opts* opts);  /* reference defaults: BL, RK4, tol 1e-7, step 1, 500000 */
/* Validates an option set without a device (the checks every render makes first; main.cpp's
   argument handling, :2548-2684, has none): KT_OK or the KT_E_* status a render would return,
   with its message copied to msg (nullable, NUL-terminated within msg_len). */
int  kt_check_opts(const kt_opts* opts, char* msg, size_t msg_len);
int  kt_hip_create(int device, kt_hip_ctx** out);
void kt_hip_destroy(kt_hip_ctx* ctx);
const char* kt_hip_last_error(const kt_hip_ctx* ctx);
/* device name, CU count, FP64 capability; name_out may be NULL */
int  kt_hip_device_info(kt_hip_ctx* ctx, char* name_out, size_t name_len, int* n_cu, int* has_fp64);

/* Set the background used by escaped rays (copied to the device). */
int kt_hip_set_background(kt_hip_ctx* ctx, const kt_background* bg);

/* Render n_frames frames in one persistent launch (host buffers).
 * rgba_out:  sum(W*H) uint32 ABGR (R in bits 0-7, alpha 0xFF; tracer.cu:482), may be NULL
 * geo_out:   sum(W*H) GeoPixel records, may be NULL
 * info_out:  n_frames kt_frame_info, may be NULL
 * stats:     may be NULL
 * Frames are laid out back to back in frame order.                                            */
int kt_hip_render(kt_hip_ctx* ctx, const kt_frame* frames, int n_frames, const kt_opts* opts,
                  uint32_t* rgba_out, kt_geopixel* geo_out, kt_frame_info* info_out, kt_stats* stats);

/* Same as kt_hip_render, but outputs stay in device memory owned by the context
 * (valid until the next call on the context); the timed path of the benchmark.
 * d_rgba / d_geo receive device pointers (may be NULL).                                      */
int kt_hip_render_device(kt_hip_ctx* ctx, const kt_frame* frames, int n_frames, const kt_opts* opts,
                         void** d_rgba, void** d_geo, kt_stats* stats);

/* Allocation-free variant of kt_hip_render_device.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call allocates nothing: the caller passes device output
 * buffers, d_geo (28 bytes per pixel, required) and d_rgba (4 bytes per pixel, or NULL to skip
 * the shading).  The queue and statistics words, the frame table, the per-frame setup scratch and
 * the tile orders live in one fixed device array of about 2.3 MB per GPU, held until the process
 * ends; the trace kernel generates each camera ray itself (no per-pixel init records).  Limits:
 * n_frames <= KT_NOALLOC_MAX_FRAMES and the frames' 8x8-tile padded ray count, the sum of
 * ceil(W/8)*ceil(H/8)*64, <= KT_NOALLOC_MAX_RAYS; beyond them the call returns KT_E_ARG.  The
 * launch uses the default register budget (tuning_occupancy is ignored).  Every call on a GPU
 * shares that GPU's array, which keeps the last call's contents: this library serialises its own
 * calls with a process-wide lock, and calls that overlap on one GPU by any other route race.   */
#define KT_NOALLOC_MAX_FRAMES 64
#define KT_NOALLOC_MAX_RAYS (1ll << 25)
int kt_hip_render_device_noalloc(kt_hip_ctx* ctx, const kt_frame* frames, int n_frames, const kt_opts* opts,
                                 void* d_rgba, void* d_geo, kt_stats* stats);

/* Phase 2 on the device: GeoPixel records -> ABGR (colorize_buffer, main.cpp:1817-1866), with the
 * same scalars colorize_buffer takes (M, r_isco, r_disk_in, r_disk_out) and the palette of opts. */
int kt_hip_colorize(kt_hip_ctx* ctx, const kt_geopixel* geo, int64_t n, double M, double r_isco,
                    double r_disk_in, double r_disk_out, const kt_opts* opts, uint32_t* rgba_out);

/* Per-frame scalars only (setup kernel): horizon, ISCO, escape radius. */
int kt_hip_frame_info(kt_hip_ctx* ctx, const kt_frame* frames, int n_frames, const kt_opts* opts,
                      kt_frame_info* info_out);

/* Procedural background: cells_x x cells_y checkerboard, colours from a seeded generator. */
int kt_make_checker_background(uint64_t seed, int32_t width, int32_t height, int32_t cells_x,
                               int32_t cells_y, uint8_t* rgb_out);

/* --camera-spp (render_image's multi-pass branch, main.cpp:2232-2322).
 * Every frame is expanded into spp passes whose sub-pixel offsets (kt_spp_offsets) are added to the
 * frame's own offsets; all passes of all frames are traced in ONE persistent launch, shaded, and
 * averaged on the device in linear light: mean_s (c_s/255)^gamma, back through ^(1/gamma) and
 * round-to-nearest, gamma = max(opts->gamma, 1e-6).  The result equals the reference's host
 * resolve bit for bit (the display rounding is evaluated through per-level thresholds computed on
 * the host with the same pow).  rgb_out: sum(W*H*3) bytes, frames back to back.               */
int kt_hip_render_spp(kt_hip_ctx* ctx, const kt_frame* frames, int n_frames, int spp, const kt_opts* opts,
                      uint8_t* rgb_out, kt_stats* stats);

/* Autotuner (no reference counterpart; the reference has no GPU launch parameters to tune).
 * Times the trace kernel for blocks_per_cu in {2,3,4,6} x chunk in {32,64,128} on a probe
 * workload: the given frames scaled down (same fov and aspect) to about 2 M rays, or the
 * frames themselves when smaller.  The fastest configuration is returned and remembered by
 * the context for this (chart, integrator, precision, bundles) class: later renders whose
 * kt_opts tuning fields are 0 use it.  Results do not depend on the configuration (FP64). */
int kt_hip_autotune(kt_hip_ctx* ctx, const kt_frame* frames, int n_frames, const kt_opts* opts, kt_tune* out);

/* Batched render plan (no reference counterpart; the reference renders animation frames one
 * at a time, main.cpp:2994-3041).  A plan owns device buffers for up to max_frames frames and
 * max_pixels pixels and, on its first render, captures the whole render -- frame-table upload
 * from pinned staging, counter fill, setup, persistent trace, shade -- in one HIP graph
 * (autotuning first when the tuning fields of opts are 0).  Every render, with any frame list
 * that fits, is then a single graph launch.  Frames must share the effective chart.
 * stats->trace_ms and total_ms are the graph's execution time.                              */
int kt_hip_plan_create(kt_hip_ctx* ctx, const kt_opts* opts, int max_frames, int64_t max_pixels,
                       kt_hip_plan** out);
int kt_hip_plan_render(kt_hip_plan* plan, const kt_frame* frames, int n_frames, uint32_t* rgba_out,
                       kt_geopixel* geo_out, kt_frame_info* info_out, kt_stats* stats);
/* Outputs stay on the device (valid until the next render of the plan). */
int kt_hip_plan_render_device(kt_hip_plan* plan, const kt_frame* frames, int n_frames, void** d_rgba,
                              void** d_geo, kt_stats* stats);
int kt_hip_plan_tuning(const kt_hip_plan* plan, kt_tune* out);
void kt_hip_plan_destroy(kt_hip_plan* plan);

/* Differentiable rendering (no reference counterpart).  The per-pixel disk intensity of the
 * shading, I = PT(r_hit)/PT(3 r_isco) g^4 (1 + 0.85 max(0, 1 - g)) with g clamped to [0.1, 10]
 * (main.cpp:1799-1810, magnification 1; 0 off the disk), of the binary64 DOPRI5 path (one frame;
 * opts->integrator must be KT_INTEGRATOR_DOPRI5, precision is forced to FP64, no bundles).     */
int kt_hip_render_intensity(kt_hip_ctx* ctx, const kt_frame* frame, const kt_opts* opts, double* intensity_out,
                            kt_stats* stats);

typedef struct kt_adjoint_stats {
    double forward_ms;        /* forward pass to 128 accepted steps: step counts, intensities,
                                 step sizes, checkpoints; rays still running then are parked */
    double record_ms;         /* that first part again, when a ray outgrew the record stride or
                                 the reverse pass the context's entry buffer (0 otherwise)     */
    double backward_ms;       /* reverse pass, with the parked rays' forward steps finished
                                 beside its first half on a second stream                      */
    int64_t steps;            /* accepted steps recorded                                       */
    int64_t checkpoints;      /* state checkpoints stored (one per 16 accepted steps)          */
    int64_t record_bytes;     /* device bytes of the step sizes and checkpoints                */
    int64_t rays_differentiated;
    int64_t seed_failures;    /* disk pixels whose last step gave no root on replay            */
    int64_t steps_differentiated; /* accepted steps of the disk pixels, swept by the reverse pass */
    int64_t reverse_entries;  /* reverse-pass entries: checkpoint segments (four sweeps each,
                                 one per basis direction of the adjoint at the segment end)
                                 and one event entry per disk pixel                            */
    int64_t rays_parked;      /* rays still running after the forward pass's first part        */
} kt_adjoint_stats;

/* Hand-written adjoint of the adaptive integrator + disk shading.  Gradients of the per-pixel
 * intensity with respect to p = (a, Q, Lambda, theta_obs [degrees], r_obs): one forward pass
 * stores every accepted step size and a checkpoint every 16 accepted steps at a fixed stride per
 * pixel (the context keeps these buffers between calls; a ray longer than the stride makes the
 * stride grow and the pass run once more), and the reverse pass replays each segment from its
 * checkpoint and propagates the adjoint back through the DOPRI5 stages (accepted step sizes
 * held fixed).  The reverse pass runs the segments in parallel: each segment is swept once
 * per basis direction of the adjoint at its end, and each pixel's chain is then composed from
 * these small linear maps, last segment first.
 *   weights:       W*H values w (dL/dI), NULL for all ones
 *   intensity_out: W*H (may be NULL);  dI_dp_out: W*H*5, pixel-major (may be NULL)
 *   grad_out:      5 values, sum_p w_p dI_p/dp (summed on the device in a fixed order)      */
int kt_hip_render_adjoint(kt_hip_ctx* ctx, const kt_frame* frame, const kt_opts* opts, const double* weights,
                          double* intensity_out, double* dI_dp_out, double* grad_out, kt_adjoint_stats* stats);

/* Host utilities (no device work; callable without a GPU). */
/* Offsets of the spp jittered passes, (x, y) pairs in pixels (build_offsets, main.cpp:2240-2279). */
int kt_spp_offsets(int32_t spp, double* xy_out);
/* Host resolve of n_pass packed-ABGR passes (pass-major) into RGB8, same rule as kt_hip_render_spp. */
int kt_spp_resolve(const uint32_t* passes, int32_t n_pass, int64_t n_pix, double gamma, uint8_t* rgb_out);
/* The 512-entry table kt_hip_render_spp uploads: [0,256) (v/255)^gamma, [256,512) the least x in
 * [0,1] whose display level lround(255 x^(1/gamma)) is >= k ([256] = -inf).                      */
int kt_spp_tables(double gamma, double* tab_out);

/* .kgeo container (save_kgeo / load_kgeo, main.cpp:96-137): "KGEO", uint32 version 1, this
 * 88-byte header, then W*H GeoPixel records.  theta/phi in degrees as the reference writes them. */
typedef struct kt_kgeo_meta {
    uint32_t W, H;
    double M, a, Q, Lambda;
    double r_isco, r_disk_in, r_disk_out;
    double theta_deg, phi_deg, r_obs;
} kt_kgeo_meta;
int kt_save_kgeo(const char* path, const kt_kgeo_meta* meta, const kt_geopixel* geo);
/* geo_out may be NULL to read the header only; geo_cap is its capacity in records. */
int kt_load_kgeo(const char* path, kt_kgeo_meta* meta, kt_geopixel* geo_out, int64_t geo_cap);
/* 8-bit RGB PNG writer (stbi_write_png's role at main.cpp:2481-2484). */
int kt_write_png(const char* path, const uint8_t* rgb, int32_t width, int32_t height);
/* Background reader: PNG (8-bit, non-interlaced) or binary PPM -> RGB8 (stbi_load's role at
 * main.cpp:145-152; JPEG is not read).  rgb_out may be NULL to query the size.                */
int kt_read_image_rgb(const char* path, uint8_t* rgb_out, int64_t cap_bytes, int32_t* width, int32_t* height);

#ifdef __cplusplus
}
#endif
#endif /* KT_HIP_H */
