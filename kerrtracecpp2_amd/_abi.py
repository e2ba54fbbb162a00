"""ctypes mirror of include/kt_hip.h (the C ABI of libkt_hip.so)."""
import ctypes

import numpy as np

KT_OK = 0
KT_NOALLOC_MAX_FRAMES = 64
KT_NOALLOC_MAX_RAYS = 1 << 25
KT_E_ARG, KT_E_DEVICE, KT_E_NOFP64, KT_E_ALLOC, KT_E_LAUNCH, KT_E_UNSUPPORTED = -1, -2, -3, -4, -5, -6
KT_CHART_BL, KT_CHART_KS = 0, 1
KT_INTEGRATOR_RK4, KT_INTEGRATOR_DOPRI5 = 0, 1
KT_PRECISION_FP64, KT_PRECISION_MIXED = 0, 1
KT_PALETTE_BLACKBODY, KT_PALETTE_INTERSTELLAR = 0, 1
KT_QUIRK_DOPRI5_PHI_FROZEN = 1 << 0
KT_QUIRK_ESCAPE_POSITIONAL = 1 << 1
KT_QUIRK_KS_INIT_NO_SHIFT = 1 << 2
KT_QUIRK_BUNDLE_HORIZON_ESC = 1 << 3
KT_QUIRK_BUNDLE_ROOT_SKIP = 1 << 4
KT_QUIRK_KS_LAMBDA_TO_BL = 1 << 5
KT_QUIRKS_REFERENCE = 0x1F

GEOPIXEL_DTYPE = np.dtype([
    ("outcome", "u1"), ("pad", "u1", 3), ("r", "<f4"), ("redshift", "<f4"), ("magnif", "<f4"),
    ("phi_disk", "<f4"), ("theta_esc", "<f4"), ("phi_esc", "<f4")])


class KtFrame(ctypes.Structure):
    _fields_ = [("M", ctypes.c_double), ("a", ctypes.c_double), ("Q", ctypes.c_double),
                ("Lambda", ctypes.c_double), ("r_obs", ctypes.c_double), ("theta_deg", ctypes.c_double),
                ("phi_deg", ctypes.c_double), ("fov_deg", ctypes.c_double), ("disk_out", ctypes.c_double),
                ("offset_x", ctypes.c_double), ("offset_y", ctypes.c_double),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class KtOpts(ctypes.Structure):
    _fields_ = [("chart", ctypes.c_int32), ("integrator", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("bundles", ctypes.c_int32), ("max_steps", ctypes.c_int32), ("quirks", ctypes.c_uint32),
                ("step_init", ctypes.c_double), ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("exposure", ctypes.c_double), ("gamma", ctypes.c_double), ("temp_scale", ctypes.c_double),
                ("tuning_blocks_per_cu", ctypes.c_int32), ("tuning_occupancy", ctypes.c_int32),
                ("tuning_chunk", ctypes.c_int32), ("tuning_priority_steps", ctypes.c_int32),
                ("mixed_noise", ctypes.c_double), ("mixed_r_island", ctypes.c_double),
                ("palette", ctypes.c_int32), ("disk_rings", ctypes.c_int32), ("disk_sectors", ctypes.c_int32),
                ("mixed_hold", ctypes.c_int32), ("disk_sigma", ctypes.c_double), ("disk_hue_offset", ctypes.c_double)]


class KtBackground(ctypes.Structure):
    _fields_ = [("rgb", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class KtFrameInfo(ctypes.Structure):
    _fields_ = [("r_horizon", ctypes.c_double), ("r_isco", ctypes.c_double), ("r_disk_in", ctypes.c_double),
                ("r_disk_out", ctypes.c_double), ("r_escape", ctypes.c_double),
                ("eff_chart", ctypes.c_int32), ("eff_integrator", ctypes.c_int32)]


class KtKgeoMeta(ctypes.Structure):
    _fields_ = [("W", ctypes.c_uint32), ("H", ctypes.c_uint32), ("M", ctypes.c_double), ("a", ctypes.c_double),
                ("Q", ctypes.c_double), ("Lambda", ctypes.c_double), ("r_isco", ctypes.c_double),
                ("r_disk_in", ctypes.c_double), ("r_disk_out", ctypes.c_double), ("theta_deg", ctypes.c_double),
                ("phi_deg", ctypes.c_double), ("r_obs", ctypes.c_double)]


class KtTune(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_int32), ("chunk", ctypes.c_int32), ("occupancy", ctypes.c_int32),
                ("candidates", ctypes.c_int32), ("best_ms", ctypes.c_double), ("default_ms", ctypes.c_double),
                ("probe_rays", ctypes.c_int64), ("priority_steps", ctypes.c_int32), ("guarded", ctypes.c_int32),
                ("mixed_noise", ctypes.c_double), ("mixed_r_island", ctypes.c_double),
                ("binary32_share", ctypes.c_double), ("max_dg", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KtAdjointStats(ctypes.Structure):
    _fields_ = [("forward_ms", ctypes.c_double), ("record_ms", ctypes.c_double), ("backward_ms", ctypes.c_double),
                ("steps", ctypes.c_int64), ("checkpoints", ctypes.c_int64), ("record_bytes", ctypes.c_int64),
                ("rays_differentiated", ctypes.c_int64), ("seed_failures", ctypes.c_int64),
                ("steps_differentiated", ctypes.c_int64), ("reverse_entries", ctypes.c_int64),
                ("rays_parked", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KtStats(ctypes.Structure):
    _fields_ = [("trace_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("rays", ctypes.c_uint64),
                ("steps_accepted", ctypes.c_uint64), ("steps_rejected", ctypes.c_uint64),
                ("rhs_fp32", ctypes.c_uint64), ("rhs_fp64", ctypes.c_uint64), ("events", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# exported symbols and their prototypes: (restype, argtypes)
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
PROTOTYPES = {
    "kt_hip_abi_version": (ctypes.c_int, []),
    "kt_hip_default_opts": (None, [_P]),
    "kt_check_opts": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_size_t]),
    "kt_hip_create": (ctypes.c_int, [_I, _P]),
    "kt_hip_destroy": (None, [_P]),
    "kt_hip_last_error": (ctypes.c_char_p, [_P]),
    "kt_hip_device_info": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_size_t, _P, _P]),
    "kt_hip_set_background": (ctypes.c_int, [_P, _P]),
    "kt_hip_render": (ctypes.c_int, [_P, _P, _I, _P, _P, _P, _P, _P]),
    "kt_hip_render_device": (ctypes.c_int, [_P, _P, _I, _P, _P, _P, _P]),
    "kt_hip_render_device_noalloc": (ctypes.c_int, [_P, _P, _I, _P, _P, _P, _P]),
    "kt_hip_colorize": (ctypes.c_int, [_P, _P, ctypes.c_int64, _D, _D, _D, _D, _P, _P]),
    "kt_hip_frame_info": (ctypes.c_int, [_P, _P, _I, _P, _P]),
    "kt_make_checker_background": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_int32, _P]),
    "kt_hip_render_spp": (ctypes.c_int, [_P, _P, _I, _I, _P, _P, _P]),
    "kt_hip_autotune": (ctypes.c_int, [_P, _P, _I, _P, _P]),
    "kt_hip_render_intensity": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "kt_hip_render_adjoint": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "kt_hip_plan_create": (ctypes.c_int, [_P, _P, _I, ctypes.c_int64, _P]),
    "kt_hip_plan_render": (ctypes.c_int, [_P, _P, _I, _P, _P, _P, _P]),
    "kt_hip_plan_render_device": (ctypes.c_int, [_P, _P, _I, _P, _P, _P]),
    "kt_hip_plan_tuning": (ctypes.c_int, [_P, _P]),
    "kt_hip_plan_destroy": (None, [_P]),
    "kt_spp_offsets": (ctypes.c_int, [ctypes.c_int32, _P]),
    "kt_spp_resolve": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int64, _D, _P]),
    "kt_spp_tables": (ctypes.c_int, [_D, _P]),
    "kt_save_kgeo": (ctypes.c_int, [ctypes.c_char_p, _P, _P]),
    "kt_load_kgeo": (ctypes.c_int, [ctypes.c_char_p, _P, _P, ctypes.c_int64]),
    "kt_write_png": (ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int32, ctypes.c_int32]),
    "kt_read_image_rgb": (ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int64, _P, _P]),
}


def bind(lib):
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def default_opts(lib=None):
    """Reference CLI defaults (main.cpp:193-197, 2505-2517): BL, RK4 doubling, tol 1e-7."""
    o = KtOpts()
    o.chart = KT_CHART_BL
    o.integrator = KT_INTEGRATOR_RK4
    o.precision = KT_PRECISION_FP64
    o.bundles = 0
    o.max_steps = 500000
    o.quirks = KT_QUIRKS_REFERENCE
    o.step_init = 1.0
    o.rtol = 1e-7
    o.atol = 1e-7
    o.exposure = 1.0
    o.gamma = 2.2
    o.temp_scale = 1.0
    o.palette = KT_PALETTE_BLACKBODY
    o.disk_rings, o.disk_sectors, o.disk_sigma, o.disk_hue_offset = 7, 14, 0.5, 0.0   # main.cpp:76-79
    return o


def make_frame(width, height, a, Q=0.0, Lambda=0.0, theta=80.0, phi=0.0, r_obs=40.0, fov=30.0, disk_out=12.0,
               M=1.0, offset_x=0.0, offset_y=0.0):
    return KtFrame(float(M), float(a), float(Q), float(Lambda), float(r_obs), float(theta), float(phi), float(fov),
                   float(disk_out), float(offset_x), float(offset_y), int(width), int(height))
