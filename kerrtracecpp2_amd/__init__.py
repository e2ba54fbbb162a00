"""kerrtracecpp2_amd -- MI355X (gfx950) HIP backend for the KerrTraceCpp2 geodesic hot path.

The compute path is libkt_hip.so (hand-written HIP kernels behind the C ABI of
include/kt_hip.h).  This package is a thin Python host over that ABI: it loads
the in-tree library and raises if it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import (GEOPIXEL_DTYPE, KT_CHART_BL, KT_CHART_KS, KT_INTEGRATOR_DOPRI5, KT_INTEGRATOR_RK4,  # noqa: F401
                   KT_PALETTE_BLACKBODY, KT_PALETTE_INTERSTELLAR, KT_PRECISION_FP64, KT_PRECISION_MIXED,
                   KT_QUIRK_BUNDLE_HORIZON_ESC, KT_QUIRK_BUNDLE_ROOT_SKIP, KT_QUIRK_DOPRI5_PHI_FROZEN,
                   KT_QUIRK_ESCAPE_POSITIONAL, KT_QUIRK_KS_INIT_NO_SHIFT, KT_QUIRK_KS_LAMBDA_TO_BL,
                   KT_QUIRKS_REFERENCE, KtFrame, KtOpts, default_opts, make_frame)

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libkt_hip.so")

_lib = None


class KtError(RuntimeError):
    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


def load_library():
    """Load libkt_hip.so from the package directory (raises KtError if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KtError(f"HIP library not built: {LIB_PATH} missing (run __graft_entry__.build())")
        _lib = _abi.bind(ctypes.CDLL(LIB_PATH))
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Context:
    """One device context (kt_hip_ctx): owns device buffers, the stream and the background."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.kt_hip_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise KtError(f"kt_hip_create({device}) failed with status {rc}")
        self.h = h

    def close(self):
        if self.h:
            self.lib.kt_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.kt_hip_last_error(self.h)
            raise KtError(f"{what} failed ({rc}): {msg.decode() if msg else ''}", rc)

    def device_info(self):
        name = ctypes.create_string_buffer(256)
        ncu = ctypes.c_int()
        f64 = ctypes.c_int()
        self._check(self.lib.kt_hip_device_info(self.h, name, 256, ctypes.byref(ncu), ctypes.byref(f64)),
                    "kt_hip_device_info")
        return {"name": name.value.decode(), "n_cu": ncu.value, "fp64": bool(f64.value)}

    def set_background(self, rgb):
        """rgb: (h, w, 3) uint8 equirectangular image, or None."""
        if rgb is None:
            self._check(self.lib.kt_hip_set_background(self.h, None), "kt_hip_set_background")
            return
        rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
        bg = _abi.KtBackground(rgb.ctypes.data, rgb.shape[1], rgb.shape[0])
        self._check(self.lib.kt_hip_set_background(self.h, ctypes.byref(bg)), "kt_hip_set_background")

    def render(self, frames, opts, want_rgba=True, want_geo=True):
        """Render frames (list of KtFrame) -> dict(rgba, geo, info, stats), frames concatenated."""
        frames = list(frames)
        arr = (_abi.KtFrame * len(frames))(*frames)
        npix = sum(f.width * f.height for f in frames)
        rgba = np.zeros(npix, np.uint32) if want_rgba else None
        geo = np.zeros(npix, GEOPIXEL_DTYPE) if want_geo else None
        info = (_abi.KtFrameInfo * len(frames))()
        st = _abi.KtStats()
        self._check(self.lib.kt_hip_render(self.h, arr, len(frames), ctypes.byref(opts), _ptr(rgba), _ptr(geo),
                                           info, ctypes.byref(st)), "kt_hip_render")
        return {"rgba": rgba, "geo": geo, "info": [_info_dict(i) for i in info], "stats": st.as_dict()}

    def render_device(self, frames, opts, shade=True):
        """Render with outputs left in device memory; returns stats (the benchmark's timed call)."""
        frames = list(frames)
        arr = (_abi.KtFrame * len(frames))(*frames)
        d_rgba = ctypes.c_void_p()
        d_geo = ctypes.c_void_p()
        st = _abi.KtStats()
        self._check(self.lib.kt_hip_render_device(self.h, arr, len(frames), ctypes.byref(opts),
                                                  ctypes.byref(d_rgba) if shade else None, ctypes.byref(d_geo),
                                                  ctypes.byref(st)), "kt_hip_render_device")
        return st.as_dict()

    def frame_info(self, frames, opts):
        frames = list(frames)
        arr = (_abi.KtFrame * len(frames))(*frames)
        info = (_abi.KtFrameInfo * len(frames))()
        self._check(self.lib.kt_hip_frame_info(self.h, arr, len(frames), ctypes.byref(opts), info),
                    "kt_hip_frame_info")
        return [_info_dict(i) for i in info]

    def render_spp(self, frames, spp, opts):
        """camera-spp: every frame as spp jittered passes in one launch, resolved on the device.
        Returns dict(rgb=(sum W*H, 3) uint8, stats)."""
        frames = list(frames)
        arr = (_abi.KtFrame * len(frames))(*frames)
        npix = sum(f.width * f.height for f in frames)
        rgb = np.zeros((npix, 3), np.uint8)
        st = _abi.KtStats()
        self._check(self.lib.kt_hip_render_spp(self.h, arr, len(frames), int(spp), ctypes.byref(opts), _ptr(rgb),
                                               ctypes.byref(st)), "kt_hip_render_spp")
        return {"rgb": rgb, "stats": st.as_dict()}

    def autotune(self, frames, opts):
        """Time the trace kernel's launch configurations on a probe of these frames; the context
        then uses the fastest one for this option class.  Returns the kt_tune fields."""
        frames = list(frames)
        arr = (_abi.KtFrame * len(frames))(*frames)
        t = _abi.KtTune()
        self._check(self.lib.kt_hip_autotune(self.h, arr, len(frames), ctypes.byref(opts), ctypes.byref(t)),
                    "kt_hip_autotune")
        return t.as_dict()

    def render_intensity(self, frame, opts):
        """Disk intensity image (H, W) float64 of the binary64 DOPRI5 path, the quantity the
        adjoint differentiates."""
        I = np.zeros(frame.width * frame.height)
        st = _abi.KtStats()
        self._check(self.lib.kt_hip_render_intensity(self.h, ctypes.byref(frame), ctypes.byref(opts), _ptr(I),
                                                     ctypes.byref(st)), "kt_hip_render_intensity")
        return I.reshape(frame.height, frame.width), st.as_dict()

    def render_adjoint(self, frame, opts, weights=None, want_pixel_grads=False):
        """Gradient of sum_p w_p I_p with respect to (a, Q, Lambda, theta_obs [deg], r_obs).
        Returns dict(grad (5,), intensity (H, W), dI_dp (H, W, 5) or None, stats)."""
        n = frame.width * frame.height
        w = None if weights is None else np.ascontiguousarray(np.asarray(weights, np.float64).reshape(n))
        I = np.zeros(n)
        dI = np.zeros((n, 5)) if want_pixel_grads else None
        g = np.zeros(5)
        st = _abi.KtAdjointStats()
        self._check(self.lib.kt_hip_render_adjoint(self.h, ctypes.byref(frame), ctypes.byref(opts), _ptr(w), _ptr(I),
                                                   _ptr(dI), _ptr(g), ctypes.byref(st)), "kt_hip_render_adjoint")
        return {"grad": g, "intensity": I.reshape(frame.height, frame.width),
                "dI_dp": None if dI is None else dI.reshape(frame.height, frame.width, 5), "stats": st.as_dict()}

    def plan(self, opts, max_frames, max_pixels):
        """A graph-captured batched render plan (see Plan)."""
        return Plan(self, opts, max_frames, max_pixels)

    def colorize(self, geo, M, r_isco, opts, r_disk_in=None, r_disk_out=12.0):
        """Phase 2 (colorize_buffer): r_disk_in defaults to r_isco as the reference writes it."""
        geo = np.ascontiguousarray(geo, dtype=GEOPIXEL_DTYPE)
        rgba = np.zeros(geo.size, np.uint32)
        r_in = float(r_isco if r_disk_in is None else r_disk_in)
        self._check(self.lib.kt_hip_colorize(self.h, _ptr(geo), geo.size, float(M), float(r_isco), r_in,
                                             float(r_disk_out), ctypes.byref(opts), _ptr(rgba)), "kt_hip_colorize")
        return rgba


class Plan:
    """kt_hip_plan: device buffers for up to max_frames / max_pixels and, from the first render
    on, the whole render captured in one HIP graph; each render is a single graph launch."""

    def __init__(self, ctx, opts, max_frames, max_pixels):
        self.ctx = ctx
        self.lib = ctx.lib
        self.opts = opts
        h = ctypes.c_void_p()
        ctx._check(self.lib.kt_hip_plan_create(ctx.h, ctypes.byref(opts), int(max_frames), int(max_pixels),
                                               ctypes.byref(h)), "kt_hip_plan_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.kt_hip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, frames, want_rgba=True, want_geo=True):
        frames = list(frames)
        arr = (_abi.KtFrame * len(frames))(*frames)
        npix = sum(f.width * f.height for f in frames)
        rgba = np.zeros(npix, np.uint32) if want_rgba else None
        geo = np.zeros(npix, GEOPIXEL_DTYPE) if want_geo else None
        info = (_abi.KtFrameInfo * len(frames))()
        st = _abi.KtStats()
        self.ctx._check(self.lib.kt_hip_plan_render(self.h, arr, len(frames), _ptr(rgba), _ptr(geo), info,
                                                    ctypes.byref(st)), "kt_hip_plan_render")
        return {"rgba": rgba, "geo": geo, "info": [_info_dict(i) for i in info], "stats": st.as_dict()}

    def render_device(self, frames):
        frames = list(frames)
        arr = (_abi.KtFrame * len(frames))(*frames)
        st = _abi.KtStats()
        d_rgba, d_geo = ctypes.c_void_p(), ctypes.c_void_p()
        self.ctx._check(self.lib.kt_hip_plan_render_device(self.h, arr, len(frames), ctypes.byref(d_rgba),
                                                           ctypes.byref(d_geo), ctypes.byref(st)),
                        "kt_hip_plan_render_device")
        return st.as_dict()

    def tuning(self):
        t = _abi.KtTune()
        self.ctx._check(self.lib.kt_hip_plan_tuning(self.h, ctypes.byref(t)), "kt_hip_plan_tuning")
        return t.as_dict()


def _info_dict(i):
    return {k: getattr(i, k) for k, _ in i._fields_}


def check_opts(opts):
    """Validate an option set without a device: raises KtError with the status and message a
    render would return for it."""
    msg = ctypes.create_string_buffer(512)
    rc = load_library().kt_check_opts(ctypes.byref(opts), msg, len(msg))
    if rc != 0:
        raise KtError(f"kt_check_opts: {msg.value.decode()} (status {rc})", rc)


def checker_background(seed=1234, width=512, height=256, cells_x=32, cells_y=16):
    """Seeded cells_x x cells_y checkerboard on the celestial sphere (RGB8, equirectangular)."""
    lib = load_library()
    out = np.zeros((height, width, 3), np.uint8)
    rc = lib.kt_make_checker_background(int(seed), int(width), int(height), int(cells_x), int(cells_y), _ptr(out))
    if rc != 0:
        raise KtError("kt_make_checker_background failed")
    return out


def rgba_to_rgb(rgba):
    """ABGR uint32 (R in bits 0-7) -> (n, 3) uint8, as main.cpp:2438-2443 unpacks it."""
    rgba = np.asarray(rgba, dtype=np.uint32)
    return np.stack([(rgba & 0xFF), (rgba >> 8) & 0xFF, (rgba >> 16) & 0xFF], axis=-1).astype(np.uint8)


def spp_offsets(spp):
    """(spp, 2) sub-pixel offsets of the jittered passes (main.cpp:2240-2279)."""
    out = np.zeros((int(spp), 2), np.float64)
    if load_library().kt_spp_offsets(int(spp), _ptr(out)) != 0:
        raise KtError("kt_spp_offsets failed")
    return out


def spp_resolve(passes_abgr, gamma=2.2):
    """Host resolve of (n_pass, n_pix) ABGR passes -> (n_pix, 3) uint8."""
    p = np.ascontiguousarray(passes_abgr, dtype=np.uint32)
    out = np.zeros((p.shape[1], 3), np.uint8)
    if load_library().kt_spp_resolve(_ptr(p), p.shape[0], p.shape[1], float(gamma), _ptr(out)) != 0:
        raise KtError("kt_spp_resolve failed")
    return out


def spp_tables(gamma=2.2):
    """(256,) linear table and (256,) display thresholds used by the device resolve."""
    tab = np.zeros(512, np.float64)
    if load_library().kt_spp_tables(float(gamma), _ptr(tab)) != 0:
        raise KtError("kt_spp_tables failed")
    return tab[:256], tab[256:]


def save_kgeo(path, geo, meta):
    """Write a .kgeo file; meta: dict with W, H, M, a, Q, Lambda, r_isco, r_disk_in, r_disk_out,
    theta_deg, phi_deg, r_obs."""
    geo = np.ascontiguousarray(geo, dtype=GEOPIXEL_DTYPE)
    m = _abi.KtKgeoMeta(**meta)
    if geo.size != m.W * m.H:
        raise KtError("geo size does not match W*H")
    if load_library().kt_save_kgeo(os.fsencode(path), ctypes.byref(m), _ptr(geo)) != 0:
        raise KtError(f"cannot write {path}")


def load_kgeo(path):
    """Read a .kgeo file -> (geo, meta dict)."""
    lib = load_library()
    m = _abi.KtKgeoMeta()
    if lib.kt_load_kgeo(os.fsencode(path), ctypes.byref(m), None, 0) != 0:
        raise KtError(f"cannot read {path}")
    geo = np.zeros(m.W * m.H, GEOPIXEL_DTYPE)
    if lib.kt_load_kgeo(os.fsencode(path), ctypes.byref(m), _ptr(geo), geo.size) != 0:
        raise KtError(f"cannot read {path}")
    return geo, {k: getattr(m, k) for k, _ in m._fields_}


def write_png(path, rgb):
    """rgb: (h, w, 3) uint8."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    if load_library().kt_write_png(os.fsencode(path), _ptr(rgb), rgb.shape[1], rgb.shape[0]) != 0:
        raise KtError(f"cannot write {path}")


def read_image_rgb(path):
    """PNG / binary PPM -> (h, w, 3) uint8."""
    lib = load_library()
    w, h = ctypes.c_int32(), ctypes.c_int32()
    if lib.kt_read_image_rgb(os.fsencode(path), None, 0, ctypes.byref(w), ctypes.byref(h)) != 0:
        raise KtError(f"cannot read {path}")
    out = np.zeros((h.value, w.value, 3), np.uint8)
    if lib.kt_read_image_rgb(os.fsencode(path), _ptr(out), out.size, ctypes.byref(w), ctypes.byref(h)) != 0:
        raise KtError(f"cannot read {path}")
    return out
