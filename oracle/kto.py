"""ctypes bindings of the CPU FP64 oracle (oracle/libkt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or as the timed CPU baseline.
The product package never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkt_oracle.so")
REF_BIN = os.path.join(_HERE, "_ref", "kerr_tracer_ref")

GEOPIXEL_DTYPE = np.dtype([
    ("outcome", "u1"), ("pad", "u1", 3), ("r", "<f4"), ("redshift", "<f4"), ("magnif", "<f4"),
    ("phi_disk", "<f4"), ("theta_esc", "<f4"), ("phi_esc", "<f4")])
assert GEOPIXEL_DTYPE.itemsize == 28


class Frame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("chart", ctypes.c_int32),
                ("integrator", ctypes.c_int32), ("bundles", ctypes.c_int32), ("max_steps", ctypes.c_int32),
                ("step_init", ctypes.c_double), ("tol", ctypes.c_double),
                ("M", ctypes.c_double), ("a", ctypes.c_double), ("Q", ctypes.c_double),
                ("Lambda", ctypes.c_double), ("theta_deg", ctypes.c_double), ("phi_deg", ctypes.c_double),
                ("r_obs", ctypes.c_double), ("fov_deg", ctypes.c_double), ("disk_out", ctypes.c_double),
                ("offset_x", ctypes.c_double), ("offset_y", ctypes.c_double),
                ("quirks", ctypes.c_uint32), ("pad", ctypes.c_int32)]


# Reference behaviours (same bit values as KT_QUIRK_* of include/kt_hip.h); a cleared bit makes
# the oracle run the corrected behaviour.
QUIRK_DOPRI5_PHI_FROZEN = 1 << 0     # geodesic.hpp:300-310
QUIRK_ESCAPE_POSITIONAL = 1 << 1     # main.cpp:360,428-433,1572,1650,1659
QUIRK_KS_INIT_NO_SHIFT = 1 << 2      # main.cpp:1373-1430
QUIRK_BUNDLE_HORIZON_ESC = 1 << 3    # ray_bundle.hpp:426-428
QUIRK_BUNDLE_ROOT_SKIP = 1 << 4      # ray_bundle.hpp:357-365
QUIRKS_REFERENCE = 0x1F


class FrameInfo(ctypes.Structure):
    _fields_ = [("r_horizon", ctypes.c_double), ("r_isco", ctypes.c_double), ("r_disk_in", ctypes.c_double),
                ("r_disk_out", ctypes.c_double), ("r_escape", ctypes.c_double),
                ("eff_chart", ctypes.c_int32), ("eff_bundles", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        D, P, I = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
        L.kto_trace_pixels.argtypes = [P, P, P, ctypes.c_int64, P, P, P, P, P, P, P, I]
        L.kto_trace_frame.argtypes = [P, P, I]
        L.kto_frame_info_get.argtypes = [P, P]
        L.kto_colorize.argtypes = [P, ctypes.c_int64, D, D, D, D, D, P, I, I, P]
        L.kto_spp_offsets.argtypes = [I, P]
        L.kto_spp_resolve.argtypes = [P, I, ctypes.c_int64, D, P]
        for n in ("kto_metric_upper_bl", "kto_metric_lower_bl"):
            getattr(L, n).argtypes = [P, D, D, P]
        L.kto_metric_upper_ks.argtypes = [P, D, D, D, P]
        L.kto_hamiltonian_bl.argtypes = [P, P]
        L.kto_hamiltonian_bl.restype = D
        L.kto_rhs_bl.argtypes = [P, P, P]
        for n in ("kto_r_horizon", "kto_r_isco"):
            getattr(L, n).argtypes = [P]
            getattr(L, n).restype = D
        L.kto_omega_k.argtypes = [P, D]
        L.kto_omega_k.restype = D
        L.kto_disk_redshift.argtypes = [P, D, D, D]
        L.kto_disk_redshift.restype = D
        L.kto_camera_pixel.argtypes = [P, P, D, D, P]
        L.kto_camera_angles.argtypes = [P, P, D, D, P]
        L.kto_adaptive_step.argtypes = [P, I, P, P, P, D]
        L.kto_bl_to_ks.argtypes = [D, D, D, D, P]
        L.kto_ks_to_bl.argtypes = [D, D, D, D, P]
        L.kto_r_ks.argtypes = [D, D, D, D]
        L.kto_r_ks.restype = D
        L.kto_hermite_first_root.argtypes = [D, D, D, D, D, D, P]
        L.kto_hermite.argtypes = [D, D, D, D, D, D]
        L.kto_hermite.restype = D
        L.kto_blackbody.argtypes = [D, P]
        L.kto_colorize_interstellar.argtypes = [P, ctypes.c_int64, D, D, D, D, I, I, D, D, D, D, P, I, I, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def make_frame(width, height, a, Q=0.0, Lambda=0.0, theta=80.0, phi=0.0, r_obs=40.0, fov=30.0, disk_out=12.0,
               chart="bl", integrator="rk4", bundles=False, max_steps=500000, step_init=1.0, tol=1e-7, M=1.0,
               offset_x=0.0, offset_y=0.0, quirks=QUIRKS_REFERENCE):
    """Frame with the reference CLI's parameter meaning (degrees, main.cpp:2548-2684 clamps).
    quirks: mask of QUIRK_* bits (default: every reference behaviour).  Bit 5 of the HIP
    library (KS with Lambda != 0 run in BL) has no counterpart: the oracle always runs such
    a frame in BL, the chart-invariant outputs (fate, r, g, theta) being the comparison."""
    return Frame(int(width), int(height), 1 if chart == "ks" else 0, 1 if integrator == "dopri5" else 0,
                 1 if bundles else 0, max(1, int(max_steps)), max(1e-10, float(step_init)),
                 max(1e-10, float(tol)), float(M), float(a), float(Q), float(Lambda), float(theta), float(phi),
                 float(r_obs), float(fov), float(disk_out), float(offset_x), float(offset_y),
                 int(quirks) & QUIRKS_REFERENCE, 0)


def frame_info(frame):
    fi = FrameInfo()
    lib().kto_frame_info_get(ctypes.byref(frame), ctypes.byref(fi))
    return fi


def trace_pixels(frame, px, py, n_threads=0, extras=False):
    px = np.ascontiguousarray(px, dtype=np.int32)
    py = np.ascontiguousarray(py, dtype=np.int32)
    n = px.size
    out = np.zeros(n, dtype=GEOPIXEL_DTYPE)
    acc = np.zeros(n, np.int64)
    rej = np.zeros(n, np.int64)
    r64 = np.zeros(n)
    g64 = np.zeros(n)
    t64 = np.zeros(n)
    p64 = np.zeros(n)
    rc = lib().kto_trace_pixels(ctypes.byref(frame), _p(px), _p(py), n, _p(out), _p(acc), _p(rej), _p(r64),
                                _p(g64), _p(t64), _p(p64), int(n_threads))
    if rc != 0:
        raise RuntimeError("kto_trace_pixels failed")
    if extras:
        return out, dict(accepted=acc, rejected=rej, r=r64, g=g64, theta_esc=t64, phi_esc=p64)
    return out


def trace_frame(frame, n_threads=0):
    out = np.zeros(frame.width * frame.height, dtype=GEOPIXEL_DTYPE)
    if lib().kto_trace_frame(ctypes.byref(frame), _p(out), int(n_threads)) != 0:
        raise RuntimeError("kto_trace_frame failed")
    return out


def colorize(geo, M, r_isco, bg=None, exposure=1.0, gamma=2.2, temp_scale=1.0):
    geo = np.ascontiguousarray(geo)
    rgb = np.zeros((geo.size, 3), np.uint8)
    if bg is not None:
        bg = np.ascontiguousarray(bg, dtype=np.uint8)
        bh, bw = bg.shape[:2]
    else:
        bh = bw = 0
    lib().kto_colorize(_p(geo), geo.size, M, r_isco, exposure, gamma, temp_scale, _p(bg), bw, bh, _p(rgb))
    return rgb


def colorize_interstellar(geo, M, r_isco, r_disk_in, r_disk_out, rings=7, sectors=14, sigma=0.5, hue_offset=0.0,
                          bg=None, exposure=1.0, gamma=2.2):
    geo = np.ascontiguousarray(geo)
    rgb = np.zeros((geo.size, 3), np.uint8)
    if bg is not None:
        bg = np.ascontiguousarray(bg, dtype=np.uint8)
        bh, bw = bg.shape[:2]
    else:
        bh = bw = 0
    lib().kto_colorize_interstellar(_p(geo), geo.size, M, r_isco, r_disk_in, r_disk_out, rings, sectors, sigma,
                                    hue_offset, exposure, gamma, _p(bg), bw, bh, _p(rgb))
    return rgb


def spp_offsets(spp):
    out = np.zeros((max(1, spp), 2))
    lib().kto_spp_offsets(int(spp), _p(out))
    return out


def spp_resolve(passes, gamma=2.2):
    passes = np.ascontiguousarray(passes, dtype=np.uint8)
    n_pass, n_pix = passes.shape[0], passes.shape[1]
    out = np.zeros((n_pix, 3), np.uint8)
    lib().kto_spp_resolve(_p(passes), n_pass, n_pix, gamma, _p(out))
    return out


def _mp(M, a, Q, L):
    return np.array([M, a, Q, L], dtype=np.float64)


def upper_bl(mp, r, th):
    out = np.zeros(16)
    lib().kto_metric_upper_bl(_p(np.asarray(mp, np.float64)), r, th, _p(out))
    return out.reshape(4, 4)


def lower_bl(mp, r, th):
    out = np.zeros(16)
    lib().kto_metric_lower_bl(_p(np.asarray(mp, np.float64)), r, th, _p(out))
    return out.reshape(4, 4)


def upper_ks(mp, X, Y, Z):
    out = np.zeros(16)
    lib().kto_metric_upper_ks(_p(np.asarray(mp, np.float64)), X, Y, Z, _p(out))
    return out.reshape(4, 4)


def hamiltonian_bl(mp, s6):
    return lib().kto_hamiltonian_bl(_p(np.asarray(mp, np.float64)), _p(np.asarray(s6, np.float64)))


def rhs_bl(mp, s6):
    out = np.zeros(5)
    lib().kto_rhs_bl(_p(np.asarray(mp, np.float64)), _p(np.asarray(s6, np.float64)), _p(out))
    return out


def r_horizon(mp):
    return lib().kto_r_horizon(_p(np.asarray(mp, np.float64)))


def r_isco(mp):
    return lib().kto_r_isco(_p(np.asarray(mp, np.float64)))


def omega_k(mp, r):
    return lib().kto_omega_k(_p(np.asarray(mp, np.float64)), r)


def disk_redshift(mp, r, pt, pp):
    return lib().kto_disk_redshift(_p(np.asarray(mp, np.float64)), r, pt, pp)


def camera_pixel(mp, cam6, px, py):
    out = np.zeros(7)
    lib().kto_camera_pixel(_p(np.asarray(mp, np.float64)), _p(np.asarray(cam6, np.float64)), px, py, _p(out))
    return out


def camera_angles(mp, cam6, al, be):
    out = np.zeros(7)
    lib().kto_camera_angles(_p(np.asarray(mp, np.float64)), _p(np.asarray(cam6, np.float64)), al, be, _p(out))
    return out


def adaptive_step(mp, method, state7, h, fsal4=None, tol=1e-7):
    st = np.array(state7, dtype=np.float64)
    hh = np.array([h], dtype=np.float64)
    fs = np.full(4, np.nan) if fsal4 is None else np.array(fsal4, dtype=np.float64)
    ok = lib().kto_adaptive_step(_p(np.asarray(mp, np.float64)), 1 if method == "dopri5" else 0, _p(st), _p(hh),
                                 _p(fs), tol)
    return bool(ok), st, float(hh[0]), fs


def bl_to_ks(r, th, ph, a):
    out = np.zeros(3)
    lib().kto_bl_to_ks(r, th, ph, a, _p(out))
    return out


def ks_to_bl(X, Y, Z, a):
    out = np.zeros(3)
    lib().kto_ks_to_bl(X, Y, Z, a, _p(out))
    return out


def r_ks(X, Y, Z, a):
    return lib().kto_r_ks(X, Y, Z, a)


def hermite(y0, y1, d0, d1, h, al):
    return lib().kto_hermite(y0, y1, d0, d1, h, al)


def hermite_first_root(y0, y1, d0, d1, h, target):
    al = np.zeros(1)
    ok = lib().kto_hermite_first_root(y0, y1, d0, d1, h, target, _p(al))
    return bool(ok), float(al[0])


def blackbody(T):
    out = np.zeros(3, np.uint8)
    lib().kto_blackbody(T, _p(out))
    return out


def read_kgeo(path):
    """Parse a .kgeo file (main.cpp:96-137): magic, version, KGeoMeta (88 B), W*H GeoPixels."""
    b = open(path, "rb").read()
    if b[:4] != b"KGEO":
        raise ValueError("bad magic")
    ver = int(np.frombuffer(b[4:8], np.uint32)[0])
    W, H = (int(v) for v in np.frombuffer(b[8:16], np.uint32))
    meta = np.frombuffer(b[16:96], np.float64).copy()
    px = np.frombuffer(b[96:96 + 28 * W * H], GEOPIXEL_DTYPE).copy()
    return dict(version=ver, W=W, H=H, meta=meta, pixels=px)
