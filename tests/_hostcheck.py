"""ctypes loader of tests/hostcheck/libkt_hostcheck.so: the product's per-ray tracer
(kerrtracecpp2_amd/csrc/kt_trace.hpp) compiled for the host.  Test-only."""
import ctypes
import os
import subprocess

import numpy as np

from kerrtracecpp2_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hostcheck", "kt_hostcheck.cpp")
LIB = os.path.join(HERE, "hostcheck", "libkt_hostcheck.so")
_lib = None


def build():
    tmp = "%s.%d.tmp" % (LIB, os.getpid())      # test workers may build at once: rename atomically
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
                           "-o", tmp, SRC])
    os.replace(tmp, LIB)


def _stale():
    """The library is missing or older than its source or any header of the device code."""
    if not os.path.exists(LIB):
        return True
    csrc = os.path.join(os.path.dirname(HERE), "kerrtracecpp2_amd", "csrc")
    deps = [SRC] + [os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".hpp")]
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(d) > t for d in deps)


def lib():
    global _lib
    if _lib is None:
        if _stale():
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.ktc_trace_pixels.argtypes = [P, P, P, P, ctypes.c_int64, P, P, P, P, P]
        L.ktc_shade.argtypes = [P, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, P, ctypes.c_int, ctypes.c_int, P]
        L.ktc_isco_exact.argtypes = [ctypes.c_double] * 4 + [P, P]
        L.ktc_shade_interstellar.argtypes = [P, ctypes.c_int64] + [ctypes.c_double] * 4 + [ctypes.c_int] * 2 + \
            [ctypes.c_double] * 4 + [P, ctypes.c_int, ctypes.c_int, P]
        L.ktc_trace_bundle_pixels.argtypes = [P, P, P, P, ctypes.c_int64, P, P]
        L.ktc_sincos.argtypes = [P, ctypes.c_int64, P, P]
        L.ktc_adjoint_pixels.argtypes = [P, P, P, P, ctypes.c_int64, P, P, P, P]
        L.ktc_adjoint_pixels_seg.argtypes = [P, P, P, P, ctypes.c_int64, P, P, P, P]
        L.ktc_camera_pair.argtypes = [P, P, P, P, ctypes.c_int64, P, P]
        D = ctypes.c_double
        L.ktc_rhs_jacobian.argtypes = [ctypes.c_int, D, P, P, D, D, P]
        L.ktc_rhs.argtypes = [ctypes.c_int, D, P, P, D, D, P]
        L.ktc_rhs_vjp.argtypes = [ctypes.c_int, D, P, P, D, D, P, P]
        L.ktc_rhs_sc.argtypes = [ctypes.c_int, D, P, P, D, D, D, D, P]
        L.ktc_rhs_vjp_sc.argtypes = [ctypes.c_int, D, P, P, D, D, D, D, P, P]
        L.ktc_dp5_step.argtypes = [ctypes.c_int, D, P, P, D, D, D, P]
        L.ktc_dp5_step_adjoint.argtypes = [ctypes.c_int, D, P, P, D, D, D, P, P]
        L.ktc_replay_intensity.argtypes = [P, P, P, P, D, D, P, ctypes.c_int, P, P]
        L.ktc_ray_steps.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P, P, P]
        L.ktc_event_filters.argtypes = [P, ctypes.c_int64, P]
        L.ktc_renormalise.argtypes = [ctypes.c_int, ctypes.c_double, P, P, ctypes.c_int64, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def trace_pixels(frame, opts, px, py):
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    n = px.size
    out = np.zeros(n, _abi.GEOPIXEL_DTYPE)
    acc = np.zeros(n, np.int64)
    r32 = np.zeros(n, np.int64)
    r64 = np.zeros(n, np.int64)
    info = np.zeros(5)
    rc = lib().ktc_trace_pixels(ctypes.byref(frame), ctypes.byref(opts), _p(px), _p(py), n, _p(out), _p(acc),
                                _p(r32), _p(r64), _p(info))
    assert rc == 0
    return out, dict(accepted=acc, rhs32=r32, rhs64=r64, r_horizon=info[0], r_isco=info[1])


def shade(geo, M, r_isco, bg=None, exposure=1.0, gamma=2.2, temp_scale=1.0):
    geo = np.ascontiguousarray(geo)
    out = np.zeros(geo.size, np.uint32)
    bh, bw = (bg.shape[0], bg.shape[1]) if bg is not None else (0, 0)
    lib().ktc_shade(_p(geo), geo.size, M, r_isco, exposure, gamma, temp_scale,
                    _p(np.ascontiguousarray(bg)) if bg is not None else None, bw, bh, _p(out))
    return out


def isco_exact(M, a, Q, Lam):
    seq = np.zeros(1)
    seg = np.zeros(1)
    lib().ktc_isco_exact(M, a, Q, Lam, _p(seq), _p(seg))
    return float(seq[0]), float(seg[0])


def trace_bundle_pixels(frame, opts, px, py):
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    out = np.zeros(px.size, _abi.GEOPIXEL_DTYPE)
    acc = np.zeros(px.size, np.int64)
    assert lib().ktc_trace_bundle_pixels(ctypes.byref(frame), ctypes.byref(opts), _p(px), _p(py), px.size, _p(out),
                                         _p(acc)) == 0
    return out, acc


def sincos(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.zeros_like(x), np.zeros_like(x)
    lib().ktc_sincos(_p(x), x.size, _p(s), _p(c))
    return s, c


def adjoint_pixels(frame, opts, px, py, want_grad=True, segmented=False):
    """Intensity, dI/d(a, Q, Lambda, theta_deg, r_obs), accepted steps and (r_h, r_isco, r_esc,
    chart, ks_ref_init) of the host build of the adjoint (segmented: through the GPU's
    segment-parallel decomposition of the reverse pass)."""
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    n = px.size
    inten = np.zeros(n)
    grad = np.zeros((n, 5)) if want_grad else None
    nacc = np.zeros(n, np.int32)
    info = np.zeros(5)
    fn = lib().ktc_adjoint_pixels_seg if segmented else lib().ktc_adjoint_pixels
    fn(ctypes.byref(frame), ctypes.byref(opts), _p(px), _p(py), n, _p(inten), _p(grad), _p(nacc), _p(info))
    return inten, grad, nacc, info


def camera_pair(frame, opts, px, py):
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    a, b = np.zeros((px.size, 7)), np.zeros((px.size, 7))
    lib().ktc_camera_pair(ctypes.byref(frame), ctypes.byref(opts), _p(px), _p(py), px.size, _p(a), _p(b))
    return a, b


def rhs_jacobian(chart, M, par, y, pt, pp):
    J = np.zeros((4, 9))
    lib().ktc_rhs_jacobian(chart, M, _p(np.asarray(par, np.float64)), _p(np.asarray(y, np.float64)), pt, pp, _p(J))
    return J


def rhs_vjp(chart, M, par, y, pt, pp, mu):
    g = np.zeros(9)
    lib().ktc_rhs_vjp(chart, M, _p(np.asarray(par, np.float64)), _p(np.asarray(y, np.float64)), pt, pp,
                      _p(np.asarray(mu, np.float64)), _p(g))
    return g


def rhs(chart, M, par, y, pt, pp):
    f = np.zeros(4)
    lib().ktc_rhs(chart, M, _p(np.asarray(par, np.float64)), _p(np.asarray(y, np.float64)), pt, pp, _p(f))
    return f


def rhs_sc(chart, M, par, y, s, c, pt, pp):
    """The right-hand side with sin and cos of theta given as independent inputs (y[1] unused)."""
    f = np.zeros(4)
    lib().ktc_rhs_sc(chart, M, _p(np.asarray(par, np.float64)), _p(np.asarray(y, np.float64)), s, c, pt, pp, _p(f))
    return f


def rhs_vjp_sc(chart, M, par, y, s, c, pt, pp, mu):
    """rhs_vjp_sc's adjoints of (sin, cos) taken as independent inputs."""
    g = np.zeros(2)
    lib().ktc_rhs_vjp_sc(chart, M, _p(np.asarray(par, np.float64)), _p(np.asarray(y, np.float64)), s, c, pt, pp,
                         _p(np.asarray(mu, np.float64)), _p(g))
    return g


def dp5_step(chart, M, par, y, pt, pp, h):
    """One step of the forward pass's DOPRI5 arithmetic from y (sin/cos of y[1])."""
    out = np.zeros(4)
    lib().ktc_dp5_step(chart, M, _p(np.asarray(par, np.float64)), _p(np.asarray(y, np.float64)), pt, pp, h, _p(out))
    return out


def dp5_step_adjoint(chart, M, par, y, pt, pp, h, lam):
    """Adjoint of dp5_step: lam at the step end -> adjoints of (y, p_t, p_phi, a, Q, Lambda)."""
    out = np.zeros(9)
    lib().ktc_dp5_step_adjoint(chart, M, _p(np.asarray(par, np.float64)), _p(np.asarray(y, np.float64)), pt, pp,
                               h, _p(np.asarray(lam, np.float64)), _p(out))
    return out


def ray_steps(frame, opts, px, py, cap=200000):
    hs = np.zeros(cap)
    n = ctypes.c_int()
    fate = ctypes.c_int()
    y0c = np.zeros(6)
    lib().ktc_ray_steps(ctypes.byref(frame), ctypes.byref(opts), int(px), int(py), _p(hs), cap, ctypes.byref(n),
                        _p(y0c), ctypes.byref(fate))
    return hs[:n.value].copy(), y0c, fate.value


def replay_intensity(frame, opts, par, y0, pt, pp, hs, adjoint=False):
    inten = np.zeros(1)
    lam = np.zeros(9) if adjoint else None
    rc = lib().ktc_replay_intensity(ctypes.byref(frame), ctypes.byref(opts), _p(np.asarray(par, np.float64)),
                                    _p(np.asarray(y0, np.float64)), pt, pp, _p(np.ascontiguousarray(hs)), hs.size,
                                    _p(inten), _p(lam))
    return (None if rc else inten[0]), lam


def shade_interstellar(geo, M, r_isco, r_in, r_out, rings=7, sectors=14, sigma=0.5, hue=0.0, bg=None, exposure=1.0,
                       gamma=2.2):
    geo = np.ascontiguousarray(geo)
    out = np.zeros(geo.size, np.uint32)
    bgp = None if bg is None else np.ascontiguousarray(bg, np.uint8)
    lib().ktc_shade_interstellar(_p(geo), geo.size, M, r_isco, r_in, r_out, rings, sectors, sigma, hue, exposure,
                                 gamma, _p(bgp), 0 if bg is None else bg.shape[1], 0 if bg is None else bg.shape[0],
                                 _p(out))
    return out


def event_filters(cases):
    """cases: (n, 11) th0, th1, dth0, dth1, r0, r1, dr0, dr1, h, r_in, r_out -> (n, 4) int32:
    may_touch, crossing found, may_reach, crossing on the disk."""
    c = np.ascontiguousarray(cases, np.float64)
    out = np.zeros((c.shape[0], 4), np.int32)
    lib().ktc_event_filters(_p(c), c.shape[0], _p(out))
    return out


def renormalise(chart, M, par, states):
    """states: (n, 6) r, theta, p_r, p_theta, p_t, p_phi -> (n, 2) p_r, p_theta after the mixed
    path's constraint projection."""
    st = np.ascontiguousarray(states, np.float64)
    out = np.zeros((st.shape[0], 2))
    lib().ktc_renormalise(chart, M, _p(np.asarray(par, np.float64)), _p(st), st.shape[0], _p(out))
    return out
