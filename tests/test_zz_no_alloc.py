"""The allocation-free entry point, kt_hip_render_device_noalloc, against the entry point it is a
variant of, kt_hip_render_device, both called through the C ABI.

Binary64 renders are independent of the launch configuration and of the tile order, so the
allocation-free render must equal the ordinary one bit for bit (GeoPixels, ABGR pixels and the
step counts); mixed-precision renders depend on which rays share a wave, so those are compared
at the parity bars.  Each case runs two calls in a row, the second with a smaller input, since
every call on a GPU reuses that GPU's scratch array."""
import ctypes

import numpy as np
import pytest

import kerrtracecpp2_amd as kt
from kerrtracecpp2_amd import _abi
from tests._util import photon_ring_band

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP.hipMemcpy.restype = ctypes.c_int
    return _HIP


def _d2h(dst, src_ptr, nbytes):
    assert _hip().hipMemcpy(dst.ctypes.data_as(ctypes.c_void_p), src_ptr, nbytes, 2) == 0   # device to host


@pytest.fixture(scope="module")
def ctx():
    c = kt.Context(0)
    yield c
    c.close()


def _existing(ctx, frames, opts):
    """kt_hip_render_device: outputs in the context's device buffers, copied back here."""
    npix = sum(f.width * f.height for f in frames)
    arr = (_abi.KtFrame * len(frames))(*frames)
    d_rgba, d_geo = ctypes.c_void_p(), ctypes.c_void_p()
    st = _abi.KtStats()
    rc = ctx.lib.kt_hip_render_device(ctx.h, arr, len(frames), ctypes.byref(opts), ctypes.byref(d_rgba),
                                      ctypes.byref(d_geo), ctypes.byref(st))
    assert rc == 0, ctx.lib.kt_hip_last_error(ctx.h)
    geo = np.zeros(npix, _abi.GEOPIXEL_DTYPE)
    rgba = np.zeros(npix, np.uint32)
    _d2h(geo, d_geo, npix * 28)
    _d2h(rgba, d_rgba, npix * 4)
    return geo, rgba, st.as_dict()


def _noalloc(ctx, frames, opts, shade=True):
    """kt_hip_render_device_noalloc into caller-owned device buffers (torch allocates them here;
    the entry point itself allocates nothing).  The buffers start as 0xAB bytes, so a pixel the
    render does not write shows."""
    npix = sum(f.width * f.height for f in frames)
    geo_t = torch.full((npix * 28,), 0xAB, dtype=torch.uint8, device="cuda")
    rgba_t = torch.full((npix * 4,), 0xAB, dtype=torch.uint8, device="cuda") if shade else None
    torch.cuda.synchronize()
    arr = (_abi.KtFrame * len(frames))(*frames)
    st = _abi.KtStats()
    rc = ctx.lib.kt_hip_render_device_noalloc(ctx.h, arr, len(frames), ctypes.byref(opts),
                                              ctypes.c_void_p(rgba_t.data_ptr()) if shade else None,
                                              ctypes.c_void_p(geo_t.data_ptr()), ctypes.byref(st))
    assert rc == 0, ctx.lib.kt_hip_last_error(ctx.h)
    geo = geo_t.cpu().numpy().view(_abi.GEOPIXEL_DTYPE)
    rgba = rgba_t.cpu().numpy().view(np.uint32) if shade else None
    return geo, rgba, st.as_dict()


def _opts(chart, method, prec="fp64", tol=1e-9, atol=1e-12, **kw):
    o = kt.default_opts()
    o.chart = kt.KT_CHART_KS if chart == "ks" else kt.KT_CHART_BL
    o.integrator = kt.KT_INTEGRATOR_DOPRI5 if method == "dopri5" else kt.KT_INTEGRATOR_RK4
    o.precision = kt.KT_PRECISION_MIXED if prec == "mixed" else kt.KT_PRECISION_FP64
    o.rtol, o.atol = tol, (atol if method == "dopri5" else tol)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _same(a, b):
    ga, ra, sa = a
    gb, rb, sb = b
    assert ga.tobytes() == gb.tobytes()
    if ra is not None and rb is not None:
        assert np.array_equal(ra, rb)
    for k in ("rays", "steps_accepted", "steps_rejected", "rhs_fp64", "rhs_fp32", "events"):
        assert sa[k] == sb[k], k


def test_binary64_equals_render_device_two_calls(ctx):
    big = [kt.make_frame(160, 90, a=0.9, Q=0.3, Lambda=1e-5, theta=85, r_obs=40),
           kt.make_frame(96, 54, a=-0.6, Q=0.1, theta=60, r_obs=30, disk_out=15),
           kt.make_frame(64, 64, a=0.5, theta=80, r_obs=40, fov=45)]
    o = _opts("ks", "dopri5")
    first = _noalloc(ctx, big, o)
    _same(first, _existing(ctx, big, o))
    print("noalloc call 1:", first[2], flush=True)
    small = [kt.make_frame(48, 27, a=0.5, theta=80)]
    o2 = _opts("bl", "rk4", tol=1e-7)
    second = _noalloc(ctx, small, o2)
    _same(second, _existing(ctx, small, o2))
    assert second[2]["rays"] == 48 * 27
    # the first frame list again: the tile orders the scratch holds are reused
    _same(_noalloc(ctx, big, o), first)


def test_geo_only_and_bundles(ctx):
    fr = [kt.make_frame(64, 36, a=0.99, theta=89)]
    o = _opts("bl", "rk4", tol=1e-7, bundles=1)
    got = _noalloc(ctx, fr, o, shade=False)
    ref = _existing(ctx, fr, o)
    assert got[0].tobytes() == ref[0].tobytes()
    assert got[2]["steps_accepted"] == ref[2]["steps_accepted"]
    fr2 = [kt.make_frame(24, 16, a=0.99, theta=89)]
    _same(_noalloc(ctx, fr2, o), _existing(ctx, fr2, o))


def test_mixed_precision_at_the_parity_bars(ctx):
    W, H = 320, 180
    fr = [kt.make_frame(W, H, a=0.9, Q=0.3, Lambda=1e-5, theta=85, r_obs=40)]
    o = _opts("ks", "dopri5", prec="mixed")
    got, _, st = _noalloc(ctx, fr, o)
    ref, _, _ = _existing(ctx, fr, o)
    assert st["rhs_fp32"] > 0
    keep = ~photon_ring_band(ref, W, H)
    assert (got["outcome"] == ref["outcome"])[keep].mean() >= 0.999
    d = keep & (got["outcome"] == 1) & (ref["outcome"] == 1)
    assert np.abs(got["redshift"][d] - ref["redshift"][d]).max() <= 1e-4
    assert np.abs(got["r"][d] - ref["r"][d]).max() <= 1e-3
    small = [kt.make_frame(80, 45, a=0.9, Q=0.3, Lambda=1e-5, theta=85, r_obs=40)]
    got2, _, _ = _noalloc(ctx, small, o)
    ref2, _, _ = _existing(ctx, small, o)
    keep2 = ~photon_ring_band(ref2, 80, 45)
    assert (got2["outcome"] == ref2["outcome"])[keep2].mean() >= 0.999


def test_limits_and_refusals(ctx):
    lib = ctx.lib
    o = _opts("ks", "dopri5")
    geo_t = torch.zeros(64 * 28, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    g = ctypes.c_void_p(geo_t.data_ptr())
    many = (_abi.KtFrame * 65)(*([kt.make_frame(8, 8, a=0.5)] * 65))
    assert lib.kt_hip_render_device_noalloc(ctx.h, many, 65, ctypes.byref(o), None, g, None) == _abi.KT_E_ARG
    assert b"64" in lib.kt_hip_last_error(ctx.h)
    huge = (_abi.KtFrame * 1)(kt.make_frame(8192, 8192, a=0.5))
    assert lib.kt_hip_render_device_noalloc(ctx.h, huge, 1, ctypes.byref(o), None, g, None) == _abi.KT_E_ARG
    assert b"maximum" in lib.kt_hip_last_error(ctx.h)
    one = (_abi.KtFrame * 1)(kt.make_frame(8, 8, a=0.5))
    assert lib.kt_hip_render_device_noalloc(ctx.h, one, 1, ctypes.byref(o), None, None, None) == _abi.KT_E_ARG
    ob = _opts("bl", "rk4", prec="mixed", bundles=1)
    assert lib.kt_hip_render_device_noalloc(ctx.h, one, 1, ctypes.byref(ob), None, g, None) == _abi.KT_E_UNSUPPORTED
    # still usable after the refusals
    fr = [kt.make_frame(8, 8, a=0.5)]
    _same(_noalloc(ctx, fr, o), _existing(ctx, fr, o))
