"""libkt_hip.so loads without a GPU and exports exactly what include/kt_hip.h declares."""
import ctypes
import os
import re

import numpy as np
import pytest

import kerrtracecpp2_amd as kt
from kerrtracecpp2_amd import _abi
from tests._util import ROOT


def _declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(kt_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = kt.load_library()
    decl = _declared()
    assert "kt_hip_render" in decl and "kt_hip_create" in decl
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) <= set(_abi.PROTOTYPES) | set(decl)
    assert set(_abi.PROTOTYPES) <= set(decl)


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_abi.KtFrame) == 11 * 8 + 2 * 4
    assert ctypes.sizeof(_abi.KtOpts) == 6 * 4 + 6 * 8 + 4 * 4 + 2 * 8 + 4 * 4 + 2 * 8
    assert (_abi.KtOpts.palette.offset, _abi.KtOpts.disk_hue_offset.offset) == (104, 128)
    assert _abi.GEOPIXEL_DTYPE.itemsize == 28


def test_host_only_entry_points_without_gpu():
    lib = kt.load_library()
    assert lib.kt_hip_abi_version() == 1
    o = _abi.KtOpts()
    lib.kt_hip_default_opts(ctypes.byref(o))
    assert (o.chart, o.integrator, o.precision, o.max_steps) == (0, 0, 0, 500000)
    assert o.rtol == 1e-7 and o.step_init == 1.0 and o.quirks == _abi.KT_QUIRKS_REFERENCE
    assert (o.palette, o.disk_rings, o.disk_sectors, o.disk_sigma, o.disk_hue_offset) == (0, 7, 14, 0.5, 0.0)
    bg = kt.checker_background(1234, 64, 32, 32, 16)
    from tests.golden.make_golden import checker
    assert np.array_equal(bg, checker(1234, 64, 32, 32, 16))
    assert lib.kt_hip_create(0, None) == _abi.KT_E_ARG
    assert lib.kt_make_checker_background(1, 0, 0, 1, 1, None) == _abi.KT_E_ARG


def test_errors_are_codes_not_exceptions():
    lib = kt.load_library()
    assert lib.kt_hip_render(None, None, 0, None, None, None, None, None) == _abi.KT_E_ARG
    assert lib.kt_hip_last_error(None) == b"null context"


def test_product_has_no_cpu_fallback(monkeypatch):
    monkeypatch.setattr(kt, "_lib", None)
    monkeypatch.setattr(kt, "LIB_PATH", "/nonexistent/libkt_hip.so")
    with pytest.raises(kt.KtError):
        kt.load_library()


def _status(opts):
    try:
        kt.check_opts(opts)
    except kt.KtError as e:
        return e.status, str(e)
    return _abi.KT_OK, ""


def test_option_validation_without_gpu():
    o = kt.default_opts()
    assert _status(o)[0] == _abi.KT_OK
    # Jacobi bundles are binary64 only: a mixed request is refused, not run in binary64
    o.bundles, o.precision = 1, kt.KT_PRECISION_MIXED
    st, msg = _status(o)
    assert st == _abi.KT_E_UNSUPPORTED and "binary64" in msg
    o.precision = kt.KT_PRECISION_FP64
    assert _status(o)[0] == _abi.KT_OK
    o = kt.default_opts()
    o.rtol = 0.0
    assert _status(o)[0] == _abi.KT_E_ARG


@pytest.mark.parametrize("rings,sectors,sigma,ok", [
    (7, 14, 0.5, True), (0, 0, 0.0, True), (150, 150, 10.0, True), (200, 200, 10.0, False),
    (7, 14, 1e9, False), (7, 14, float("inf"), False), (7, 14, float("nan"), False),
    (100000, 14, 0.5, False), (300, 300, 0.5, False)])
def test_interstellar_palette_loop_bound(rings, sectors, sigma, ok):
    """The shade kernel's cell loop is (rings + 2s + 1)(sectors + 2s + 1), s = ceil(3 sigma) + 1
    (main.cpp:1717-1797, which has no bound): sizes beyond 65536 cells per pixel are KT_E_ARG."""
    o = kt.default_opts()
    o.palette = kt.KT_PALETTE_INTERSTELLAR
    o.disk_rings, o.disk_sectors, o.disk_sigma = rings, sectors, sigma
    assert (_status(o)[0] == _abi.KT_OK) == ok
    o.palette = kt.KT_PALETTE_BLACKBODY      # the cell parameters are unused by the blackbody palette
    assert _status(o)[0] == _abi.KT_OK


def test_ppm_header_size_overflow_is_rejected(tmp_path):
    """A P6 header whose width * height overflows a signed 64-bit product is refused, not read."""
    lib = kt.load_library()
    p = tmp_path / "big.ppm"
    p.write_bytes(b"P6\n9999999999 9999999999\n255\n" + b"\0" * 64)
    buf = np.zeros(64, np.uint8)
    w, h = ctypes.c_int32(0), ctypes.c_int32(0)
    assert lib.kt_read_image_rgb(str(p).encode(), buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                 ctypes.byref(w), ctypes.byref(h)) != _abi.KT_OK
    p.write_bytes(b"P6\n70000 2\n255\n" + b"\0" * 64)
    assert lib.kt_read_image_rgb(str(p).encode(), buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                 ctypes.byref(w), ctypes.byref(h)) != _abi.KT_OK
    p.write_bytes(b"P6\n4 2\n255\n" + bytes(range(24)))
    assert lib.kt_read_image_rgb(str(p).encode(), buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                 ctypes.byref(w), ctypes.byref(h)) == _abi.KT_OK
    assert (w.value, h.value) == (4, 2) and np.array_equal(buf[:24], np.arange(24))
