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
