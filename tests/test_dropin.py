"""The drop-in as the reference's host code would use it (tests/dropin/dropin_check.cpp):
the USE_HIP branch of INTEGRATION.md compiled verbatim inside a render_image() of the
reference's signature (main.cpp:2209-2231), kthip::hip_render() in the cuda_render() call shape
(gpu/cuda/tracer.cuh:23-26), and kthip::hip_render_ex(), on configuration 1 (320x180 BL RK4)."""
import os
import subprocess

import numpy as np
import pytest

from kerrtracecpp2_amd import _abi
from oracle import kto
from tests._util import load_golden, photon_ring_band
from tests.dropin import build as dropin

W, H = 320, 180


def test_integration_snippet_compiles(tmp_path):
    """CPU: the documented integration compiles against include/kt_hip.h and hip_backend.hpp."""
    exe = dropin.build(str(tmp_path))
    assert os.path.exists(exe)
    assert "kthip::hip_render" not in dropin.snippet() and "kt_hip_render(ctx.get()" in dropin.snippet()


@pytest.mark.gpu
def test_dropin_paths_vs_reference_c1(tmp_path):
    exe = os.path.join(os.path.dirname(dropin.__file__), "_build", "dropin_check")
    assert os.path.exists(exe), "build first: __graft_entry__.build()"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    ref_geo = load_golden("c1_bl_rk4_320x180")["pixels"]
    keep = ~photon_ring_band(ref_geo, W, H)
    # 1. the INTEGRATION.md branch: GeoPixels against the reference binary's .kgeo, RGB against its PNG
    geo = np.fromfile(tmp_path / "snippet.geo", dtype=_abi.GEOPIXEL_DTYPE)
    assert geo.size == W * H
    assert (geo["outcome"] == ref_geo["outcome"])[keep].mean() >= 0.999
    d = keep & (geo["outcome"] == 1) & (ref_geo["outcome"] == 1)
    assert np.abs(geo["redshift"][d] - ref_geo["redshift"][d]).max() <= 5e-4
    assert np.abs(geo["r"][d] - ref_geo["r"][d]).max() <= 2e-3
    rgb = np.fromfile(tmp_path / "snippet.rgb", dtype=np.uint8).reshape(-1, 3).astype(int)
    want = load_golden("c1_rgb_320x180")["rgb"].reshape(-1, 3).astype(int)
    assert np.mean((np.abs(rgb - want).max(axis=1) <= 2)[keep]) >= 0.999
    # 2. cuda_render's shape: no background (escapes black), packed ABGR with alpha 0xFF
    abgr = np.fromfile(tmp_path / "cuda_shape.abgr", dtype=np.uint32)
    assert np.all(abgr >> 24 == 0xFF)
    got = np.stack([abgr & 0xFF, (abgr >> 8) & 0xFF, (abgr >> 16) & 0xFF], axis=1).astype(int)
    want_black = kto.colorize(ref_geo, 1.0, float(load_golden("c1_bl_rk4_320x180")["meta"][4])).astype(int)
    assert np.mean((np.abs(got - want_black).max(axis=1) <= 2)[keep]) >= 0.999
    assert np.all(got[ref_geo["outcome"] == 0][keep[ref_geo["outcome"] == 0]] == 0)
    # 3. hip_render_ex with the same controls and background is the snippet's image
    ex = np.fromfile(tmp_path / "render_ex.abgr", dtype=np.uint32)
    ex_rgb = np.stack([ex & 0xFF, (ex >> 8) & 0xFF, (ex >> 16) & 0xFF], axis=1).astype(int)
    assert np.array_equal(ex_rgb, rgb)
