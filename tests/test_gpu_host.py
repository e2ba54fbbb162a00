"""GPU: the callers either side of the hot path -- camera-spp folded into one launch, .kgeo
emission / colour-only re-shading, and the kerr_tracer_hip CLI (single frame, spp, animation).

Device results are compared bit for bit with the same kernels reached another way (the
per-pass render + host resolve, the render's own GeoPixels), and the spp image with the
reference's camera-spp=3 PNG under the same per-pixel bar as the single-pass image.
"""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import kerrtracecpp2_amd as kt
from oracle import kto
from tests._util import load_golden, photon_ring_band

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(kt.__file__), "bin", "kerr_tracer_hip")


def _opts(**kw):
    o = kt.default_opts()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _sky():
    from tests.golden.make_golden import checker
    return checker()


def _passes(ctx, frame, spp, opts):
    frames = []
    for ox, oy in kt.spp_offsets(spp):
        f = kt.make_frame(frame.width, frame.height, a=frame.a, Q=frame.Q, Lambda=frame.Lambda, theta=frame.theta_deg,
                          phi=frame.phi_deg, r_obs=frame.r_obs, fov=frame.fov_deg, disk_out=frame.disk_out,
                          offset_x=frame.offset_x + ox, offset_y=frame.offset_y + oy)
        frames.append(f)
    out = ctx.render(frames, opts, want_rgba=True, want_geo=False)
    return out["rgba"].reshape(spp, -1)


@pytest.mark.parametrize("spp", [2, 3, 4, 9])
def test_render_spp_equals_resolved_passes(gpu_ctx, spp):
    gpu_ctx.set_background(_sky())
    try:
        fr = kt.make_frame(80, 45, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12)
        o = _opts()
        got = gpu_ctx.render_spp([fr], spp, o)["rgb"]
        want = kt.spp_resolve(_passes(gpu_ctx, fr, spp, o), o.gamma)
    finally:
        gpu_ctx.set_background(None)
    np.testing.assert_array_equal(got, want)


def test_render_spp_multi_frame_layout(gpu_ctx):
    """Frames of different sizes in one spp call land at their own offsets."""
    gpu_ctx.set_background(_sky())
    try:
        o = _opts(chart=kt.KT_CHART_KS, integrator=kt.KT_INTEGRATOR_DOPRI5, rtol=1e-8, atol=1e-8, gamma=1.8)
        f1 = kt.make_frame(40, 24, a=0.9, theta=70, disk_out=10)
        f2 = kt.make_frame(24, 40, a=-0.6, Q=0.2, theta=85, phi=30, disk_out=14)
        both = gpu_ctx.render_spp([f1, f2], 5, o)["rgb"]
        one = gpu_ctx.render_spp([f1], 5, o)["rgb"]
        two = gpu_ctx.render_spp([f2], 5, o)["rgb"]
    finally:
        gpu_ctx.set_background(None)
    np.testing.assert_array_equal(both, np.concatenate([one, two]))


def test_render_spp3_vs_reference_png(gpu_ctx):
    W, H = 80, 45
    gpu_ctx.set_background(_sky())
    try:
        got = gpu_ctx.render_spp([kt.make_frame(W, H, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12)], 3, _opts())["rgb"]
    finally:
        gpu_ctx.set_background(None)
    want = load_golden("spp3_rgb_80x45")["rgb"].reshape(-1, 3).astype(int)
    band = np.zeros(W * H, bool)
    for ox, oy in kt.spp_offsets(3):      # a pixel is in the band if any of its passes is
        geo = kto.trace_frame(kto.make_frame(W, H, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12,
                                             offset_x=ox, offset_y=oy))
        band |= photon_ring_band(geo, W, H)
    diff = np.abs(got.astype(int) - want).max(axis=1)
    assert np.mean(diff[~band] <= 2) >= 0.999, np.mean(diff[~band] <= 2)


def _run(args, tmp_path):
    env = dict(os.environ, KT_OUT_DIR=str(tmp_path / "out"))
    print("CLI:", " ".join(args), flush=True)
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.fixture(scope="module")
def sky_png(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("sky") / "sky.png")
    kt.write_png(p, _sky())
    return p


def test_cli_geo_only_and_color_only(gpu_ctx, tmp_path, sky_png):
    W, H = 64, 36
    kg = str(tmp_path / "f.kgeo")
    args = ["--bl", "--custom-res", str(W), str(H), "--a", "0.7", "--charge", "0.2", "--theta", "75", "--phi", "15",
            "--fov", "40", "--disk-out", "14", "--bg", sky_png]
    out = _run(args + ["--geo-only", "--geo-file", kg], tmp_path)
    assert "Trace:" in out and "Geo saved:" in out
    geo, meta = kt.load_kgeo(kg)
    fr = kt.make_frame(W, H, a=0.7, Q=0.2, theta=75, phi=15, fov=40, disk_out=14)
    ref = gpu_ctx.render([fr], _opts(), want_rgba=False, want_geo=True)
    assert geo.tobytes() == ref["geo"].tobytes()
    info = ref["info"][0]
    assert (meta["W"], meta["H"]) == (W, H)
    assert meta["r_isco"] == info["r_isco"] and meta["r_disk_in"] == info["r_isco"]
    assert (meta["a"], meta["Q"], meta["theta_deg"], meta["phi_deg"], meta["r_obs"], meta["r_disk_out"]) == \
        (0.7, 0.2, 75.0, 15.0, 40.0, 14.0)
    # colour-only: Phase 2 from the file, with the same background
    out = _run(["--color-only", kg, "--bg", sky_png, "--exposure", "1.3"], tmp_path)
    png = re.search(r"Saved: (\S+_recolor_\d{8}-\d{6}\.png)", out).group(1)
    gpu_ctx.set_background(_sky())
    try:
        want = kt.rgba_to_rgb(gpu_ctx.colorize(geo, 1.0, meta["r_isco"], _opts(exposure=1.3)))
    finally:
        gpu_ctx.set_background(None)
    np.testing.assert_array_equal(kt.read_image_rgb(png).reshape(-1, 3), want)


def test_cli_interstellar_palette(gpu_ctx, tmp_path, sky_png):
    W, H = 48, 32
    cells = ["--disk-interstellar", "--disk-rings", "5", "--disk-sectors", "9", "--disk-sigma", "0.8",
             "--disk-hue-offset", "0.3"]
    out = _run(["--ks", "--custom-res", str(W), str(H), "--a", "0.5", "--theta", "80", "--fov", "45",
                "--bg", sky_png] + cells, tmp_path)
    png = re.search(r"Saved: (\S+\.png)", out).group(1)
    gpu_ctx.set_background(_sky())
    try:
        o = _opts(chart=kt.KT_CHART_KS, palette=kt.KT_PALETTE_INTERSTELLAR, disk_rings=5, disk_sectors=9,
                  disk_sigma=0.8, disk_hue_offset=0.3)
        ref = gpu_ctx.render([kt.make_frame(W, H, a=0.5, theta=80, fov=45, disk_out=12)], o)
    finally:
        gpu_ctx.set_background(None)
    np.testing.assert_array_equal(kt.read_image_rgb(png).reshape(-1, 3), kt.rgba_to_rgb(ref["rgba"]))
    disk = ref["geo"]["outcome"] == 1
    assert disk.sum() > 50
    # a kgeo re-coloured with the same cell options gives the same pixels
    kg = str(tmp_path / "i.kgeo")
    _run(["--ks", "--custom-res", str(W), str(H), "--a", "0.5", "--theta", "80", "--fov", "45", "--geo-only",
          "--geo-file", kg], tmp_path)
    out = _run(["--color-only", kg, "--bg", sky_png] + cells, tmp_path)
    png2 = re.search(r"Saved: (\S+_recolor_\d{8}-\d{6}\.png)", out).group(1)
    np.testing.assert_array_equal(kt.read_image_rgb(png2), kt.read_image_rgb(png))


def test_cli_full_render_naming_and_pixels(gpu_ctx, tmp_path, sky_png):
    W, H = 48, 32
    kg = str(tmp_path / "g.kgeo")
    out = _run(["--ks", "--dopri5", "--tol", "1e-8", "--custom-res", str(W), str(H), "--a", "0.9", "--charge", "0.3",
                "--lambda", "1e-5", "--theta", "85", "--bg", sky_png, "--geo-file", kg, "--gpu-fp64"], tmp_path)
    png = re.search(r"Saved: (\S+\.png)", out).group(1)
    # this backend integrates KS with Lambda != 0 (the reference falls back to BL and names it so)
    assert re.fullmatch(r"custom_48x32_ks-standard-dopri5_gpu-hip_gpufp64req_\d{8}-\d{6}\.png", os.path.basename(png))
    gpu_ctx.set_background(_sky())
    try:
        o = _opts(chart=kt.KT_CHART_KS, integrator=kt.KT_INTEGRATOR_DOPRI5, rtol=1e-8, atol=1e-8)
        ref = gpu_ctx.render([kt.make_frame(W, H, a=0.9, Q=0.3, Lambda=1e-5, theta=85, disk_out=12)], o)
    finally:
        gpu_ctx.set_background(None)
    np.testing.assert_array_equal(kt.read_image_rgb(png).reshape(-1, 3), kt.rgba_to_rgb(ref["rgba"]))
    geo, _ = kt.load_kgeo(kg)
    assert geo.tobytes() == ref["geo"].tobytes()


def test_cli_camera_spp(gpu_ctx, tmp_path, sky_png):
    W, H = 40, 30
    out = _run(["--bl", "--custom-res", str(W), str(H), "--a", "0.5", "--fov", "45", "--camera-spp", "4",
                "--bg", sky_png], tmp_path)
    png = re.search(r"Saved: (\S+\.png)", out).group(1)
    assert "_gpu-hip_spp4_" in os.path.basename(png)
    gpu_ctx.set_background(_sky())
    try:
        want = gpu_ctx.render_spp([kt.make_frame(W, H, a=0.5, theta=80, fov=45, disk_out=12)], 4, _opts())["rgb"]
    finally:
        gpu_ctx.set_background(None)
    np.testing.assert_array_equal(kt.read_image_rgb(png).reshape(-1, 3), want)


def test_cli_animation_through_plan_vs_oracle(gpu_ctx, tmp_path, sky_png):
    """--anim (main.cpp:2951-3062) renders its batches through one kt_hip_plan: the first batch is
    captured in a HIP graph and replayed for the next, shorter one.  Every frame is checked against
    the oracle's render of that frame's interpolated parameters (fate band excluded, RGB within 2
    levels), not against the kernel itself."""
    W, H, n = 96, 54, 3
    fdir = str(tmp_path / "frames")
    out = _run(["--bl", "--custom-res", str(W), str(H), "--anim", "--frames", str(n), "--anim-batch", "2",
                "--no-encode", "--phi-start", "0", "--phi-end", "20", "--a-start", "0.3", "--a-end", "0.7",
                "--theta-start", "70", "--theta-end", "84", "--frames-dir", fdir, "--bg", sky_png], tmp_path)
    assert "Plan: 2 frames per graph launch" in out
    files = sorted(glob.glob(os.path.join(fdir, "frame_*.png")))
    assert [os.path.basename(f) for f in files] == ["frame_%05d.png" % k for k in range(n)]
    sky = _sky()
    for k, f in enumerate(files):
        t = k / (n - 1)                          # the CLI's interpolation, evaluated the same way
        a, th, ph = 0.3 + (0.7 - 0.3) * t, 70.0 + (84.0 - 70.0) * t, 0.0 + (20.0 - 0.0) * t
        fr = kto.make_frame(W, H, a=a, theta=th, phi=ph, disk_out=12)
        geo = kto.trace_frame(fr)
        want = kto.colorize(geo, 1.0, kto.frame_info(fr).r_isco, bg=sky).astype(int)
        got = kt.read_image_rgb(f).reshape(-1, 3).astype(int)
        keep = ~photon_ring_band(geo, W, H)
        ok = (np.abs(got - want).max(axis=1) <= 2)[keep].mean()
        print("anim frame", k, "a=%.2f theta=%.1f phi=%.1f: RGB within 2 levels on %.4f off-band" % (a, th, ph, ok),
              flush=True)
        assert ok >= 0.999
