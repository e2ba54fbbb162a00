"""Host-side formats and the camera-spp rule, through the C ABI (no GPU needed).

  spp offsets ......... equal to the oracle's build_offsets restatement (main.cpp:2240-2279)
  spp resolve ......... host resolve equal to the oracle's (main.cpp:2281-2320) on random passes
  spp thresholds ...... the device resolve's per-level thresholds reproduce lround(255 x^(1/g))
  .kgeo ............... byte layout of save_kgeo (main.cpp:106-117), read back by the oracle's reader
  PNG / PPM ........... writer read by PIL; reader against PIL for every 8-bit colour type
  CLI ................. unsupported switches fail before any device work
"""
import os
import struct
import subprocess

import numpy as np
import pytest
from PIL import Image

import kerrtracecpp2_amd as kt
from oracle import kto
from tests._util import load_golden

CLI = os.path.join(os.path.dirname(kt.__file__), "bin", "kerr_tracer_hip")


def _abgr(rgb):
    rgb = rgb.astype(np.uint32)
    return rgb[..., 0] | (rgb[..., 1] << 8) | (rgb[..., 2] << 16) | np.uint32(0xFF000000)


@pytest.mark.parametrize("spp", [1, 2, 3, 4, 5, 8, 14, 15, 16, 17, 33])
def test_spp_offsets_match_oracle(spp):
    np.testing.assert_array_equal(kt.spp_offsets(spp), kto.spp_offsets(spp))


def test_spp_offsets_known_values():
    np.testing.assert_array_equal(kt.spp_offsets(2), [[-0.35, 0.0], [0.35, 0.0]])
    np.testing.assert_array_equal(kt.spp_offsets(3), [[0.0, 0.0], [0.35, 0.0], [-0.35, -0.0]])
    o = kt.spp_offsets(32)
    assert np.all(np.abs(o) < 0.5)                       # inside the pixel
    np.testing.assert_allclose(o[0::2] + o[1::2], 0.0, atol=0)   # mirrored pairs (even spp)


@pytest.mark.parametrize("gamma", [2.2, 1.0, 0.45, 3.7])
@pytest.mark.parametrize("n_pass", [2, 3, 7, 16])
def test_spp_resolve_matches_oracle(gamma, n_pass):
    rng = np.random.default_rng(n_pass * 100 + int(gamma * 10))
    passes = rng.integers(0, 256, size=(n_pass, 30000, 3), dtype=np.uint8)
    want = kto.spp_resolve(passes, gamma)
    np.testing.assert_array_equal(kt.spp_resolve(_abgr(passes), gamma), want)


@pytest.mark.parametrize("gamma", [2.2, 1.0, 0.45, 3.7, 1e-9])
def test_spp_thresholds_reproduce_display_rounding(gamma):
    """The device resolve picks the level by searching these thresholds; on random and on
    boundary-adjacent linear values it must agree with the oracle's pow + lround."""
    lin, thr = kt.spp_tables(gamma)
    g = max(gamma, 1e-6)
    np.testing.assert_array_equal(lin, [(v / 255.0) ** g for v in range(256)])   # libm pow, as the oracle
    assert thr[0] == -np.inf and np.all(np.diff(thr) > 0)
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.random(200000), thr[1:], np.nextafter(thr[1:], -1.0), [0.0, 1.0]])
    xs = xs[(xs >= 0) & (xs <= 1)]
    level = np.searchsorted(thr, xs, side="right") - 1
    # the oracle's rule on a single pass of x^(1/g) levels: resolve of one pass is identity, so use
    # the closed form directly with the same library pow the oracle uses
    want = np.array([int(np.floor(255.0 * v ** (1.0 / g) + 0.5)) for v in xs[:2000]])
    np.testing.assert_array_equal(level[:2000], want)
    # and against the oracle's resolve for pass averages that land on these x values
    two = rng.integers(0, 256, size=(2, 50000, 3), dtype=np.uint8)
    acc = (lin[two[0]] + lin[two[1]]) * 0.5
    lev = np.searchsorted(thr, np.clip(acc, 0, 1), side="right") - 1
    np.testing.assert_array_equal(lev.astype(np.uint8), kto.spp_resolve(two, gamma))


def test_spp_resolve_reproduces_reference_png_from_oracle_passes():
    from tests.golden.make_golden import checker
    bg = checker()
    passes = []
    for ox, oy in kt.spp_offsets(3):
        fr = kto.make_frame(80, 45, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12, offset_x=ox, offset_y=oy)
        passes.append(kto.colorize(kto.trace_frame(fr), 1.0, kto.frame_info(fr).r_isco, bg=bg))
    got = kt.spp_resolve(_abgr(np.stack(passes)))
    np.testing.assert_array_equal(got, load_golden("spp3_rgb_80x45")["rgb"].reshape(-1, 3))


def _meta(W, H):
    return dict(W=W, H=H, M=1.0, a=0.9, Q=0.3, Lambda=1e-5, r_isco=2.75, r_disk_in=2.75, r_disk_out=12.0,
                theta_deg=85.0, phi_deg=10.0, r_obs=40.0)


def test_kgeo_layout_and_round_trip(tmp_path):
    W, H = 17, 5
    g = load_golden("c1_bl_rk4_320x180")["pixels"][: W * H]
    p = str(tmp_path / "x.kgeo")
    kt.save_kgeo(p, g, _meta(W, H))
    raw = open(p, "rb").read()
    assert raw[:4] == b"KGEO" and struct.unpack("<I", raw[4:8])[0] == 1
    assert len(raw) == 8 + 88 + 28 * W * H
    head = struct.unpack("<II10d", raw[8:96])
    assert head[:2] == (W, H) and head[2:] == (1.0, 0.9, 0.3, 1e-5, 2.75, 2.75, 12.0, 85.0, 10.0, 40.0)
    assert raw[96:] == g.tobytes()
    g2, m2 = kt.load_kgeo(p)
    assert g2.tobytes() == g.tobytes() and m2["W"] == W and m2["theta_deg"] == 85.0
    # the oracle's reader (the one that reads the reference's own .kgeo output) agrees
    k = kto.read_kgeo(p)
    assert k["version"] == 1 and (k["W"], k["H"]) == (W, H)
    assert k["pixels"].tobytes() == g.tobytes() and tuple(k["meta"]) == head[2:]


def test_kgeo_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.kgeo"
    p.write_bytes(b"KGEX" + b"\0" * 200)
    with pytest.raises(kt.KtError):
        kt.load_kgeo(str(p))
    p.write_bytes(b"KGEO" + struct.pack("<I", 2) + b"\0" * 200)
    with pytest.raises(kt.KtError):
        kt.load_kgeo(str(p))
    good = tmp_path / "good.kgeo"
    kt.save_kgeo(str(good), load_golden("c1_bl_rk4_320x180")["pixels"][:12], _meta(4, 3))
    p.write_bytes(good.read_bytes()[:-5])     # truncated pixel data
    with pytest.raises(kt.KtError):
        kt.load_kgeo(str(p))


def test_png_writer_read_by_pil(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, size=(37, 53, 3), dtype=np.uint8)
    p = str(tmp_path / "a.png")
    kt.write_png(p, img)
    with Image.open(p) as im:
        assert im.mode == "RGB" and im.size == (53, 37)
        np.testing.assert_array_equal(np.asarray(im), img)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P"])
def test_png_reader_against_pil(tmp_path, mode):
    rng = np.random.default_rng(5)
    base = Image.fromarray(rng.integers(0, 256, size=(29, 41, 3), dtype=np.uint8), "RGB")
    if mode == "RGB":
        # a smooth image so PIL's encoder picks Sub/Up/Average/Paeth filters
        y, x = np.mgrid[0:29, 0:41]
        base = Image.fromarray(np.stack([x * 6, y * 8, (x + y) * 3], -1).astype(np.uint8), "RGB")
    im = base.convert(mode) if mode != "P" else base.quantize(64)
    p = str(tmp_path / f"b_{mode}.png")
    im.save(p, optimize=False)
    want = np.asarray(Image.open(p).convert("RGB"))
    np.testing.assert_array_equal(kt.read_image_rgb(p), want)


def test_ppm_reader(tmp_path):
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, size=(11, 7, 3), dtype=np.uint8)
    p = str(tmp_path / "c.ppm")
    Image.fromarray(img).save(p)
    np.testing.assert_array_equal(kt.read_image_rgb(p), img)


def test_checker_background_png_round_trip(tmp_path):
    bg = kt.checker_background()
    p = str(tmp_path / "sky.png")
    kt.write_png(p, bg)
    np.testing.assert_array_equal(kt.read_image_rgb(p), bg)


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("flag", ["--semi-analytic", "--wormhole", "--elliptic-closed", "--intersection-linear"])
def test_cli_rejects_unprovided_modes(flag):
    r = subprocess.run([CLI, "--custom-res", "8", "8", "--bg", "", flag], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "not provided by the HIP backend" in r.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("args,msg", [
    (["--disk-interstellar", "--disk-rings", "100000"], "invalid options: disk_rings/disk_sectors above"),
    (["--bundles", "--precision", "mixed"], "--bundles runs in binary64 only"),
])
def test_cli_rejects_invalid_options_before_opening_a_device(args, msg):
    """The CLI applies kt_check_opts (the check of every render entry point) before it creates a
    context, so these fail the same way with or without a GPU."""
    r = subprocess.run([CLI, "--custom-res", "8", "8", "--bg", ""] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert msg in r.stderr
    assert "Device:" not in r.stdout
