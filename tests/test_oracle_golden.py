"""Pin the CPU oracle against the reference built from source.

Parity status of the oracle (oracle/kt_oracle.cpp), function by function:
  metric g_mn / g^mn (BL, KS), Hamiltonian, FD RHS, Keplerian Omega, horizon (both
  branches), ISCO (both branches), camera ray, RK4 doubling and DOPRI5 steps, BL<->KS
  maps, Hermite event root ......... bit-exact vs function_vectors.json (reference headers)
  trace_bl / trace_ks / trace_bundle / GeoPixel fill ........ bit-exact vs the reference
                                      binary's .kgeo for 7 configurations
  colorize (blackbody, Page-Thorne, tonemap, sky sample) .... vs the reference PNGs
  camera-spp offsets + linear-space resolve ................ vs the reference spp=3 PNG
"""
import numpy as np
import pytest

from tests._util import golden_manifest, load_golden
from oracle import kto


def _vectors():
    import json
    import os
    from tests._util import GOLDEN
    with open(os.path.join(GOLDEN, "function_vectors.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", sorted(golden_manifest()["frames"].keys()))
def test_oracle_frame_bit_exact(name):
    man = golden_manifest()["frames"][name]
    if "bundles" in name:
        pytest.skip("bundle frame checked in test_oracle_bundle_sample (reference bundle path is slow)")
    g = load_golden(name)
    fr = kto.make_frame(**man["frame"])
    out = kto.trace_frame(fr)
    assert out.tobytes() == g["pixels"].tobytes(), name
    fi = kto.frame_info(fr)
    meta = g["meta"]   # M, a, Q, Lam, r_isco, r_disk_in, r_disk_out, theta, phi, r_obs
    assert fi.r_isco == meta[4]


def test_oracle_bundle_sample():
    """Bundle mode on a pixel sample (the full 64x36 frame takes minutes on the CPU: some rays
    run to max_steps = 500000 near the horizon).  Sample mixes disk hits and escapes."""
    man = golden_manifest()["frames"]["bundles_bl_64x36"]
    g = load_golden("bundles_bl_64x36")
    W, H = int(g["W"]), int(g["H"])
    pts = [(3, 4), (20, 4), (40, 18), (32, 18), (50, 10), (5, 30), (33, 12), (30, 24)]
    px = np.array([p[0] for p in pts])
    py = np.array([p[1] for p in pts])
    out = kto.trace_pixels(kto.make_frame(**man["frame"]), px, py)
    ref = g["pixels"].reshape(H, W)[py, px]
    assert out.tobytes() == ref.tobytes()


def test_oracle_function_vectors():
    for c in _vectors():
        k, i, want = c["kind"], c["inputs"], np.array(c["output"])
        if k in ("upper", "lower"):
            got = (kto.upper_bl if k == "upper" else kto.lower_bl)(i["mp"], i["r"], i["th"]).ravel()
        elif k == "upper_ks":
            got = kto.upper_ks(i["mp"], i["X"], i["Y"], i["Z"]).ravel()
        elif k == "ham":
            got = np.array([kto.hamiltonian_bl(i["mp"], i["s6"])])
        elif k == "rhs":
            got = kto.rhs_bl(i["mp"], i["s6"])
        elif k == "horizon":
            got = np.array([kto.r_horizon(i["mp"])])
        elif k == "isco":
            got = np.array([kto.r_isco(i["mp"])])
        elif k == "omega":
            got = np.array([kto.omega_k(i["mp"], i["r"])])
        elif k == "camera":
            st = kto.camera_angles(i["mp"], i["cam"], i["al"], i["be"])
            got = st
        elif k == "step":
            ok, st, h, _ = kto.adaptive_step(i["mp"], "dopri5" if i["method"] else "rk4", i["state"], i["h"],
                                             None, i["tol"])
            got = np.array([1.0 if ok else 0.0, st[0], st[1], st[2], st[3], st[4], h])
        elif k == "bl2ks":
            got = kto.bl_to_ks(i["r"], i["th"], i["ph"], i["a"])
        elif k == "ks2bl":
            got = kto.ks_to_bl(i["X"], i["Y"], i["Z"], i["a"])
        elif k == "hermite_root":
            ok, al = kto.hermite_first_root(i["y0"], i["y1"], i["d0"], i["d1"], i["h"], i["target"])
            got = np.array([1.0 if ok else 0.0, al if ok else -1.0])
        else:
            raise AssertionError(k)
        assert got.shape == want.shape, k
        assert np.array_equal(got, want), (k, i, got, want)


def _sky():
    from tests.golden.make_golden import checker
    return checker()


def test_oracle_colorize_matches_reference_png():
    man = golden_manifest()
    bg = _sky()
    for geo_name, img_name in (("c1_bl_rk4_320x180", "c1_rgb_320x180"), ("c2_ks_dopri5_lambda_160x90", "c2_rgb_160x90")):
        g = load_golden(geo_name)
        want = load_golden(img_name)["rgb"].reshape(-1, 3)
        rgb = kto.colorize(g["pixels"], 1.0, g["meta"][4], bg=bg)
        assert np.array_equal(rgb, want), geo_name
    assert man["background"]["cells_x"] == 32


INTERSTELLAR = [("c1_bl_rk4_320x180", "c1_interstellar_rgb_320x180", {}),
                ("ks_rk4_spin_pos_160x90", "ks_interstellar_rgb_160x90",
                 dict(rings=5, sectors=9, sigma=0.8, hue_offset=0.3))]


@pytest.mark.parametrize("geo_name,img_name,cells", INTERSTELLAR)
def test_oracle_interstellar_matches_reference_png(geo_name, img_name, cells):
    """Interstellar palette (main.cpp:1689-1797) over the reference's own GeoPixels: bit-exact
    against the reference binary's PNG (default cells, and every cell option changed)."""
    g = load_golden(geo_name)
    M, r_isco, r_in, r_out = g["meta"][0], g["meta"][4], g["meta"][5], g["meta"][6]
    want = load_golden(img_name)["rgb"].reshape(-1, 3)
    rgb = kto.colorize_interstellar(g["pixels"], M, r_isco, r_in, r_out, bg=_sky(), **cells)
    assert np.array_equal(rgb, want)
    disk = g["pixels"]["outcome"] == 1
    assert len(np.unique(rgb[disk], axis=0)) > 500
    assert not np.array_equal(rgb[disk], kto.colorize(g["pixels"], M, r_isco, bg=_sky())[disk])


def test_oracle_spp_resolve_matches_reference_png():
    """camera-spp=3: three offset passes resolved in linear space (main.cpp:2232-2322)."""
    bg = _sky()
    want = load_golden("spp3_rgb_80x45")["rgb"].reshape(-1, 3)
    offs = kto.spp_offsets(3)
    np.testing.assert_array_equal(offs, [[0.0, 0.0], [0.35, 0.0], [-0.35, -0.0]])
    passes = []
    for ox, oy in offs:
        fr = kto.make_frame(80, 45, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12, offset_x=ox, offset_y=oy)
        geo = kto.trace_frame(fr)
        passes.append(kto.colorize(geo, 1.0, kto.frame_info(fr).r_isco, bg=bg))
    got = kto.spp_resolve(np.stack(passes))
    assert np.array_equal(got, want)


def test_bundle_colorize_vs_reference_png():
    """Phase 2 in bundle mode: the oracle's shading of the reference's bundle GeoPixels (flux
    weighted by clamp(1/magnif, 0.05, 5), main.cpp:1799-1814) against the reference's own PNG of the
    same frame, bit for bit."""
    from tests.golden.make_golden import checker
    g = load_golden("bundles_bl_64x36")
    want = load_golden("bundles_rgb_64x36")["rgb"].reshape(-1, 3)
    got = kto.colorize(g["pixels"], 1.0, float(g["meta"][4]), bg=checker())
    assert np.array_equal(got, want)
    assert (g["pixels"]["magnif"][g["pixels"]["outcome"] == 1] != 1.0).all()
