"""The product's per-ray tracer (kt_trace.hpp / kt_setup.hpp / kt_shade.hpp), compiled for
the host, against the oracle.  Same sources as the HIP kernel; catches logic errors on the
CPU before any GPU run.  Tolerances are those of the GPU parity tests."""
import numpy as np
import pytest

from kerrtracecpp2_amd import _abi
from oracle import kto
from tests import _hostcheck as hc
from tests._util import angle_diff, golden_manifest, load_golden, photon_ring_band


def _grid(W, H, step):
    ys, xs = np.mgrid[0:H:step, 0:W:step]
    return xs.ravel().astype(np.int32), ys.ravel().astype(np.int32)


def _compare(ref, got, band, dg_tol, dr_tol, fate_min=0.999):
    keep = ~band
    fate_ok = (ref["outcome"] == got["outcome"])[keep].mean()
    assert fate_ok >= fate_min, fate_ok
    d = keep & (ref["outcome"] == 1) & (got["outcome"] == 1)
    if d.any():
        assert np.abs(ref["redshift"][d] - got["redshift"][d]).max() <= dg_tol
        assert np.abs(ref["r"][d] - got["r"][d]).max() <= dr_tol


@pytest.mark.parametrize("method", ["rk4", "dopri5"])
def test_c1_bl_full_frame(method):
    W, H = 320, 180
    fo = kto.make_frame(W, H, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12, chart="bl", integrator=method,
                        tol=1e-7)
    ref = load_golden("c1_bl_rk4_320x180")["pixels"] if method == "rk4" else kto.trace_frame(fo)
    px, py = _grid(W, H, 1)
    op = _abi.default_opts()
    op.integrator = 1 if method == "dopri5" else 0
    got, ex = hc.trace_pixels(_abi.make_frame(W, H, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12), op, px, py)
    band = photon_ring_band(ref, W, H)
    _compare(ref, got, band, dg_tol=1e-4, dr_tol=2e-3)
    assert ex["r_isco"] == kto.frame_info(fo).r_isco


def test_c2_ks_lambda_matches_bl_oracle():
    """KS chart with Lambda != 0 (no reference counterpart) against the reference's BL fallback."""
    g = load_golden("c2_ks_dopri5_lambda_160x90")
    W, H = 160, 90
    px, py = _grid(W, H, 1)
    op = _abi.default_opts()
    op.chart, op.integrator, op.rtol, op.atol = 1, 1, 1e-9, 1e-12
    fr = _abi.make_frame(W, H, a=0.9, Q=0.3, Lambda=1e-5, theta=85, r_obs=40, disk_out=12)
    got, ex = hc.trace_pixels(fr, op, px, py)
    ref = g["pixels"]
    band = photon_ring_band(ref, W, H)
    _compare(ref, got, band, dg_tol=1e-4, dr_tol=1e-3)
    assert ex["r_horizon"] == kto.r_horizon([1.0, 0.9, 0.3, 1e-5])
    assert ex["r_isco"] == g["meta"][4]
    esc = (~band) & (ref["outcome"] == 0) & (got["outcome"] == 0)
    # DOPRI5 keeps phi frozen (phi_obs = 0) and escapes fill theta_esc <- phi: exact zeros
    assert np.all(got["theta_esc"][esc] == 0.0) and np.all(ref["theta_esc"][esc] == 0.0)
    assert np.abs(got["phi_disk"][esc] - ref["phi_disk"][esc]).max() < 2e-3     # theta at escape


@pytest.mark.parametrize("name,a", [("ks_rk4_spin_pos_160x90", 0.5), ("ks_rk4_spin_neg_160x90", -0.5)])
def test_ks_lambda0_matches_reference_ks(name, a):
    ref = load_golden(name)["pixels"]
    W, H = 160, 90
    px, py = _grid(W, H, 1)
    op = _abi.default_opts()
    op.chart = 1
    got, _ = hc.trace_pixels(_abi.make_frame(W, H, a=a, theta=80, r_obs=40, fov=45, disk_out=12), op, px, py)
    band = photon_ring_band(ref, W, H)
    _compare(ref, got, band, dg_tol=3e-4, dr_tol=3e-3)
    esc = (~band) & (ref["outcome"] == 0) & (got["outcome"] == 0)
    assert angle_diff(got["theta_esc"][esc], ref["theta_esc"][esc]).max() < 5e-3


def test_setup_isco_exact_scan_matches_sequential():
    for mp in ((1.0, 0.9, 0.3, 1e-5), (1.0, 0.0, 0.4, 0.0), (1.0, -0.6, 0.1, 0.0), (1.0, 0.3, 0.0, -1e-4)):
        seq, seg = hc.isco_exact(*mp)
        assert seq == seg == kto.r_isco(list(mp)), mp


def test_shading_matches_oracle():
    g = load_golden("c1_bl_rk4_320x180")
    from tests.golden.make_golden import checker
    bg = checker()
    want = load_golden("c1_rgb_320x180")["rgb"].reshape(-1, 3)
    rgba = hc.shade(g["pixels"], 1.0, g["meta"][4], bg=bg)
    got = np.stack([rgba & 0xFF, (rgba >> 8) & 0xFF, (rgba >> 16) & 0xFF], -1)
    assert np.abs(got.astype(int) - want.astype(int)).max() <= 1
    assert (got != want).any(axis=1).mean() < 1e-3


@pytest.mark.parametrize("geo_name,img_name,cells", [
    ("c1_bl_rk4_320x180", "c1_interstellar_rgb_320x180", {}),
    ("ks_rk4_spin_pos_160x90", "ks_interstellar_rgb_160x90", dict(rings=5, sectors=9, sigma=0.8, hue=0.3))])
def test_interstellar_shading_matches_reference(geo_name, img_name, cells):
    """The shade kernel's Interstellar palette (its own sincos in the cell hash) built for the
    host, against the reference binary's PNGs: bit-exact."""
    from tests.golden.make_golden import checker
    g = load_golden(geo_name)
    m = g["meta"]
    rgba = hc.shade_interstellar(g["pixels"], m[0], m[4], m[5], m[6], bg=checker(), **cells)
    got = np.stack([rgba & 0xFF, (rgba >> 8) & 0xFF, (rgba >> 16) & 0xFF], -1)
    assert np.array_equal(got, load_golden(img_name)["rgb"].reshape(-1, 3))


def test_max_steps_one_positional_results():
    """--max-steps 1: every ray stops after one accepted step with the reference's
    positional TraceResult fill {ESCAPED, r, 1, theta, phi} (main.cpp:433)."""
    W, H = 16, 9
    px, py = _grid(W, H, 1)
    fo = kto.make_frame(W, H, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12, max_steps=1)
    ref = kto.trace_pixels(fo, px, py)
    op = _abi.default_opts()
    op.max_steps = 1
    got, _ = hc.trace_pixels(_abi.make_frame(W, H, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12), op, px, py)
    assert np.all(got["outcome"] == ref["outcome"])
    np.testing.assert_allclose(got["r"], ref["r"], rtol=1e-6)
    np.testing.assert_allclose(got["phi_disk"], ref["phi_disk"], atol=1e-6)   # theta
    np.testing.assert_allclose(got["theta_esc"], ref["theta_esc"], atol=1e-6)  # phi


BUNDLE_PTS = [(3, 4), (20, 4), (40, 18), (32, 18), (50, 10), (5, 30), (33, 12), (30, 24)]


def test_bundles_vs_reference_sample():
    """Jacobi bundles (dual-number J_f W) vs the reference's FD-Hessian bundles: same fate, r_hit,
    Doppler factor; magnification |det J| within 10 % -- the reference's own value at tol 1e-7
    moves by 9 % when the tolerance is tightened to 1e-8 (W is not error controlled)."""
    g = load_golden("bundles_bl_64x36")["pixels"].reshape(36, 64)
    px = np.array([p[0] for p in BUNDLE_PTS])
    py = np.array([p[1] for p in BUNDLE_PTS])
    op = _abi.default_opts()
    op.bundles = 1
    got, _ = hc.trace_bundle_pixels(_abi.make_frame(64, 36, a=0.99, theta=89, r_obs=40, disk_out=12), op, px, py)
    ref = g[py, px]
    assert np.array_equal(got["outcome"], ref["outcome"])
    d = ref["outcome"] == 1
    assert np.abs(got["r"][d] - ref["r"][d]).max() < 1e-3
    assert np.abs(got["redshift"][d] - ref["redshift"][d]).max() < 1e-4
    assert np.all(np.abs(got["magnif"][d] / ref["magnif"][d] - 1.0) < 0.10)
    e = ~d
    np.testing.assert_allclose(got["theta_esc"][e], ref["theta_esc"][e], atol=1e-4)
    np.testing.assert_allclose(got["phi_esc"][e], ref["phi_esc"][e], atol=1e-4)


def test_bundle_magnification_converged_vs_oracle():
    """At tol 1e-9 both magnifications are converged: 1 % agreement."""
    pts = [(20, 4), (40, 18), (32, 18), (50, 10)]
    px = np.array([p[0] for p in pts])
    py = np.array([p[1] for p in pts])
    ref = kto.trace_pixels(kto.make_frame(64, 36, a=0.99, theta=89, r_obs=40, disk_out=12, bundles=True, tol=1e-9),
                           px, py)
    op = _abi.default_opts()
    op.bundles, op.rtol, op.atol = 1, 1e-9, 1e-9
    got, _ = hc.trace_bundle_pixels(_abi.make_frame(64, 36, a=0.99, theta=89, r_obs=40, disk_out=12), op, px, py)
    assert np.array_equal(got["outcome"], ref["outcome"])
    assert np.all(np.abs(got["magnif"] / ref["magnif"] - 1.0) < 0.01)


def test_c3_bundle_sample_like_for_like():
    """Configuration 3 (1920x1080, a=0.99, theta=89, bundles, root-skip quirk cleared on both
    sides): the 16 384-pixel oracle sample that the full-size GPU test uses.  The host build of
    the kernels' code agrees on every fate, with no band excluded."""
    g = load_golden("c3_bundles_sample")
    op = _abi.default_opts()
    op.bundles = 1
    op.quirks = int(g["quirks"])
    got, _ = hc.trace_bundle_pixels(_abi.make_frame(int(g["W"]), int(g["H"]), a=0.99, theta=89, r_obs=40,
                                                    disk_out=12), op, g["px"], g["py"])
    ref = g["pixels"]
    assert np.array_equal(got["outcome"], ref["outcome"])
    d = ref["outcome"] == 1
    assert np.abs(got["r"][d] - ref["r"][d]).max() < 1e-3
    assert np.quantile(np.abs(got["redshift"][d] - ref["redshift"][d]), 0.999) < 2.5e-4
    rel = np.abs(got["magnif"][d] / ref["magnif"][d] - 1.0)
    assert np.quantile(rel, 0.5) < 0.005 and np.mean(rel < 0.05) >= 0.97
    e = (ref["outcome"] == 0) & (ref["theta_esc"] != 0)
    assert np.abs(got["theta_esc"][e] - ref["theta_esc"][e]).max() < 1e-3
    assert np.abs(got["phi_esc"][e] - ref["phi_esc"][e]).max() < 2e-3


def test_kernel_sincos_accuracy():
    """The kernels' inline sin/cos (Cody-Waite reduction + Taylor series) against x87
    extended precision: <= 2 ulp over the angles a ray meets, exact special values."""
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-4, 4, 200000), rng.uniform(-200, 200, 200000), rng.uniform(-1e6, 1e6, 20000),
                        np.arange(-64, 65) * np.pi / 4, [1e-300, 5e-324, np.pi, 3 * np.pi / 2]])
    s, c = hc.sincos(x)
    xl = x.astype(np.longdouble)
    for got, ref in ((s, np.sin(xl)), (c, np.cos(xl))):
        ulp = np.spacing(np.abs(ref.astype(np.float64))).astype(np.longdouble)
        assert float((np.abs(got.astype(np.longdouble) - ref) / ulp).max()) <= 2.0
    s, c = hc.sincos(np.array([0.0, -0.0, np.pi / 2, np.nan, np.inf, -np.inf]))
    assert np.array_equal(s[:2], [0.0, -0.0]) and np.signbit(s[1]) and np.array_equal(c[:2], [1.0, 1.0])
    assert s[2] == 1.0 and c[2] == 6.123233995736766e-17        # glibc's cos(M_PI/2)
    assert np.all(np.isnan(s[3:])) and np.all(np.isnan(c[3:]))


def test_event_pretests_never_skip_a_crossing():
    """The two pre-tests that let a step skip the 8-bin / 8-bisection disk search
    (geodesic.hpp:49-140) may only skip steps on which the search finds nothing that counts:
    hermite_may_touch is true whenever the search finds a crossing of the equator, and
    hermite_may_reach is true whenever that crossing's Hermite radius lies on the disk.  Random
    steps near the equator, plus steps built to graze the plane and the disk's edges."""
    rng = np.random.default_rng(11)
    n = 200000
    hp = np.pi / 2
    c = np.empty((n, 11))
    h = rng.uniform(0.01, 5.0, n)
    c[:, 0] = hp + rng.uniform(-0.4, 0.4, n)
    c[:, 1] = hp + rng.uniform(-0.4, 0.4, n)
    c[:, 2] = rng.uniform(-0.5, 0.5, n) / h
    c[:, 3] = rng.uniform(-0.5, 0.5, n) / h
    # a quarter graze the plane: endpoints on one side, slopes that bend the cubic back to it
    g = rng.random(n) < 0.25
    side = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    eps = 10.0 ** rng.uniform(-14, -3, n)
    c[g, 0] = hp + side[g] * eps[g]
    c[g, 1] = hp + side[g] * rng.uniform(0.0, 0.01, g.sum())
    c[g, 2] = -side[g] * rng.uniform(0.0, 0.05, g.sum()) / h[g]
    c[g, 3] = side[g] * rng.uniform(0.0, 0.05, g.sum()) / h[g]
    c[:, 4] = rng.uniform(1.5, 20.0, n)
    c[:, 5] = c[:, 4] + rng.uniform(-3.0, 3.0, n)
    c[:, 6] = rng.uniform(-5.0, 5.0, n) / h
    c[:, 7] = rng.uniform(-5.0, 5.0, n) / h
    # a quarter of the radii end near an edge of the disk [3.87, 12]
    e = rng.random(n) < 0.25
    edge = np.where(rng.random(n) < 0.5, 3.87, 12.0)
    c[e, 4] = edge[e] + rng.uniform(-0.05, 0.05, e.sum())
    c[e, 5] = edge[e] + rng.uniform(-0.05, 0.05, e.sum())
    c[:, 8] = h
    c[:, 9] = 3.87
    c[:, 10] = 12.0
    out = hc.event_filters(c)
    touch, found, reach, on_disk = (out[:, k].astype(bool) for k in range(4))
    assert found.sum() > 1000 and on_disk.sum() > 1000           # the cases exercise both tests
    assert not (found & ~touch).any(), np.where(found & ~touch)[0][:5]
    assert not (on_disk & ~reach).any(), np.where(on_disk & ~reach)[0][:5]
    # and they do skip: most steps without a crossing, and of those with one most off the disk
    assert (~touch & ~found).sum() > 0.5 * (~found).sum()
    assert (~reach & found & ~on_disk).sum() > 0.3 * (found & ~on_disk).sum()


def test_device_rhs_matches_reference_finite_differences():
    """The kernels' analytic Boyer-Lindquist right-hand side (kt_device.hpp rhs<CHART_BL>) against
    the reference's (geodesic.hpp:143-166: g^rr p_r, g^thth p_th, and -dH/dr, -dH/dth by central
    differences with eps_r = 1e-5 max(|r|, 0.01), eps_th = 1e-6), restated in the oracle.  The
    difference is the reference's truncation and rounding, ~1e-9 relative; a dropped term or a
    sign error is O(1)."""
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(400):
        a = rng.uniform(-0.95, 0.95)
        Q = rng.uniform(0.0, 0.3) * np.sqrt(max(0.0, 1.0 - a * a))
        Lam = rng.choice([0.0, rng.uniform(0.0, 1e-4)])
        mp = [1.0, a, Q, Lam]
        rh = kto.r_horizon(mp)
        r = rng.uniform(1.2 * rh, 40.0)
        th = rng.uniform(0.15, np.pi - 0.15)
        pr, pth = rng.uniform(-2.0, 2.0), rng.uniform(-6.0, 6.0)
        pt, pp = -1.0, rng.uniform(-6.0, 6.0)
        ref = np.asarray(kto.rhs_bl(mp, [r, th, pr, pth, pt, pp]))[:4]
        got = hc.rhs(0, 1.0, [a, Q, Lam], [r, th, pr, pth], pt, pp)
        err = np.abs(got - ref) / (np.abs(ref) + 1e-2)
        worst = max(worst, float(err.max()))
        assert np.all(np.abs(got - ref) <= 2e-7 * np.abs(ref) + 2e-9), (mp, r, th, pr, pth, pp, got, ref)
    print("worst relative difference", worst)


def test_constraint_projection_restores_null_condition():
    """The mixed path's binary64 projection onto H = 0 (kt_trace.hpp renormalise, BL), judged by
    the reference's own Hamiltonian and right-hand side (knds_metric.hpp:122-132,
    geodesic.hpp:143-166, restated in the oracle): from camera states of random spacetimes pushed
    off the shell by 1e-8..1e-5 in (p_r, p_theta), one projection cuts |H| by at least 1e4; the
    correction is parallel to grad_p H = (g^rr p_r, g^thth p_theta); a state on the shell is
    left as it is."""
    rng = np.random.default_rng(9)
    n_checked = 0
    for _ in range(60):
        a = rng.uniform(-0.9, 0.9)
        Q = rng.uniform(0.0, 0.3) * np.sqrt(1 - a * a)
        Lam = float(rng.choice([0.0, rng.uniform(0.0, 1e-5)]))
        mp = [1.0, a, Q, Lam]
        cam = [rng.uniform(20, 60), rng.uniform(30, 85), 0.0, 30.0, 64, 36]
        s = kto.camera_pixel(mp, cam, float(rng.integers(0, 64)), float(rng.integers(0, 36)))
        r, th, pr, pth, pt, pp = s[0], s[1], s[3], s[4], s[5], s[6]
        H0 = kto.hamiltonian_bl(mp, [r, th, pr, pth, pt, pp])
        # on the shell: unchanged
        same = hc.renormalise(0, 1.0, [a, Q, Lam], [[r, th, pr, pth, pt, pp]])[0]
        assert abs(H0) < 1e-10 and same[0] == pr and same[1] == pth
        eps = 10.0 ** rng.uniform(-8, -5)
        pr1, pth1 = pr * (1 + eps * rng.uniform(-1, 1)), pth * (1 + eps * rng.uniform(-1, 1)) + eps
        H1 = kto.hamiltonian_bl(mp, [r, th, pr1, pth1, pt, pp])
        if abs(H1) < 1e-11:
            continue
        out = hc.renormalise(0, 1.0, [a, Q, Lam], [[r, th, pr1, pth1, pt, pp]])[0]
        H2 = kto.hamiltonian_bl(mp, [r, th, out[0], out[1], pt, pp])
        assert abs(H2) <= 1e-4 * abs(H1) + 1e-15, (mp, cam, H1, H2)
        g = np.asarray(kto.rhs_bl(mp, [r, th, pr1, pth1, pt, pp]))[:2]      # g^rr p_r, g^thth p_theta
        dp = np.array([out[0] - pr1, out[1] - pth1])
        cosang = abs(dp @ g) / (np.linalg.norm(dp) * np.linalg.norm(g))
        assert cosang > 1 - 1e-6, (mp, cam, cosang)
        n_checked += 1
    assert n_checked >= 50


@pytest.mark.parametrize("chart", [0, 1])
def test_parked_lane_state_is_a_fixed_point(chart):
    """Idle lanes of the persistent trace kernel run every attempt on a parked state (r at the
    escape radius, theta 1 rad, every momentum zero): the right-hand side must vanish there
    exactly, so the lane never moves (no infall through the horizon cut, no disk search)."""
    for par in ((0.9, 0.3, 1e-5), (0.99, 0.0, 0.0), (-0.5, 0.2, 0.0)):
        f = hc.rhs(chart, 1.0, par, [42.0, 1.0, 0.0, 0.0], 0.0, 0.0)
        assert np.array_equal(f, np.zeros(4)), (chart, par, f)
