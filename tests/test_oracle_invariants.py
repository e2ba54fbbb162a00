"""Mathematical invariants of the oracle that the reference's own tests hold
(tests/core_tests.cpp:18-96) plus closed forms that a plausible transcription
mistake would break (dropped term, sign, swapped index)."""
import numpy as np

from oracle import kto


def test_metric_inverse_bl():                      # core_tests.cpp:18-34
    mp = [1.0, 0.5, 0.1, 1e-4]
    G = kto.lower_bl(mp, 10.0, 1.2)
    U = kto.upper_bl(mp, 10.0, 1.2)
    np.testing.assert_allclose(U @ G, np.eye(4), atol=1e-8)


def test_metric_inverse_random_points():
    rng = np.random.default_rng(3)
    for _ in range(50):
        mp = [1.0, rng.uniform(-0.99, 0.99), rng.uniform(0, 0.3), rng.uniform(-1e-4, 1e-4)]
        r, th = rng.uniform(3, 50), rng.uniform(0.2, 2.9)
        np.testing.assert_allclose(kto.upper_bl(mp, r, th) @ kto.lower_bl(mp, r, th), np.eye(4), atol=1e-9)


def test_ks_metric_is_inverse_and_null_vector():
    # g^ab = eta^ab - H l^a l^b with eta(l,l) = 0 (knds_metric.hpp:193-218)
    rng = np.random.default_rng(4)
    for _ in range(20):
        a, Q = rng.uniform(-0.99, 0.99), rng.uniform(0, 0.3)
        X, Y, Z = rng.uniform(-10, 10, 3)
        U = kto.upper_ks([1.0, a, Q, 0.0], X, Y, Z)
        r = kto.r_ks(X, Y, Z, a)
        l = np.array([1.0, (r * X + a * Y) / (r * r + a * a), (r * Y - a * X) / (r * r + a * a), Z / r])
        eta = np.diag([-1.0, 1, 1, 1])
        assert abs(l @ eta @ l) < 1e-12
        Hs = (2 * r - Q * Q) / (r * r + a * a * Z * Z / (r * r))
        G = eta + Hs * np.outer(l, l)
        np.testing.assert_allclose(U @ G, np.eye(4), atol=1e-10)


def test_camera_ray_null_future_directed():          # core_tests.cpp:36-45
    mp = [1.0, -0.5, 0.0, 0.0]
    s = kto.camera_angles(mp, [40.0, 80.0, 15.0, 45.0, 640, 360], 0.0, 0.0)
    H = kto.hamiltonian_bl(mp, [s[0], s[1], s[3], s[4], s[5], s[6]])
    assert abs(H) < 1e-10 and s[5] < 0.0 and abs(s[2] - np.radians(15.0)) < 1e-14


def test_one_step_preserves_null_residual():          # core_tests.cpp:47-75
    mp = [1.0, -0.5, 0.0, 0.0]
    for method in ("rk4", "dopri5"):
        s = kto.camera_angles(mp, [40.0, 80.0, 0.0, 45.0, 640, 360], 0.03, -0.02)
        h, fs, ok = 1.0, None, False
        for _ in range(32):
            ok, s, h, fs = kto.adaptive_step(mp, method, s, h, fs, 1e-7)
            if ok:
                break
        assert ok
        assert abs(kto.hamiltonian_bl(mp, [s[0], s[1], s[3], s[4], s[5], s[6]])) < 5e-6


def test_keplerian_omega_screen_convention():        # core_tests.cpp:77-84
    op = kto.omega_k([1.0, 0.5, 0.0, 0.0], 8.0)
    on = kto.omega_k([1.0, -0.5, 0.0, 0.0], 8.0)
    assert op < 0.0 < on and abs(abs(op) - abs(on)) < 1e-14
    # closed form |Omega| = 1/(r^1.5 - |a|) for the s = -sign(a) branch at M = 1
    assert abs(abs(op) - 1.0 / (8.0 ** 1.5 - 0.5)) < 1e-15


def test_ks_radius_recovery():                        # core_tests.cpp:86-96
    X, Y, Z = kto.bl_to_ks(8.0, 1.2, 2.1, 0.5)
    assert abs(kto.r_ks(X, Y, Z, 0.5) - 8.0) < 1e-9 and abs(Z - 8.0 * np.cos(1.2)) < 1e-12
    r, th, ph = kto.ks_to_bl(X, Y, Z, 0.5)
    assert abs(r - 8.0) < 1e-9 and abs(th - 1.2) < 1e-9 and abs(ph - 2.1) < 1e-9


def test_horizon_closed_forms():
    for a, Q in ((0.0, 0.0), (0.5, 0.0), (0.9, 0.3), (0.0, 0.4)):
        assert kto.r_horizon([1.0, a, Q, 0.0]) == 1.0 + np.sqrt(1.0 - a * a - Q * Q)
    # Lambda != 0: a root of Delta_r close to the Lambda = 0 horizon, below the cosmological one
    for lam in (1e-5, 1e-4):
        rh = kto.r_horizon([1.0, 0.9, 0.3, lam])
        d = (rh * rh + 0.81) * (1 - lam * rh * rh / 3) - 2 * rh + 0.09
        assert abs(d) < 1e-9 and abs(rh - (1 + np.sqrt(1 - 0.81 - 0.09))) < 1e-3


def test_isco_closed_forms():
    assert kto.r_isco([1.0, 0.0, 0.0, 0.0]) == 6.0
    # Bardeen et al. 1972: a -> 1 prograde ISCO -> 1; a = 0.5 -> 4.2330...
    assert abs(kto.r_isco([1.0, 0.999999, 0.0, 0.0]) - 1.0) < 0.05
    assert abs(kto.r_isco([1.0, 0.5, 0.0, 0.0]) - 4.233002531) < 1e-8


def test_numeric_isco_is_first_decrease_of_L():
    """For Q != 0 the scan of knds_metric.hpp:330-348 stops at the first r where L_circ
    decreases, i.e. one scan step outside the radius where the reference's Keplerian orbit
    (Omega^2 = (M - Q^2/(2r))/r^3, knds_metric.hpp:354-363) becomes null.  For a = 0 that
    radius solves r^2 - 3 M r + 1.5 Q^2 = 0."""
    for Q in (0.2, 0.4):
        rnull = (3.0 + np.sqrt(9.0 - 6.0 * Q * Q)) / 2.0
        r = kto.r_isco([1.0, 0.0, Q, 0.0])
        assert rnull < r < rnull + 2e-4, (Q, r, rnull)


def test_schwarzschild_shadow_edge():
    """Equatorial rays with impact parameter b < 3 sqrt(3) M fall in, b > 3 sqrt(3) M escape."""
    W = 400
    fr = kto.make_frame(W, 1, a=0.0, theta=90.0, r_obs=400.0, fov=3.0, disk_out=1.0)   # no disk
    px = np.arange(W, dtype=np.int32)
    py = np.zeros(W, np.int32)
    out = kto.trace_pixels(fr, px, py)
    bcrit = 3.0 * np.sqrt(3.0)
    for i in range(W):
        s = kto.camera_pixel([1.0, 0.0, 0.0, 0.0], [400.0, 90.0, 0.0, 3.0, W, 1], float(i), 0.0)
        b = abs(s[6] / s[5])
        if b < bcrit * 0.995:
            assert out["outcome"][i] == 2, (i, b)
        elif b > bcrit * 1.005:
            assert out["outcome"][i] == 0, (i, b)


def test_carter_constant_conserved_kerr():
    """Kerr (Q = Lambda = 0): Q_c = p_th^2 + cos^2 th (-a^2 E^2 + L^2/sin^2 th) is conserved."""
    mp = [1.0, 0.7, 0.0, 0.0]
    s = kto.camera_angles(mp, [40.0, 70.0, 0.0, 30.0, 320, 180], 0.05, 0.08)
    E, L = -s[5], s[6]

    def carter(st):
        c, sn = np.cos(st[1]), np.sin(st[1])
        return st[4] ** 2 + c * c * (-(0.7 ** 2) * E * E + L * L / (sn * sn))

    q0 = carter(s)
    h, fs = 1.0, None
    for _ in range(400):
        ok, s, h, fs = kto.adaptive_step(mp, "dopri5", s, h, fs, 1e-10)
        if s[0] > 60 or s[0] < 2.5:
            break
    assert abs(carter(s) - q0) < 1e-6 * max(1.0, abs(q0))


def test_dopri5_local_order():
    """One DOPRI5 step vs two half steps differ by O(h^6) (local error of a 5th-order
    method): halving h shrinks the difference by ~64x in the asymptotic range."""
    mp = [1.0, 0.5, 0.0, 0.0]
    s = kto.camera_angles(mp, [20.0, 80.0, 0.0, 30.0, 320, 180], 0.1, 0.05)
    diffs = []
    for h in (4.0, 2.0, 1.0):
        _, y1, _, _ = kto.adaptive_step(mp, "dopri5", s, h, None, 1e3)        # tol 1e3: always accepted
        _, ya, _, _ = kto.adaptive_step(mp, "dopri5", s, h / 2, None, 1e3)
        _, yb, _, _ = kto.adaptive_step(mp, "dopri5", ya, h / 2, None, 1e3)
        diffs.append(np.abs(y1[[0, 1, 3, 4]] - yb[[0, 1, 3, 4]]).max())
    r1, r2 = diffs[0] / diffs[1], diffs[1] / diffs[2]
    assert 30 < r1 < 100 and 30 < r2 < 100, diffs
    # RK4 step doubling: O(h^5) local error -> ~32x
    diffs = []
    for h in (2.0, 1.0, 0.5):
        _, y1, _, _ = kto.adaptive_step(mp, "rk4", s, h, None, 1e3)
        _, ya, _, _ = kto.adaptive_step(mp, "rk4", s, h / 2, None, 1e3)
        _, yb, _, _ = kto.adaptive_step(mp, "rk4", ya, h / 2, None, 1e3)
        diffs.append(np.abs(y1[[0, 1, 3, 4]] - yb[[0, 1, 3, 4]]).max())
    r1, r2 = diffs[0] / diffs[1], diffs[1] / diffs[2]
    assert 15 < r1 < 50 and 15 < r2 < 50, diffs
