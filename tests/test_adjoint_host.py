"""The adjoint's arithmetic on the CPU (the kernels' per-ray code built for the host).

  local Jacobian ........... dual-number J of the right-hand side vs central differences
  templated camera map ..... bit-identical to the kernels' camera + covector (T = double)
  fixed-step reverse pass .. vs central differences of the same fixed-step replay (all 9
                             inputs: state, conserved p_t / p_phi, a, Q, Lambda)
  full adaptive adjoint .... dI/dp vs central differences of the binary64 forward, h = 1e-6,
                             at the differentiable-render configuration (KS, a=0.7, Q=0.2,
                             theta=60) and at Lambda != 0, BL and Kerr (Bardeen ISCO) frames;
                             bar: 1e-3 relative
"""
import numpy as np
import pytest

import kerrtracecpp2_amd as kt
from tests import _hostcheck as hc

EXACT_ROOT = 1 << 30          # internal mode of the differentiable path (converged disk root)
PX = np.array([128, 100, 150, 60, 200, 128, 90, 170, 140, 110, 180])
PY = np.array([128, 140, 110, 128, 120, 160, 100, 150, 170, 125, 135])


def _opts(chart=1, rtol=1e-9, atol=1e-12):
    o = kt.default_opts()
    o.chart, o.integrator, o.rtol, o.atol = chart, kt.KT_INTEGRATOR_DOPRI5, rtol, atol
    o.quirks |= EXACT_ROOT
    return o


@pytest.mark.parametrize("chart", [0, 1])
def test_rhs_jacobian_vs_central_differences(chart):
    rng = np.random.default_rng(chart)
    for _ in range(5):
        par = [rng.uniform(-0.9, 0.9), rng.uniform(0, 0.3), rng.uniform(0, 1e-5)]
        y = [rng.uniform(3, 30), rng.uniform(0.3, 2.8), rng.uniform(-2, 2), rng.uniform(-5, 5)]
        pt, pp = -1.0, rng.uniform(-4, 4)
        J = hc.rhs_jacobian(chart, 1.0, par, y, pt, pp)
        x0 = np.array(y + [pt, pp] + par)
        fd = np.zeros((4, 9))
        for k in range(9):
            e = max(abs(x0[k]), 1e-3) * 1e-6
            xp, xm = x0.copy(), x0.copy()
            xp[k] += e
            xm[k] -= e
            fd[:, k] = (hc.rhs(chart, 1.0, xp[6:], xp[:4], xp[4], xp[5]) -
                        hc.rhs(chart, 1.0, xm[6:], xm[:4], xm[4], xm[5])) / (2 * e)
        assert np.abs(J - fd).max() <= 1e-6 * np.abs(J).max()


@pytest.mark.parametrize("chart,Lam", [(0, 0.0), (1, 0.0), (1, 1e-5)])
def test_templated_camera_is_the_kernel_camera(chart, Lam):
    fr = kt.make_frame(64, 36, a=0.9, Q=0.3, Lambda=Lam, theta=70, fov=40)
    px, py = np.arange(0, 64, 5), np.arange(0, 36, 3)
    px, py = np.meshgrid(px, py)
    a, b = hc.camera_pair(fr, _opts(chart), px.ravel(), py.ravel())
    assert np.array_equal(a, b)


@pytest.mark.parametrize("pix", [(60, 128), (128, 160), (200, 120)])
def test_fixed_step_reverse_pass(pix):
    o = _opts()
    fr = kt.make_frame(256, 256, a=0.7, Q=0.2, theta=60)
    hs, y0c, fate = hc.ray_steps(fr, o, *pix)
    assert fate == 1
    par = np.array([0.7, 0.2, 0.0])
    I, lam = hc.replay_intensity(fr, o, par, y0c[:4], y0c[4], y0c[5], hs, adjoint=True)
    x0 = np.concatenate([y0c, par])
    fd = np.zeros(9)
    for k in range(9):
        e = max(abs(x0[k]), 1e-2) * 1e-7
        xp, xm = x0.copy(), x0.copy()
        xp[k] += e
        xm[k] -= e
        Ip, _ = hc.replay_intensity(fr, o, xp[6:], xp[:4], xp[4], xp[5], hs)
        Im, _ = hc.replay_intensity(fr, o, xm[6:], xm[:4], xm[4], xm[5], hs)
        fd[k] = (Ip - Im) / (2 * e)
    np.testing.assert_allclose(lam, fd, rtol=1e-5, atol=1e-9 * abs(I))


def _adjoint_vs_fd(base, keys, h, chart=1, px=PX, py=PY):
    o = _opts(chart)
    W = H = 256

    def frame(**kw):
        p = dict(base)
        p.update(kw)
        return kt.make_frame(W, H, **p)

    I, G, N, _ = hc.adjoint_pixels(frame(), o, px, py)
    m = I > 0
    assert m.sum() >= 4
    names = ["a", "Q", "Lambda", "theta", "r_obs"]
    worst = 0.0
    for k in keys:
        q = names.index(k)
        Ip, _, _, _ = hc.adjoint_pixels(frame(**{k: base[k] + h[k]}), o, px, py, False)
        Im, _, _, _ = hc.adjoint_pixels(frame(**{k: base[k] - h[k]}), o, px, py, False)
        fd = (Ip - Im) / (2 * h[k])
        rel = np.abs(G[m, q] - fd[m]) / np.maximum(np.abs(fd[m]), 1e-6 * I[m])
        worst = max(worst, float(rel.max()))
    return worst


def test_adjoint_vs_fd_kerr_newman_ks():
    base = dict(a=0.7, Q=0.2, Lambda=0.0, theta=60.0, r_obs=40.0)
    assert _adjoint_vs_fd(base, ["a", "Q", "theta", "r_obs"], dict(a=1e-6, Q=1e-6, theta=1e-6, r_obs=1e-6)) <= 1e-3


def test_adjoint_vs_fd_lambda():
    base = dict(a=0.7, Q=0.2, Lambda=5e-6, theta=60.0, r_obs=40.0)
    assert _adjoint_vs_fd(base, ["Lambda", "a"], dict(Lambda=1e-8, a=1e-6)) <= 1e-3


def test_adjoint_vs_fd_bl_and_kerr_isco():
    # BL chart; Q = Lambda = 0: the ISCO is the Bardeen closed form, differentiated in a
    base = dict(a=0.9, Q=0.0, Lambda=0.0, theta=75.0, r_obs=35.0)
    assert _adjoint_vs_fd(base, ["a", "theta", "r_obs"], dict(a=1e-6, theta=1e-6, r_obs=1e-6), chart=0) <= 1e-3


@pytest.mark.parametrize("chart", [0, 1])
def test_rhs_vjp_equals_forward_jacobian(chart):
    """The adjoint's hand-written reverse mode of the right-hand side (rhs_vjp) against the
    forward-mode Jacobian in dual numbers (rhs_jacobian): J^T mu over (r, theta, p_r, p_theta, p_t,
    p_phi, a, Q, Lambda) for random states, spacetimes (Lambda 0 and 1e-5 scale) and covectors.
    A dropped or sign-flipped term in the reverse sweep shows as an O(1) error."""
    rng = np.random.default_rng(11 + chart)
    for _ in range(500):
        a = rng.uniform(-0.95, 0.95)
        Q = rng.uniform(0, 0.4) * np.sqrt(1 - a * a)
        L = rng.choice([0.0, rng.uniform(0, 1e-5)])
        y = [rng.uniform(2.5, 40), rng.uniform(0.2, np.pi - 0.2), rng.normal(), 3 * rng.normal()]
        pt, pp = -1.0, 3 * rng.normal()
        mu = rng.normal(size=4)
        J = hc.rhs_jacobian(chart, 1.0, [a, Q, L], y, pt, pp)
        g = hc.rhs_vjp(chart, 1.0, [a, Q, L], y, pt, pp, mu)
        scale = np.abs(J.T) @ np.abs(mu)
        assert np.all(np.abs(g - J.T @ mu) <= 1e-10 * scale + 1e-300)


@pytest.mark.parametrize("chart", [0, 1])
def test_segmented_reverse_pass_equals_the_sequential_sweep(chart):
    """The GPU's reverse pass (each checkpoint segment swept once per basis direction of the
    adjoint at its end, the last step with the disk seed as an entry of its own, then the
    pixel's chain composed from these linear maps) against one sequential sweep per pixel: the
    same derivative summed in another order."""
    rng = np.random.default_rng(7 + chart)
    o = _opts(chart)
    cases = []
    # the frame of the differentiable render, and a camera at 14 M whose disk rays end in a few steps
    for kw in (dict(a=0.7, Q=0.2, theta=60), dict(a=0.7, Q=0.2, theta=75, r_obs=14.0, fov=60)):
        fr = kt.make_frame(256, 256, **kw)
        px, py = rng.integers(0, 256, 500), rng.integers(0, 256, 500)
        I, G, N, _ = hc.adjoint_pixels(fr, o, px, py)
        I2, G2, N2, _ = hc.adjoint_pixels(fr, o, px, py, segmented=True)
        np.testing.assert_array_equal(N, N2)
        np.testing.assert_array_equal(I, I2)
        m = I > 0
        assert m.sum() > 100
        assert np.isfinite(G[m]).all() and np.isfinite(G2[m]).all()
        scale = np.abs(G[m]).max(axis=0)
        assert np.all(np.abs(G2[m] - G[m]) <= 1e-9 * np.abs(G[m]) + 1e-12 * scale)
        cases.append(N[m])
    n = np.concatenate(cases)
    r = (n - 1) % 16
    # the split's edge cases are among them: the event in the first segment (no segment before it),
    # a last segment of 16 steps with a one-step event entry, and a 16-step event replay
    assert (n <= 16).any() and (n > 16).any() and (r == 0).any() and (r == 15).any()


def _richardson(f, x0, k, e):
    """Central difference in input k with one Richardson step (error O(e^4))."""
    xs = [x0.copy() for _ in range(4)]
    xs[0][k] += e
    xs[1][k] -= e
    xs[2][k] += 2 * e
    xs[3][k] -= 2 * e
    d1 = (f(xs[0]) - f(xs[1])) / (2 * e)
    d2 = (f(xs[2]) - f(xs[3])) / (4 * e)
    return (4 * d1 - d2) / 3


def _random_point(rng):
    a = rng.uniform(-0.95, 0.95)
    par = np.array([a, rng.uniform(0, 0.4) * np.sqrt(1 - a * a), rng.choice([0.0, rng.uniform(0, 1e-5)])])
    y = np.array([rng.uniform(4, 30), rng.uniform(0.3, np.pi - 0.3), 0.5 * rng.normal(), 2 * rng.normal()])
    return par, y, -1.0, 3 * rng.normal()


@pytest.mark.parametrize("chart", [0, 1])
def test_rhs_vjp_sin_cos_inputs(chart):
    """rhs_vjp_sc's adjoints of sin and cos taken as independent inputs (the integrated stage
    sin/cos are not on the unit circle) against central differences of the right-hand side in s
    and in c separately, at points off the circle."""
    rng = np.random.default_rng(31 + chart)
    for _ in range(200):
        par, y, pt, pp = _random_point(rng)
        s, c, mu = rng.uniform(0.2, 1.0), rng.uniform(-1.0, 1.0), rng.normal(size=4)
        g = hc.rhs_vjp_sc(chart, 1.0, par, y, s, c, pt, pp, mu)
        f = lambda x: mu @ hc.rhs_sc(chart, 1.0, par, y, x[0], x[1], pt, pp)
        x0 = np.array([s, c])
        fd = np.array([_richardson(f, x0, k, 1e-4) for k in range(2)])
        nat = np.abs(fd).max() + np.abs(mu) @ np.abs(hc.rhs_sc(chart, 1.0, par, y, s, c, pt, pp))
        assert np.all(np.abs(g - fd) <= 1e-9 * nat)


@pytest.mark.parametrize("chart", [0, 1])
def test_dp5_step_adjoint_vs_central_differences(chart):
    """The adjoint of one DOPRI5 step (dp5_step_adjoint: the integrated stage sin/cos carried as
    adjoint components) against central differences of the forward pass's own step arithmetic
    (dp5_replay) in all nine inputs: state, p_t, p_phi, a, Q, Lambda.  Steps of 0.3 % to 30 % of
    r; the errors are measured on each input's natural scale (r, 1, 1, 1, 1, 1, 0.1, 0.1, 1e-5)
    against the largest such sensitivity.  An adjoint of a per-stage rotation instead of the
    integrated sin/cos differs from the step's derivative at the level of its truncation error
    and fails the 1e-9 bar."""
    rng = np.random.default_rng(41 + chart)
    for n in range(150):
        par, y, pt, pp = _random_point(rng)
        h = (0.003, 0.03, 0.3)[n % 3] * y[0] * rng.uniform(0.5, 1.0)
        lam = rng.normal(size=4)
        g = hc.dp5_step_adjoint(chart, 1.0, par, y, pt, pp, h, lam)
        x0 = np.concatenate([y, [pt, pp], par])
        scale = np.array([y[0], 1, 1, 1, 1, 1, 0.1, 0.1, 1e-5])
        f = lambda x: lam @ hc.dp5_step(chart, 1.0, x[6:], x[:4], x[4], x[5], h)
        fd = np.array([_richardson(f, x0, k, 1e-4 * scale[k]) for k in range(9)])
        nat = np.max(np.abs(fd) * scale)
        assert np.all(np.abs(g - fd) * scale <= 1e-9 * nat), (n, h, (g - fd) * scale / nat)
