"""GPU: the hand-written adjoint (checkpointed reverse pass) and its torch.autograd wrapper, at
the differentiable-render configuration: 256x256, Kerr-Newman a=0.7, Q=0.2, Lambda=0, KS,
theta_obs=60 deg, DOPRI5 (rtol 1e-9, atol 1e-12).

  GPU adjoint == host build of the same reverse pass (per pixel)
  per-pixel dI/dp vs central differences (h = 1e-6) of the GPU binary64 forward
  loss gradient: L = sum over interior disk pixels of (I(p) - I_target)^2, target rendered from
      seeded perturbed parameters (da, dQ ~ U(+-0.02), dtheta ~ U(+-2 deg)); adjoint vs central
      differences, 1e-3 relative, over the pixels whose intensity is smooth across the stencil
  torch: KerrIntensity.apply(p).backward() == the adjoint's gradient

"Interior disk pixels": I > 0 at the pixel and its 8 neighbours in the base image, so the fixed
mask holds no disk edge or photon-ring pixel (where I jumps and no derivative exists).
"""
import numpy as np
import pytest

import kerrtracecpp2_amd as kt
from tests import _hostcheck as hc

pytestmark = pytest.mark.gpu

KN = dict(a=0.7, Q=0.2, Lambda=0.0, theta=60.0, r_obs=40.0)
W = H = 256
NAMES = ["a", "Q", "Lambda", "theta", "r_obs"]


def _opts():
    o = kt.default_opts()
    o.chart, o.integrator, o.rtol, o.atol = kt.KT_CHART_KS, kt.KT_INTEGRATOR_DOPRI5, 1e-9, 1e-12
    return o


def _frame(**kw):
    p = dict(KN)
    p.update(kw)
    return kt.make_frame(W, H, fov=30.0, disk_out=12.0, **p)


def _interior(I):
    m = I > 0
    out = m.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            out &= np.roll(np.roll(m, dy, 0), dx, 1)
    out[0, :] = out[-1, :] = out[:, 0] = out[:, -1] = False
    return out


@pytest.fixture(scope="module")
def base(gpu_ctx):
    out = gpu_ctx.render_adjoint(_frame(), _opts(), want_pixel_grads=True)
    print("256x256 adjoint stats", out["stats"], flush=True)
    return out


def test_adjoint_records_and_counts(gpu_ctx, base):
    st = base["stats"]
    I = base["intensity"]
    assert st["seed_failures"] == 0
    assert st["rays_differentiated"] == int((I > 0).sum()) > 10000
    assert st["checkpoints"] * 16 >= st["steps"] and st["record_bytes"] == st["steps"] * 8 + st["checkpoints"] * 48
    # the forward pass parks the rays that pass 128 accepted steps: the frame has such rays, and
    # they cannot hold more than the steps past 128 that were recorded
    assert 0 < st["rays_parked"] and st["rays_parked"] * 129 <= st["steps"]
    I2, _ = gpu_ctx.render_intensity(_frame(), _opts())
    assert np.array_equal(I2, I)                  # the intensity pass and the adjoint's forward agree


def test_adjoint_record_stride_kept_between_calls(base):
    """The records sit at a per-pixel stride that the context keeps: a fresh context may run the
    forward pass twice on its first call (a ray longer than the starting stride), never on the
    next; a larger frame then grows the buffers.  Every call gives the same gradients bit for
    bit, since the binary64 path takes the same steps whatever the stride."""
    ctx = kt.Context(0)
    try:
        first = ctx.render_adjoint(_frame(), _opts(), want_pixel_grads=True)
        again = ctx.render_adjoint(_frame(), _opts(), want_pixel_grads=True)
        print("strides: first call", first["stats"], "second", again["stats"], flush=True)
        assert again["stats"]["record_ms"] == 0.0
        for out in (first, again):
            assert np.array_equal(out["dI_dp"], base["dI_dp"])
            assert np.array_equal(out["intensity"], base["intensity"])
            assert out["stats"]["steps"] == base["stats"]["steps"]
        big = kt.make_frame(320, 288, fov=30.0, disk_out=12.0, **KN)
        grown = ctx.render_adjoint(big, _opts(), want_pixel_grads=True)
    finally:
        ctx.close()
    ctx2 = kt.Context(0)
    try:
        fresh = ctx2.render_adjoint(big, _opts(), want_pixel_grads=True)
    finally:
        ctx2.close()
    assert np.array_equal(grown["dI_dp"], fresh["dI_dp"])
    assert grown["stats"]["rays_differentiated"] == fresh["stats"]["rays_differentiated"] > 10000


def test_grad_is_the_weighted_sum_of_pixel_grads(gpu_ctx, base):
    """grad_out is summed on the device in a fixed order: it equals the per-pixel dI/dp summed with
    the weights (to rounding), and repeats bit for bit."""
    dI = base["dI_dp"].reshape(-1, 5)
    np.testing.assert_allclose(base["grad"], dI.sum(axis=0), rtol=1e-12, atol=1e-12 * np.abs(dI).sum(axis=0).max())
    w = np.random.default_rng(5).standard_normal(W * H)
    w[::7] = 0.0
    out = gpu_ctx.render_adjoint(_frame(), _opts(), weights=w, want_pixel_grads=True)
    want = (w[:, None] * out["dI_dp"].reshape(-1, 5)).sum(axis=0)
    np.testing.assert_allclose(out["grad"], want, rtol=1e-11, atol=1e-12 * np.abs(out["dI_dp"]).sum(axis=(0, 1)).max())
    again = gpu_ctx.render_adjoint(_frame(), _opts(), weights=w)
    assert np.array_equal(again["grad"], out["grad"])


def test_gpu_adjoint_equals_host_adjoint(base):
    o = _opts()
    o.quirks |= 1 << 30
    rng = np.random.default_rng(11)
    idx = rng.choice(np.flatnonzero(base["intensity"].ravel() > 0), size=60, replace=False)
    px, py = (idx % W).astype(np.int32), (idx // W).astype(np.int32)
    I, G, _, _ = hc.adjoint_pixels(_frame(), o, px, py)
    gd = base["dI_dp"].reshape(-1, 5)[idx]
    np.testing.assert_allclose(base["intensity"].ravel()[idx], I, rtol=1e-7)
    scale = np.abs(G).max(axis=0) + 1e-30
    assert np.all(np.abs(gd - G) <= 1e-5 * scale + 1e-5 * np.abs(G))


def test_per_pixel_adjoint_vs_central_differences(gpu_ctx, base):
    o = _opts()
    G = base["dI_dp"]
    mask = _interior(base["intensity"])
    report = {}
    for q, k in enumerate(NAMES):
        if k == "Lambda":
            continue     # Lambda = 0 switches the KS initial covector form (reference quirk): no two-sided FD
        h = 1e-6
        Ip, _ = gpu_ctx.render_intensity(_frame(**{k: KN[k] + h}), o)
        Im, _ = gpu_ctx.render_intensity(_frame(**{k: KN[k] - h}), o)
        fd = (Ip - Im) / (2 * h)
        m = mask & (Ip > 0) & (Im > 0)
        err = np.abs(G[..., q] - fd)[m]
        rel = err / np.maximum(np.abs(fd[m]), 1e-4 * base["intensity"][m])
        report[k] = (float(np.quantile(rel, 0.99)), float(np.mean(rel <= 1e-3)))
    print("per-pixel rel err (q99, frac <= 1e-3)", report, flush=True)
    for k, (q99, frac) in report.items():
        assert frac >= 0.99, (k, frac)


def _loss_setup(gpu_ctx):
    rng = np.random.default_rng(2025)
    d = rng.uniform(-1, 1, 3) * np.array([0.02, 0.02, 2.0])
    target, _ = gpu_ctx.render_intensity(_frame(a=KN["a"] + d[0], Q=KN["Q"] + d[1], theta=KN["theta"] + d[2]),
                                         _opts())
    I0, _ = gpu_ctx.render_intensity(_frame(), _opts())
    return target, _interior(I0)


def test_c5_loss_gradient_vs_central_differences(gpu_ctx):
    """The loss gradient against central differences of the binary64 forward, h = 1e-6.

    The forward's adaptive step control makes each I(p) piecewise smooth: where a step is accepted
    on one side of p and rejected on the other, the step sequence changes, and with it the step in
    which the ray crosses the disk and the Hermite interpolation of the hit radius inside it.  I
    then jumps, by up to ~2e-5 I at this configuration, which over 2h = 2e-6 is a slope of order
    10 I: one such pixel moved the loss's difference quotient by 3.6e-3 in r_obs.  A pixel whose
    difference quotients at h and 2h disagree has a jump inside the stencil; those pixels (a few
    in 30 000) leave both sides of the comparison: the difference quotient sums the others, and
    the adjoint is weighted over the same pixels."""
    target, mask = _loss_setup(gpu_ctx)
    I0, _ = gpu_ctx.render_intensity(_frame(), _opts())
    w = 2.0 * (I0 - target) * mask
    out = gpu_ctx.render_adjoint(_frame(), _opts(), weights=w, want_pixel_grads=True)
    G = out["dI_dp"]
    rel, dropped = {}, {}
    for q, k in enumerate(NAMES):
        if k == "Lambda":
            continue
        h = 1e-6
        I = {m: gpu_ctx.render_intensity(_frame(**{k: KN[k] + m * h}), _opts())[0] for m in (-2, -1, 1, 2)}
        d1 = (I[1] - I[-1]) / (2 * h)
        d2 = (I[2] - I[-2]) / (4 * h)
        jump = mask & (np.abs(d1 - d2) > 1e-3 * np.abs(d2) + 1e-6 * np.abs(d2[mask]).max())
        keep = mask & ~jump
        fd = float(((I[1] - target) ** 2 - (I[-1] - target) ** 2)[keep].sum()) / (2 * h)
        g = float((w * G[..., q])[keep].sum())
        rel[k] = abs(g - fd) / abs(fd)
        dropped[k] = int(jump.sum())
    print("256x256 loss gradient: rel err vs FD", rel, "pixels with a jump in the stencil", dropped,
          "of", int(mask.sum()), flush=True)
    assert max(rel.values()) <= 1e-3
    assert max(dropped.values()) <= 1e-3 * mask.sum()
    # the full adjoint gradient is the weighted sum over every pixel
    np.testing.assert_allclose(out["grad"], (w[..., None] * G).sum(axis=(0, 1)), rtol=1e-10,
                               atol=1e-12 * np.abs(w[..., None] * G).sum(axis=(0, 1)).max())


def test_torch_autograd_function(gpu_ctx):
    torch = pytest.importorskip("torch")
    from kerrtracecpp2_amd.autograd import KerrIntensity
    target, mask = _loss_setup(gpu_ctx)
    spec = dict(width=W, height=H, fov=30.0, disk_out=12.0)
    p = torch.tensor([KN[k] for k in NAMES], dtype=torch.float64, requires_grad=True)
    img = KerrIntensity.apply(p, gpu_ctx, spec, _opts())
    tt, mm = torch.from_numpy(target), torch.from_numpy(mask.astype(np.float64))
    loss = (((img - tt) ** 2) * mm).sum()
    loss.backward()
    w = 2.0 * (img.detach().numpy() - target) * mask
    want = gpu_ctx.render_adjoint(_frame(), _opts(), weights=w)["grad"]
    np.testing.assert_array_equal(p.grad.numpy(), want)
    assert np.all(np.isfinite(p.grad.numpy())) and abs(p.grad.numpy()[3]) > 0


@pytest.mark.parametrize("chart,kw", [
    (kt.KT_CHART_BL, dict(a=0.6, Q=0.25, Lambda=5e-6, theta=75.0, r_obs=30.0)),
    (kt.KT_CHART_KS, dict(a=0.95, Q=0.0, Lambda=2e-6, theta=40.0, r_obs=25.0)),
])
def test_gpu_adjoint_other_spacetimes_equal_host_adjoint(gpu_ctx, chart, kw):
    """The segment-parallel reverse pass in the BL chart and at Lambda != 0, against the host
    build's sequential sweep per pixel (intensity and dI/dp), on 96 seeded disk pixels of a
    160x120 frame."""
    fr = kt.make_frame(160, 120, fov=40.0, disk_out=14.0, **kw)
    o = _opts()
    o.chart = chart
    out = gpu_ctx.render_adjoint(fr, o, want_pixel_grads=True)
    st = out["stats"]
    I_gpu = out["intensity"].ravel()
    assert st["seed_failures"] == 0 and st["rays_differentiated"] == int((I_gpu > 0).sum()) > 500
    ho = _opts()
    ho.chart = chart
    ho.quirks |= 1 << 30
    rng = np.random.default_rng(23 + chart)
    idx = rng.choice(np.flatnonzero(I_gpu > 0), size=96, replace=False)
    px, py = (idx % 160).astype(np.int32), (idx // 160).astype(np.int32)
    I, G, _, _ = hc.adjoint_pixels(fr, ho, px, py)
    np.testing.assert_allclose(I_gpu[idx], I, rtol=1e-7)
    gd = out["dI_dp"].reshape(-1, 5)[idx]
    scale = np.abs(G).max(axis=0) + 1e-30
    assert np.all(np.abs(gd - G) <= 1e-5 * scale + 1e-5 * np.abs(G))


def test_gpu_reverse_pass_every_disk_pixel_of_a_close_frame(gpu_ctx):
    """The GPU's reverse pass lays its entries out by cost class (full segments; last segments of
    16 .. 1 steps; event entries replaying 16 .. 1 steps).  Every disk pixel of a 96x96 frame with
    the camera at 14 M, where many rays end within a few steps, against the host build's
    sequential sweep: intensity and dI/dp.  The pixels cover every class, rays with no full
    segment, and step counts of 16 k + 1 and 16 k + 16."""
    fr = kt.make_frame(96, 96, fov=60.0, disk_out=12.0, a=0.7, Q=0.2, Lambda=0.0, theta=75.0, r_obs=14.0)
    o = _opts()
    out = gpu_ctx.render_adjoint(fr, o, want_pixel_grads=True)
    I_gpu = out["intensity"].ravel()
    idx = np.flatnonzero(I_gpu > 0)
    assert out["stats"]["seed_failures"] == 0 and idx.size > 1000
    ho = _opts()
    ho.quirks |= 1 << 30
    px, py = (idx % 96).astype(np.int32), (idx // 96).astype(np.int32)
    I, G, N, _ = hc.adjoint_pixels(fr, ho, px, py)
    np.testing.assert_allclose(I_gpu[idx], I, rtol=1e-7)
    gd = out["dI_dp"].reshape(-1, 5)[idx]
    scale = np.abs(G).max(axis=0) + 1e-30
    assert np.all(np.abs(gd - G) <= 1e-5 * scale + 1e-5 * np.abs(G))
    nseg = (N - 1 + 15) // 16
    last = (N - 1) - 16 * (nseg - 1)
    replay = N - 16 * ((N - 1) // 16)
    print("disk pixels", idx.size, "steps", N.min(), N.max(), flush=True)
    assert (nseg == 1).any() and ((N % 16) == 1).any() and ((N % 16) == 0).any()
    assert set(last[nseg >= 1]) == set(range(1, 17)) and set(replay) == set(range(1, 17))
