"""Seeded random spacetimes and cameras against the oracle, beyond the fixed configurations:
spin in (-0.95, 0.95), charge up to 0.4 sqrt(1 - a^2), Lambda 0 or up to 1e-5, inclination 20-89
deg, distance 20-80 M, disk edge 8-20 M; 96x54 pixels, DOPRI5 at tol 1e-9 (atol = rtol, the
reference's semantics), Boyer-Lindquist and Kerr-Schild charts, binary64 and mixed; RK4 step
doubling at tol 1e-7 in BL, and in KS where Lambda = 0; and all 16 in one batched plan.

The oracle runs what the reference runs: BL DOPRI5; for KS at Lambda = 0 its own KS path (the
initial covector without the shift terms, main.cpp:1373-1430, and RK4, main.cpp:1539-1545), which
the kernels reproduce by default; for KS at Lambda != 0 BL again (main.cpp:2034).  Fate, r_hit and
g do not depend on the chart.  With a < 0 and Q or Lambda != 0 the reference's ISCO scan (from
1.05 r_+ up to 30 M, knds_metric.hpp:330-348) finds no minimum and returns 30 M, so the disk is
empty there; the test checks that too."""
import numpy as np
import pytest

import kerrtracecpp2_amd as kt
from oracle import kto
from tests._util import photon_ring_band

pytestmark = pytest.mark.gpu

W, H = 96, 54


def _configs(n=16, seed=2026):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(-0.95, 0.95)
        Q = rng.uniform(0, 0.4) * np.sqrt(1 - a * a)
        Lam = float(rng.choice([0.0, rng.uniform(0, 1e-5)]))
        out.append(dict(a=a, Q=Q, Lambda=Lam, theta=rng.uniform(20, 89), r_obs=rng.uniform(20, 80),
                        disk_out=rng.uniform(8, 20)))
    return out


CONFIGS = _configs()


def _opts(chart, prec):
    o = kt.default_opts()
    o.chart, o.integrator, o.precision = chart, kt.KT_INTEGRATOR_DOPRI5, prec
    o.rtol = o.atol = 1e-9
    return o


@pytest.mark.parametrize("i", range(len(CONFIGS)))
def test_random_config_vs_oracle(gpu_ctx, i):
    cfg = CONFIGS[i]
    bl = kto.trace_frame(kto.make_frame(W, H, chart="bl", integrator="dopri5", tol=1e-9, **cfg))
    ks = bl if cfg["Lambda"] != 0.0 else kto.trace_frame(kto.make_frame(W, H, chart="ks", integrator="dopri5",
                                                                         tol=1e-9, **cfg))
    r_isco = kto.r_isco([1.0, cfg["a"], cfg["Q"], cfg["Lambda"]])
    assert ((bl["outcome"] == 1).any()) == (r_isco < cfg["disk_out"])
    for chart, ref in ((kt.KT_CHART_BL, bl), (kt.KT_CHART_KS, ks)):
        keep = ~photon_ring_band(ref, W, H)
        for prec in (kt.KT_PRECISION_FP64, kt.KT_PRECISION_MIXED):
            g = gpu_ctx.render([kt.make_frame(W, H, **cfg)], _opts(chart, prec), want_rgba=False)["geo"]
            what = f"config {i} {cfg} chart {chart} precision {prec}"
            assert (ref["outcome"] == g["outcome"])[keep].mean() >= 0.999, what
            d = keep & (ref["outcome"] == 1) & (g["outcome"] == 1)
            if d.any():
                assert np.abs(ref["redshift"][d] - g["redshift"][d]).max() <= 1e-4, what
                assert np.abs(ref["r"][d] - g["r"][d]).max() <= 1e-3, what


@pytest.mark.parametrize("i", range(len(CONFIGS)))
def test_random_config_rk4_vs_oracle(gpu_ctx, i):
    """The same spacetimes with RK4 step doubling at the reference's default tolerance (1e-7, BL):
    |dg| <= 5e-4 and |dr| <= 2e-3, the tol-1e-7 bars."""
    cfg = CONFIGS[i]
    ref = kto.trace_frame(kto.make_frame(W, H, chart="bl", integrator="rk4", tol=1e-7, **cfg))
    keep = ~photon_ring_band(ref, W, H)
    for prec in (kt.KT_PRECISION_FP64, kt.KT_PRECISION_MIXED):
        o = kt.default_opts()
        o.chart, o.integrator, o.precision = kt.KT_CHART_BL, kt.KT_INTEGRATOR_RK4, prec
        o.rtol = o.atol = 1e-7
        g = gpu_ctx.render([kt.make_frame(W, H, **cfg)], o, want_rgba=False)["geo"]
        what = f"config {i} {cfg} precision {prec}"
        assert (ref["outcome"] == g["outcome"])[keep].mean() >= 0.999, what
        d = keep & (ref["outcome"] == 1) & (g["outcome"] == 1)
        if d.any():
            assert np.abs(ref["redshift"][d] - g["redshift"][d]).max() <= 5e-4, what
            assert np.abs(ref["r"][d] - g["r"][d]).max() <= 2e-3, what


def test_random_configs_in_one_plan_equal_single_renders(gpu_ctx):
    """All 16 spacetimes in one graph-captured plan (the multi-frame kernel, per-lane frame
    constants, one setup pass for 16 horizons and ISCOs) against one-frame renders: binary64
    GeoPixels equal bit for bit, twice in a row."""
    frames = [kt.make_frame(W, H, **cfg) for cfg in CONFIGS]
    o = _opts(kt.KT_CHART_KS, kt.KT_PRECISION_FP64)
    singles = [gpu_ctx.render([f], o, want_rgba=False)["geo"] for f in frames]
    plan = gpu_ctx.plan(o, len(frames), W * H * len(frames))
    try:
        for _ in range(2):
            geo = plan.render(frames, want_rgba=False)["geo"]
            for k, g in enumerate(singles):
                assert geo[k * W * H:(k + 1) * W * H].tobytes() == g.tobytes(), f"frame {k}"
    finally:
        plan.close()


KS_CASES = [i for i, c in enumerate(CONFIGS) if c["Lambda"] == 0.0]


@pytest.mark.parametrize("i", KS_CASES)
def test_random_config_ks_rk4_vs_oracle(gpu_ctx, i):
    """RK4 step doubling at tol 1e-7 in the Kerr-Schild chart, for the spacetimes with Lambda = 0
    (where the reference runs its KS path, main.cpp:1539-1545), against the oracle's KS path:
    the tol-1e-7 bars."""
    cfg = CONFIGS[i]
    ref = kto.trace_frame(kto.make_frame(W, H, chart="ks", integrator="rk4", tol=1e-7, **cfg))
    keep = ~photon_ring_band(ref, W, H)
    for prec in (kt.KT_PRECISION_FP64, kt.KT_PRECISION_MIXED):
        o = kt.default_opts()
        o.chart, o.integrator, o.precision = kt.KT_CHART_KS, kt.KT_INTEGRATOR_RK4, prec
        o.rtol = o.atol = 1e-7
        g = gpu_ctx.render([kt.make_frame(W, H, **cfg)], o, want_rgba=False)["geo"]
        what = f"config {i} {cfg} precision {prec}"
        assert (ref["outcome"] == g["outcome"])[keep].mean() >= 0.999, what
        d = keep & (ref["outcome"] == 1) & (g["outcome"] == 1)
        if d.any():
            assert np.abs(ref["redshift"][d] - g["redshift"][d]).max() <= 5e-4, what
            assert np.abs(ref["r"][d] - g["r"][d]).max() <= 2e-3, what
