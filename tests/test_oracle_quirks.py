"""The oracle's reference-behaviour switches (oracle/kt_oracle.cpp QK_*, oracle/kto.py QUIRK_*).

With every bit set the oracle is bit-exact to the reference binary (tests/test_oracle_golden.py).
Each cleared bit runs a corrected behaviour that the reference never runs, so no reference output
pins it; these tests pin each one against an independent path of the same physics:

  DOPRI5 phi advanced (geodesic.hpp:300-310) ... the RK4 doubling integrator, which advances phi
                                                 (geodesic.hpp:184-191), at a tight tolerance
  non-positional escape fill (main.cpp:360,428-433) ... a permutation of the positional fill
  exact KS initial covector (main.cpp:1373-1430) ... the BL chart: the same photon, so the
                                                 chart-invariant outputs (fate, r_hit, g) agree
  bundle horizon / root skip (ray_bundle.hpp:357-365,426-428) ... the plain BL tracer: with both
                                                 bits cleared a bundle's geodesic and events are
                                                 the plain tracer's, bit for bit
"""
import numpy as np
import pytest

from oracle import kto

R = kto.QUIRKS_REFERENCE
W, H = 48, 27


def _frame(**kw):
    return kto.make_frame(W, H, **kw)


def test_dopri5_phi_advanced_matches_rk4_phi():
    q = R & ~kto.QUIRK_DOPRI5_PHI_FROZEN & ~kto.QUIRK_ESCAPE_POSITIONAL
    gd = kto.trace_frame(_frame(a=0.6, theta=75, integrator="dopri5", tol=1e-10, quirks=q))
    gr = kto.trace_frame(_frame(a=0.6, theta=75, integrator="rk4", tol=1e-10, quirks=q))
    assert np.array_equal(gd["outcome"], gr["outcome"])
    esc = gd["outcome"] == 0
    assert esc.sum() > 200
    # escape angles come from a linear interpolation to r_escape between steps, which the two
    # integrators place differently: agreement to ~1e-4, against phi itself of order 1
    assert np.abs(gd["phi_esc"][esc] - gr["phi_esc"][esc]).max() < 5e-4
    assert np.abs(gd["theta_esc"][esc] - gr["theta_esc"][esc]).max() < 5e-4
    assert np.abs(gr["phi_esc"][esc]).max() > 0.5
    disk = gd["outcome"] == 1
    assert np.abs(gd["phi_disk"][disk] - gr["phi_disk"][disk]).max() < 1e-3
    # with the bit set phi stays at the observer's (0 here)
    gz = kto.trace_frame(_frame(a=0.6, theta=75, integrator="dopri5", tol=1e-10,
                                quirks=R & ~kto.QUIRK_ESCAPE_POSITIONAL))
    assert np.all(gz["phi_esc"][gz["outcome"] == 0] == 0.0)
    assert np.all(gz["phi_disk"][gz["outcome"] == 1] == 0.0)


@pytest.mark.parametrize("chart", ["bl", "ks"])
def test_escape_fill_positional_is_a_permutation(chart):
    kw = dict(a=0.5, theta=80, chart=chart, integrator="rk4", tol=1e-8)
    on = kto.trace_frame(_frame(quirks=R, **kw))
    off = kto.trace_frame(_frame(quirks=R & ~kto.QUIRK_ESCAPE_POSITIONAL, **kw))
    assert np.array_equal(on["outcome"], off["outcome"])
    esc = on["outcome"] == 0
    assert esc.sum() > 100
    assert np.array_equal(off["theta_esc"][esc], on["phi_disk"][esc])
    assert np.array_equal(off["phi_esc"][esc], on["theta_esc"][esc])
    assert np.all(on["phi_esc"][esc] == 0.0) and np.all(off["phi_disk"][esc] == 0.0)
    # theta at escape is near the observer's latitude band, not an azimuth
    assert np.all((off["theta_esc"][esc] > 0.0) & (off["theta_esc"][esc] < np.pi))
    rest = ~esc
    for f in ("r", "redshift", "phi_disk"):
        assert np.array_equal(on[f][rest], off[f][rest])


def test_ks_exact_initial_covector_is_the_bl_photon():
    kw = dict(a=0.5, Q=0.1, theta=80, tol=1e-9)
    res = {}
    for name, q in (("ref", R), ("exact", R & ~kto.QUIRK_KS_INIT_NO_SHIFT)):
        gk = kto.trace_frame(_frame(chart="ks", quirks=q, **kw))
        gb = kto.trace_frame(_frame(chart="bl", quirks=q, **kw))
        d = (gk["outcome"] == 1) & (gb["outcome"] == 1)
        res[name] = ((gk["outcome"] == gb["outcome"]).mean(), np.abs(gk["r"][d] - gb["r"][d]).max(),
                     np.abs(gk["redshift"][d] - gb["redshift"][d]).max())
    fate, dr, dg = res["exact"]
    assert fate == 1.0 and dr < 1e-3 and dg < 1e-4
    # the reference's covector (shift terms left out) is a different photon of the same pixel
    fate_ref, dr_ref, _ = res["ref"]
    assert fate_ref < 0.98 and dr_ref > 1.0


def test_bundles_without_quirks_equal_the_plain_tracer():
    kw = dict(a=0.99, theta=89)
    gb = kto.trace_frame(kto.make_frame(32, 18, bundles=True,
                                        quirks=R & ~kto.QUIRK_BUNDLE_HORIZON_ESC & ~kto.QUIRK_BUNDLE_ROOT_SKIP, **kw))
    gp = kto.trace_frame(kto.make_frame(32, 18, quirks=R & ~kto.QUIRK_ESCAPE_POSITIONAL, **kw))
    assert np.array_equal(gb["outcome"], gp["outcome"])
    assert set(np.unique(gp["outcome"])) == {0, 1, 2}
    for f, sel in (("r", 1), ("phi_disk", 1), ("theta_esc", 0), ("phi_esc", 0)):
        m = gp["outcome"] == sel
        assert np.array_equal(gb[f][m], gp[f][m]), f
    # the reference's bundle pixel of a horizon ray: outcome 0, r 0, angles 0
    gq = kto.trace_frame(kto.make_frame(32, 18, bundles=True, quirks=R & ~kto.QUIRK_BUNDLE_ROOT_SKIP, **kw))
    hole = gb["outcome"] == 2
    assert np.all(gq["outcome"][hole] == 0) and np.all(gq["r"][hole] == 0)
    assert np.all(gq["theta_esc"][hole] == 0) and np.all(gq["phi_esc"][hole] == 0)
    assert np.all(gb["r"][hole] == 0) and np.all(gb["redshift"][hole] == 0)
    assert np.array_equal(gq["outcome"][~hole], gb["outcome"][~hole])


def test_bundle_root_skip_changes_only_escape_bookkeeping():
    """With the root-skip bit set a near-equatorial step whose Hermite search fails skips the
    horizon and escape tests (ray_bundle.hpp:357-365), so rays end later or run to the step cap;
    disk hits found by the search are the same either way."""
    kw = dict(a=0.99, theta=89, max_steps=20000)
    on = kto.trace_frame(kto.make_frame(16, 9, bundles=True, quirks=R, **kw))
    off = kto.trace_frame(kto.make_frame(16, 9, bundles=True, quirks=R & ~kto.QUIRK_BUNDLE_ROOT_SKIP, **kw))
    both = (on["outcome"] == 1) & (off["outcome"] == 1)
    assert both.sum() > 20
    assert np.array_equal(on["r"][both], off["r"][both])
    esc_on = (on["outcome"] == 0) & (on["theta_esc"] != 0)
    esc_off = (off["outcome"] == 0) & (off["theta_esc"] != 0)
    assert esc_off.sum() >= esc_on.sum()
