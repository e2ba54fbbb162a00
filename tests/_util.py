"""Shared helpers of the test suite."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_manifest():
    import json
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def fate_boundary_band(fate, W, H):
    """Pixels with any 8-neighbour of a different fate in the oracle image: the photon-ring /
    disk-edge / shadow-edge band excluded from per-pixel tolerance checks."""
    f = np.asarray(fate).reshape(H, W)
    band = np.zeros_like(f, dtype=bool)
    pad = np.pad(f, 1, mode="edge")
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            band |= pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] != f
    return band.ravel()


def angle_diff(a, b):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return np.abs((d + np.pi) % (2 * np.pi) - np.pi)




def steep_redshift_band(geo, W, H, thr=0.05):
    """Disk pixels whose oracle redshift differs from an 8-neighbour disk pixel by more than thr:
    near-critical rays on the disk's inner rim next to the photon orbit, where g varies by ~3 per M
    of r_hit and the oracle's own error at tol 1e-9 exceeds 1e-4."""
    g = np.where(np.asarray(geo["outcome"]) == 1, np.asarray(geo["redshift"], np.float64), np.nan).reshape(H, W)
    pad = np.pad(g, 1, constant_values=np.nan)
    band = np.zeros((H, W), dtype=bool)
    with np.errstate(invalid="ignore"):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                band |= np.abs(pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] - g) > thr
    return band.ravel()


def photon_ring_band(geo, W, H):
    """The stated exclusion band of the per-pixel parity checks: fate-boundary pixels (photon
    ring, shadow and disk edges) plus the steep inner rim of the disk."""
    return fate_boundary_band(geo["outcome"], W, H) | steep_redshift_band(geo, W, H)
