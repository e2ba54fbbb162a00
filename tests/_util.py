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


def disk_edge_band(geo, W, H, rel=0.005):
    """Disk pixels on the image of the disk's inner or outer edge: the pixel or an 8-neighbour
    hits the disk within rel of r_in or r_out (taken as the image's smallest and largest hit
    radius).  Where a ray crosses the equator just outside the edge it hits on a later crossing
    instead, so r_hit jumps there between the direct image of the disk and a higher-order one
    (fate stays "disk"), and a perturbation of the ray at the 1e-5 level moves a pixel across:
    at 1280x720 a mixed-precision render put one such pixel at r = 11.99997 against 9.6746."""
    out = np.asarray(geo["outcome"]).reshape(H, W)
    r = np.asarray(geo["r"], np.float64).reshape(H, W)
    disk = out == 1
    if not disk.any():
        return np.zeros(W * H, dtype=bool)
    r_in, r_out = r[disk].min(), r[disk].max()
    near = disk & ((np.abs(r - r_out) <= rel * r_out) | (np.abs(r - r_in) <= rel * r_in))
    pad = np.pad(near, 1, constant_values=False)
    band = np.zeros((H, W), dtype=bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            band |= pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
    return (band & disk).ravel()


def photon_ring_band(geo, W, H):
    """The stated exclusion band of the per-pixel parity checks: fate-boundary pixels (photon
    ring, shadow and disk edges), the steep inner rim of the disk, and the images of the disk's
    edges where the hit moves between crossings."""
    return fate_boundary_band(geo["outcome"], W, H) | steep_redshift_band(geo, W, H) | disk_edge_band(geo, W, H)
