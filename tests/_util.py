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


