"""Synthetic workloads of the batched path.

c4_frames(seed): 64 frames of the uneven batched workload -- square sides in 64*{1..16},
a ~ U(0, 0.998), Q ~ U(0, 0.4 sqrt(1 - a^2)), Lambda ~ U(0, 1e-5), theta_obs ~ U(5, 89) deg,
r_obs ~ U(20, 80) M, disk outer radius ~ U(8, 20) M, fov 30 deg, all drawn from one
numpy default_rng(seed) stream in that order per frame.
"""
import numpy as np

C4_SEED = 20240601
C4_FRAMES = 64


def c4_params(seed=C4_SEED, n=C4_FRAMES):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = 64 * int(rng.integers(1, 17))
        a = float(rng.uniform(0.0, 0.998))
        Q = float(rng.uniform(0.0, 0.4 * np.sqrt(1.0 - a * a)))
        Lam = float(rng.uniform(0.0, 1e-5))
        theta = float(rng.uniform(5.0, 89.0))
        r_obs = float(rng.uniform(20.0, 80.0))
        disk_out = float(rng.uniform(8.0, 20.0))
        out.append(dict(width=side, height=side, a=a, Q=Q, Lambda=Lam, theta=theta, r_obs=r_obs, fov=30.0,
                        disk_out=disk_out))
    return out


def c4_frames(seed=C4_SEED, n=C4_FRAMES):
    from . import make_frame
    return [make_frame(**p) for p in c4_params(seed, n)]
