"""Write the golden fixtures of tests/golden/ from the reference built from source.

Run in the build container only (needs /root/reference):
    make -C oracle ref && python tests/golden/make_golden.py
    python tests/golden/make_golden.py --images NAME...     (only those image fixtures)
    python tests/golden/make_golden.py --c3-sample          (oracle sample of the 1920x1080 bundle frame)

The reference CPU binary (oracle/_ref/kerr_tracer_ref, IEEE flags) produces
.kgeo files and PNGs for small configurations, and oracle/_ref/ref_harness
(the reference headers compiled where they lie) produces function-level
known answers.  Only inputs and outputs are stored: no reference source.
"""
import glob
import json
import os
import subprocess
import sys

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import kto  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "kerr_tracer_ref")
HARN = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUTD = os.path.join(ROOT, "oracle", "_ref", "out")
BG_SEED = 1234


def checker(seed=BG_SEED, w=512, h=256, cx=32, cy=16):
    """Same generator as kt_make_checker_background (splitmix64 palette)."""
    M64 = (1 << 64) - 1
    s = seed
    pal = np.zeros((cy, cx, 3), np.uint8)
    for j in range(cy):
        for i in range(cx):
            s = (s + 0x9E3779B97F4A7C15) & M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            v = z ^ (z >> 31)
            dark = ((i + j) & 1) != 0
            lo, span = (16, 64) if dark else (128, 128)
            for ch in range(3):
                pal[j, i, ch] = lo + ((v >> (8 * ch)) & 0xFF) * span // 256
    jj = (np.arange(h) * cy // h)
    ii = (np.arange(w) * cx // w)
    return pal[jj][:, ii]


def run_ref(args, geo=True):
    os.makedirs(OUTD, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count()))
    if geo:
        path = os.path.join(OUTD, "g.kgeo")
        subprocess.check_call([REF, "--geo-only", "--geo-file", path] + args, env=env,
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        return kto.read_kgeo(path)
    before = set(glob.glob(os.path.join(OUTD, "*.png")))
    subprocess.check_call([REF] + args, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    new = sorted(set(glob.glob(os.path.join(OUTD, "*.png"))) - before, key=os.path.getmtime)
    img = np.asarray(Image.open(new[-1]).convert("RGB"))
    for f in new:
        os.remove(f)
    return img


FRAMES = {
    # name: (reference CLI args, kto.make_frame kwargs)
    "c1_bl_rk4_320x180": (
        "--bl --custom-res 320 180 --a 0.5 --charge 0 --lambda 0 --r-obs 40 --theta 80 --phi 0 --fov 45 --disk-out 12",
        dict(width=320, height=180, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12, chart="bl")),
    "bl_dopri5_160x90": (
        "--bl --custom-res 160 90 --a 0.5 --r-obs 40 --theta 80 --fov 45 --disk-out 12 --dopri5 --tol 1e-8",
        dict(width=160, height=90, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12, chart="bl",
             integrator="dopri5", tol=1e-8)),
    "c2_ks_dopri5_lambda_160x90": (
        "--ks --custom-res 160 90 --a 0.9 --charge 0.3 --lambda 1e-5 --r-obs 40 --theta 85 --disk-out 12 "
        "--dopri5 --tol 1e-9",
        dict(width=160, height=90, a=0.9, Q=0.3, Lambda=1e-5, theta=85, r_obs=40, disk_out=12, chart="ks",
             integrator="dopri5", tol=1e-9)),
    "ks_rk4_spin_pos_160x90": (
        "--ks --custom-res 160 90 --a 0.5 --theta 80 --phi 0 --r-obs 40 --fov 45 --disk-out 12",
        dict(width=160, height=90, a=0.5, theta=80, r_obs=40, fov=45, disk_out=12, chart="ks")),
    "ks_rk4_spin_neg_160x90": (
        "--ks --custom-res 160 90 --a -0.5 --theta 80 --phi 0 --r-obs 40 --fov 45 --disk-out 12",
        dict(width=160, height=90, a=-0.5, theta=80, r_obs=40, fov=45, disk_out=12, chart="ks")),
    "kn_ks_64x64": (
        "--ks --custom-res 64 64 --a 0.7 --charge 0.2 --lambda 0 --theta 60 --dopri5",
        dict(width=64, height=64, a=0.7, Q=0.2, theta=60, chart="ks", integrator="dopri5")),
    "bundles_bl_64x36": (
        "--bl --bundles --custom-res 64 36 --a 0.99 --theta 89 --r-obs 40 --disk-out 12",
        dict(width=64, height=36, a=0.99, theta=89, r_obs=40, disk_out=12, chart="bl", bundles=True)),
}

IMAGES = {
    # full renders (Phase 1 + 2) with the checkerboard background
    "c1_rgb_320x180": "--bl --custom-res 320 180 --a 0.5 --r-obs 40 --theta 80 --phi 0 --fov 45 --disk-out 12",
    "c2_rgb_160x90": "--ks --custom-res 160 90 --a 0.9 --charge 0.3 --lambda 1e-5 --r-obs 40 --theta 85 "
                     "--disk-out 12 --dopri5 --tol 1e-9",
    "spp3_rgb_80x45": "--bl --custom-res 80 45 --a 0.5 --r-obs 40 --theta 80 --fov 45 --disk-out 12 --camera-spp 3",
    # Interstellar palette: the c1 frame with default cells, and a KS frame with every cell option set
    "c1_interstellar_rgb_320x180": "--bl --custom-res 320 180 --a 0.5 --r-obs 40 --theta 80 --phi 0 --fov 45 "
                                   "--disk-out 12 --disk-interstellar",
    "ks_interstellar_rgb_160x90": "--ks --custom-res 160 90 --a 0.5 --theta 80 --phi 0 --r-obs 40 --fov 45 "
                                  "--disk-out 12 --disk-interstellar --disk-rings 5 --disk-sectors 9 "
                                  "--disk-sigma 0.8 --disk-hue-offset 0.3",
    # Jacobi ray bundles: the disk flux is weighted by clamp(1/magnif, 0.05, 5) (main.cpp:1799-1814)
    "bundles_rgb_64x36": "--bl --bundles --custom-res 64 36 --a 0.99 --theta 89 --r-obs 40 --disk-out 12",
}


def harness(lines):
    p = subprocess.run([HARN], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    out = []
    for ln in p.stdout.strip().split("\n"):
        parts = ln.split()
        out.append([float(v) for v in parts[1:]])
    return out


def function_vectors():
    rng = np.random.default_rng(20261203)
    metrics = [(1.0, 0.5, 0.0, 0.0), (1.0, 0.9, 0.3, 1e-5), (1.0, -0.7, 0.2, 0.0), (1.0, 0.998, 0.0, 0.0),
               (1.0, 0.0, 0.0, 0.0), (1.0, 0.5, 0.1, 1e-4), (1.0, 0.3, 0.0, -1e-4), (1.0, 0.0, 0.4, 0.0)]
    cases = []
    for mp in metrics:
        m = " ".join(repr(v) for v in mp)
        for _ in range(6):
            r = float(rng.uniform(2.5, 60.0)); th = float(rng.uniform(0.05, 3.09))
            cases.append(("upper", f"upper {m} {r!r} {th!r}", dict(mp=mp, r=r, th=th)))
            cases.append(("lower", f"lower {m} {r!r} {th!r}", dict(mp=mp, r=r, th=th)))
            s6 = [r, th, float(rng.uniform(-2, 2)), float(rng.uniform(-6, 6)), -1.0, float(rng.uniform(-7, 7))]
            cases.append(("ham", f"ham {m} " + " ".join(repr(v) for v in s6), dict(mp=mp, s6=s6)))
            cases.append(("rhs", f"rhs {m} " + " ".join(repr(v) for v in s6), dict(mp=mp, s6=s6)))
            rr = float(rng.uniform(3.0, 20.0))
            cases.append(("omega", f"omega {m} {rr!r}", dict(mp=mp, r=rr)))
        cases.append(("horizon", f"horizon {m}", dict(mp=mp)))
        cases.append(("isco", f"isco {m}", dict(mp=mp)))
        for _ in range(4):
            cam = [float(rng.uniform(20, 80)), float(rng.uniform(5, 175)), float(rng.uniform(-30, 30)),
                   float(rng.uniform(20, 60)), 320.0, 180.0]
            al, be = float(rng.uniform(-0.3, 0.3)), float(rng.uniform(-0.2, 0.2))
            cases.append(("camera", f"camera {m} " + " ".join(repr(v) for v in cam) + f" {al!r} {be!r}",
                          dict(mp=mp, cam=cam, al=al, be=be)))
        if mp[3] == 0.0:
            for _ in range(3):
                X, Y, Z = (float(v) for v in rng.uniform(-15, 15, 3))
                cases.append(("upper_ks", f"upper_ks {m} {X!r} {Y!r} {Z!r}", dict(mp=mp, X=X, Y=Y, Z=Z)))
    # one adaptive step per method from camera rays
    for mp in metrics[:4]:
        m = " ".join(repr(v) for v in mp)
        for method in (0, 1):
            for tol in (1e-7, 1e-9):
                cam = [40.0, 80.0, 0.0, 45.0, 640.0, 360.0]
                al, be = float(rng.uniform(-0.1, 0.1)), float(rng.uniform(-0.1, 0.1))
                s = harness([f"camera {m} " + " ".join(repr(v) for v in cam) + f" {al!r} {be!r}"])[0]
                h = float(rng.choice([0.5, 1.0, 4.0]))
                line = f"step {m} {method} " + " ".join(repr(v) for v in s) + f" {h!r} {tol!r}"
                cases.append(("step", line, dict(mp=mp, method=method, state=s, h=h, tol=tol)))
    for _ in range(12):
        a = float(rng.uniform(-0.99, 0.99))
        r, th, ph = float(rng.uniform(2, 40)), float(rng.uniform(0.1, 3.0)), float(rng.uniform(-3, 3))
        cases.append(("bl2ks", f"bl2ks 1 {a!r} 0 0 {r!r} {th!r} {ph!r}", dict(a=a, r=r, th=th, ph=ph)))
        X, Y, Z = (float(v) for v in rng.uniform(-20, 20, 3))
        cases.append(("ks2bl", f"ks2bl 1 {a!r} 0 0 {X!r} {Y!r} {Z!r}", dict(a=a, X=X, Y=Y, Z=Z)))
    for _ in range(24):
        y0, y1 = float(rng.uniform(1.0, 2.1)), float(rng.uniform(1.0, 2.1))
        d0, d1 = float(rng.uniform(-0.3, 0.3)), float(rng.uniform(-0.3, 0.3))
        h = float(rng.uniform(0.1, 5.0))
        cases.append(("hermite_root", f"hermite_root 1 0 0 0 {y0!r} {y1!r} {d0!r} {d1!r} {h!r} {np.pi/2!r}",
                      dict(y0=y0, y1=y1, d0=d0, d1=d1, h=h, target=np.pi / 2)))
    outs = harness([c[1] for c in cases])
    return [dict(kind=c[0], inputs=c[2], output=o) for c, o in zip(cases, outs)]


def images_only(names):
    """Re-make only the named image fixtures (the rest of the manifest is kept)."""
    bg_path = os.path.join(OUTD, "checker_bg.png")
    os.makedirs(OUTD, exist_ok=True)
    Image.fromarray(checker()).save(bg_path)
    with open(os.path.join(HERE, "manifest.json")) as f:
        manifest = json.load(f)
    for name in names:
        args = IMAGES[name]
        img = run_ref(args.split() + ["--bg", bg_path], geo=False)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), rgb=img)
        manifest["images"][name] = dict(args=args)
        print(name, img.shape, flush=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


def c3_bundle_sample(n=16384, seed=3):
    """Configuration 3 at full size (1920x1080, a=0.99, theta=89, BL, RK4 doubling, Jacobi
    bundles): n seeded pixels traced by the oracle with the bundle root-skip quirk cleared
    (KT_QUIRK_BUNDLE_ROOT_SKIP, ray_bundle.hpp:357-365), the semantics the GPU runs this frame
    with; every other reference behaviour is kept.  The oracle's bundle path is bit-exact to the
    reference binary on the 64x36 bundle frame with every quirk set."""
    W, H = 1920, 1080
    quirks = kto.QUIRKS_REFERENCE & ~kto.QUIRK_BUNDLE_ROOT_SKIP
    fr = kto.make_frame(W, H, a=0.99, theta=89, r_obs=40, disk_out=12, chart="bl", integrator="rk4", bundles=True,
                        quirks=quirks)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(W * H, n, replace=False))
    px = (idx % W).astype(np.int32)
    py = (idx // W).astype(np.int32)
    g = kto.trace_pixels(fr, px, py)
    np.savez_compressed(os.path.join(HERE, "c3_bundles_sample.npz"), px=px, py=py, pixels=g, W=W, H=H,
                        quirks=quirks)
    print("c3 bundle sample", n, np.bincount(g["outcome"], minlength=4))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--c3-sample":
        return c3_bundle_sample()
    if not (os.path.exists(REF) and os.path.exists(HARN)):
        sys.exit("build the reference first: make -C oracle ref")
    if len(sys.argv) > 2 and sys.argv[1] == "--images":
        return images_only(sys.argv[2:])
    bg = checker()
    bg_path = os.path.join(OUTD, "checker_bg.png")
    os.makedirs(OUTD, exist_ok=True)
    Image.fromarray(bg).save(bg_path)
    manifest = {"background": dict(seed=BG_SEED, width=512, height=256, cells_x=32, cells_y=16), "frames": {},
                "images": {}}
    for name, (args, kw) in FRAMES.items():
        g = run_ref(args.split())
        np.savez_compressed(os.path.join(HERE, name + ".npz"), pixels=g["pixels"], meta=g["meta"],
                            W=g["W"], H=g["H"])
        manifest["frames"][name] = dict(args=args, frame=kw)
        print(name, g["W"], g["H"], flush=True)
    for name, args in IMAGES.items():
        img = run_ref(args.split() + ["--bg", bg_path], geo=False)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), rgb=img)
        manifest["images"][name] = dict(args=args)
        print(name, img.shape, flush=True)
    fv = function_vectors()
    with open(os.path.join(HERE, "function_vectors.json"), "w") as f:
        json.dump(fv, f)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("function vectors", len(fv))


if __name__ == "__main__":
    main()
