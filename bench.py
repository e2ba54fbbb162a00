"""Benchmark: rays/s of the KNdS geodesic hot path on one MI355X.

Workload: Kerr-Newman-de Sitter a=0.9, Q=0.3, Lambda=1e-5,
r_obs=40 M, theta_obs=85 deg, fov 30 deg (CLI default), disk r_isco..12 M, 1280x720,
ingoing Kerr-Schild chart, DOPRI5 rtol=1e-9 atol=1e-12.  A "step" is one full render of
the frame (device setup + persistent trace + shading) with its parameters already on the
device and the outputs (GeoPixel + RGBA) left in HBM.  `value` is the mixed-precision
rays/s; the FP64 path is timed the same way and reported beside it.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--only mixed|fp64]
    python bench.py --copy-calibration     # known-size device copy, to scale the PMC byte counters

cpu_baseline is the reference's own CPU renderer (oracle/_ref/kerr_tracer_ref, built from the
reference sources by oracle/Makefile) on the whole frame; cpu_port is the oracle
restatement on strided rows of the frame itself (also the source of accuracy_vs_oracle).

roofline.traffic (HBM bytes per trace launch) comes from the committed rocprofv3 --pmc summary
profiles/kt_trace_kernel_c2_1280x720_traffic.txt (FETCH_SIZE + WRITE_SIZE of the trace kernel,
each scaled by the same counter's reading on the calibration copy); null when that file is absent.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/sec @1280x720 a=0.9 KS DOPRI5: mixed vs FP64 GPU at max|Δg|≤1e-4"
W, H = 1280, 720
FRAME = dict(a=0.9, Q=0.3, Lambda=1e-5, theta=85.0, r_obs=40.0, fov=30.0, disk_out=12.0)
RTOL, ATOL = 1e-9, 1e-12

# Algorithmic floating-point work per unit (add/sub/mul 1, fma 2, div/sqrt/rcp 1):
#   right-hand side, counted by tests/hostcheck/flopcount.cpp on the kernel source,
#   DOPRI5 stage combination + embedded error + update per attempt,
#   integrated stage sin/cos per attempt (u = cos theta', v = -sin theta' for stages 1-5 and
#   their tableau combinations for stages 2-6),
#   one sin/cos rotation per attempt (the FSAL stage; |d| <= 1/8 series),
#   Hermite 8-bin/8-bisection disk search.
F_RHS = {"bl": 83, "ks": 78}
F_DP5_COMB = 256
F_SCI = 80
F_ROT = 31
F_HERMITE = 370
PEAK_VALU_TFLOPS = 78.6   # MI355X vector FP64; FP32 without packed math issues at the same rate

# Adjoint reverse pass (KS chart), per accepted step of one sweep:
#   segment replay (the forward pass's own attempt, dp5_attempt_h, without its error sum): five
#   h-scaled right-hand sides (F_RHS + 1) and the FSAL one, integrated stage sin/cos F_SCI, the
#   stage and 5th-order combinations (134), the FSAL rotation F_ROT;
#   the step adjoint (dp5_step_adjoint): the stage points again (first right-hand side and its
#   h-scaling 82; per stage st = 1..5 the combinations 12 st and an h-scaled right-hand side),
#   then per stage st = 5..0 a reverse-mode vector-Jacobian product (F_VJP, counted by
#   tests/hostcheck/flopcount.cpp) with its combinations (26 + 10 (5 - st)), and the sin/cos
#   adjoint folded into theta's (4).
F_VJP = {"bl": 367, "ks": 348}
F_REPLAY_STEP = 5 * (F_RHS["ks"] + 1) + F_RHS["ks"] + F_SCI + 134 + F_ROT
F_STEP_ADJ = (F_RHS["ks"] + 4) + sum(12 * st + F_RHS["ks"] + 1 for st in range(1, 6)) \
    + sum(F_VJP["ks"] + 26 + 10 * (5 - st) for st in range(6)) + 4
F_ADJ_STEP = F_REPLAY_STEP + F_STEP_ADJ
# The reverse pass runs the checkpoint segments in parallel (kt_adjoint_segments_kernel): every
# step of a disk pixel is swept four times, once per basis direction of its segment's end adjoint.
F_SEG_STEP = 4 * F_ADJ_STEP
# Jacobi bundles (kt_bundle_kernel, BL, RK4 doubling, root-skip quirk off), per attempt: the full
# step (3 plain right-hand sides F_RHS["bl"] + integrated stage sin/cos + sums: 3 x (F_RHS + 6 + 3)
# + 54), the two half steps (3 stages each of one dual-number right-hand side F_DUAL_BL (counted by
# flopcount.cpp) + 4 for its sin/cos tangents + integrated sin/cos 6 + state/W updates 24 + stage
# sums 26, then 14 for the increments), the derivative at the midpoint (F_DUAL_BL + 4 + rotation 31
# + 12) and the error norm (22); an accepted attempt adds the derivative at the new point
# (F_DUAL_BL + 4 + 31 + 21).  Integrated stage sin/cos: s' = s0 + c h (c_prev k_theta),
# c' = c0 - c h (s_prev k_theta), 3 flops each.
F_DUAL_BL = 392
F_SCI_STAGE = 6
F_BUNDLE_ATTEMPT = 3 * (F_RHS["bl"] + F_SCI_STAGE + 3) + 54 \
    + 2 * (3 * (F_DUAL_BL + 4 + F_SCI_STAGE + 24 + 26) + 14) + (F_DUAL_BL + 4 + F_ROT + 12) + 22
F_BUNDLE_ACCEPT = F_DUAL_BL + 4 + F_ROT + 21


def flops_of(st, chart="ks"):
    attempts = st["steps_accepted"] + st["steps_rejected"]
    return ((st["rhs_fp64"] + st["rhs_fp32"]) * F_RHS[chart] + attempts * (F_DP5_COMB + F_SCI + F_ROT)
            + st["events"] * F_HERMITE)


def gpu_run(ctx, kt, prec, steps, warmup):
    fr = kt.make_frame(W, H, **FRAME)
    o = c2_opts(kt, prec)
    for _ in range(warmup):
        ctx.render_device([fr], o)
    t0 = time.perf_counter()
    stats = [ctx.render_device([fr], o) for _ in range(steps)]
    wall = time.perf_counter() - t0
    out = ctx.render([fr], o, want_rgba=False, want_geo=True)
    return wall, stats, out["geo"], out["info"][0]


def c2_opts(kt, prec):
    o = kt.default_opts()
    o.chart, o.integrator, o.precision = kt.KT_CHART_KS, kt.KT_INTEGRATOR_DOPRI5, prec
    o.rtol, o.atol = RTOL, ATOL
    return o


def autotune_c2(ctx, kt, prec):
    """The project's autotuner on the benchmark frame (its own probe: the whole frame): launch
    configuration, wave-priority steps and, for mixed precision, the island threshold and radius
    under its accuracy guard.  The timed renders then use the result (their fields are 0)."""
    t = ctx.autotune([kt.make_frame(W, H, **FRAME)], c2_opts(kt, prec))
    t["gain_vs_builtin"] = t["default_ms"] / t["best_ms"] if t["best_ms"] > 0 else None
    return t


def c4_run(kt, steps):
    """The uneven batched workload: 64 seeded frames (24.2 M rays), KS DOPRI5 mixed, rendered by
    one graph launch of an autotuned plan; timed with the outputs left on the device.  A context of
    its own, so that the plan tunes on these frames (a context keeps one result per option class)."""
    from kerrtracecpp2_amd import workloads
    ctx = kt.Context(0)
    frames = workloads.c4_frames()
    npix = sum(f.width * f.height for f in frames)
    o = kt.default_opts()
    o.chart, o.integrator, o.precision = kt.KT_CHART_KS, kt.KT_INTEGRATOR_DOPRI5, kt.KT_PRECISION_MIXED
    o.rtol, o.atol = RTOL, ATOL
    plan = ctx.plan(o, len(frames), npix)
    plan.render_device(frames)          # autotune + capture + first launch
    t0 = time.perf_counter()
    sts = [plan.render_device(frames) for _ in range(steps)]
    wall = (time.perf_counter() - t0) / steps
    tune = plan.tuning()
    plan.close()
    ctx.close()
    return {"workload": f"{len(frames)} frames, sides 64..1024, seed {workloads.C4_SEED}, KS DOPRI5 mixed, one graph launch",
            "rays": npix, "ms_per_batch": 1e3 * wall, "graph_ms": float(np.mean([s["trace_ms"] for s in sts])),
            "rays_per_s": npix / wall, "steps_accepted": sts[-1]["steps_accepted"], "autotune": tune}


def c5_run(ctx, kt, reps):
    """Differentiable render: 256x256 Kerr-Newman a=0.7 Q=0.2 KS theta=60 DOPRI5; one adjoint call =
    one forward pass (step sizes, checkpoints every 16 steps, at the record stride the context keeps
    from its earlier calls) + reverse pass.  The untimed first call sizes that stride.  The rays
    still running after 128 steps finish on a second stream during the reverse pass, so
    backward_ms holds their time too."""
    fr = kt.make_frame(256, 256, a=0.7, Q=0.2, Lambda=0.0, theta=60.0, r_obs=40.0, fov=30.0, disk_out=12.0)
    o = kt.default_opts()
    o.chart, o.integrator, o.rtol, o.atol = kt.KT_CHART_KS, kt.KT_INTEGRATOR_DOPRI5, RTOL, ATOL
    ctx.render_adjoint(fr, o)
    t0 = time.perf_counter()
    outs = [ctx.render_adjoint(fr, o) for _ in range(reps)]
    wall = (time.perf_counter() - t0) / reps
    st = outs[-1]["stats"]
    bw_ms = float(np.mean([o["stats"]["backward_ms"] for o in outs]))
    fl = st["steps_differentiated"] * F_SEG_STEP
    ach = fl / (bw_ms * 1e-3) / 1e12
    return {"workload": "256x256 KN a=0.7 Q=0.2 KS theta=60 DOPRI5, dI/d(a,Q,Lambda,theta,r_obs) per pixel",
            "ms_per_call": 1e3 * wall, "forward_ms": st["forward_ms"], "record_ms": st["record_ms"],
            "backward_ms": bw_ms, "rays_differentiated": st["rays_differentiated"],
            "steps_recorded": st["steps"], "checkpoints": st["checkpoints"], "record_bytes": st["record_bytes"],
            "steps_differentiated": st["steps_differentiated"], "reverse_entries": st["reverse_entries"],
            "rays_parked": st["rays_parked"],
            "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_VALU_TFLOPS,
                         "kernel": "reverse pass (kt_adjoint_segments_kernel<KS> + table + composition)",
                         "flops_per_launch": fl, "flops_per_step": F_SEG_STEP}}


def c3_run(ctx, kt, reps):
    """Ray-bundle mode: 1920x1080 Kerr a=0.99, BL, theta_obs=89 deg, RK4 step doubling, binary64,
    Jacobi bundle (4x2 deviation matrix) per pixel; rays/s counts pixels.  Run with the bundle
    root-skip quirk off: with it (ray_bundle.hpp:357-365) every step within 0.35 rad of the
    equator that finds no disk root skips the horizon and escape tests, and at theta_obs = 89 deg
    about 23 % of the rays then run to the 500 000-step cap (hours of reference CPU time)."""
    W3, H3 = 1920, 1080
    fr = kt.make_frame(W3, H3, a=0.99, Q=0.0, Lambda=0.0, theta=89.0, r_obs=40.0, fov=30.0, disk_out=12.0)
    o = kt.default_opts()
    o.chart, o.integrator, o.bundles = kt.KT_CHART_BL, kt.KT_INTEGRATOR_RK4, 1
    o.quirks &= ~kt.KT_QUIRK_BUNDLE_ROOT_SKIP
    ctx.render_device([fr], o)
    t0 = time.perf_counter()
    sts = [ctx.render_device([fr], o) for _ in range(reps)]
    wall = (time.perf_counter() - t0) / reps
    tr = float(np.mean([s["trace_ms"] for s in sts]))
    last = sts[-1]
    fl = ((last["steps_accepted"] + last["steps_rejected"]) * F_BUNDLE_ATTEMPT
          + last["steps_accepted"] * F_BUNDLE_ACCEPT + last["events"] * F_HERMITE)
    ach = fl / (tr * 1e-3) / 1e12
    return {"workload": "1920x1080 Kerr a=0.99 BL theta=89 RK4 doubling tol 1e-7, Jacobi bundles, FP64, "
                        "bundle root-skip quirk off",
            "pixels": W3 * H3, "ms_per_frame": 1e3 * wall, "trace_ms": tr,
            "pixels_per_s": W3 * H3 / wall, "steps_accepted": last["steps_accepted"],
            "steps_rejected": last["steps_rejected"],
            "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_VALU_TFLOPS, "kernel": "kt_bundle_kernel<single,1>",
                         "flops_per_launch": fl, "flops_per_attempt": F_BUNDLE_ATTEMPT}}


def accuracy(geo, ref, px, py, band_geo):
    """Per-pixel agreement with the oracle outside the photon-ring band (fate boundaries, the
    steep inner rim |g - g_neighbour| > 0.05, and the images of the disk's edges), the band taken
    from the full FP64 GPU image."""
    from tests._util import photon_ring_band
    band = photon_ring_band(band_geo, W, H).reshape(H, W)[py, px]
    sub = geo.reshape(H, W)[py, px]
    keep = ~band
    d = keep & (sub["outcome"] == 1) & (ref["outcome"] == 1)
    dg = np.abs(sub["redshift"][d] - ref["redshift"][d])
    dr = np.abs(sub["r"][d] - ref["r"][d])
    return {"pixels_compared": int(px.size), "fate_agreement_off_band": float((sub["outcome"] == ref["outcome"])[keep].mean()),
            "max_abs_dg_off_band": float(dg.max()) if d.any() else 0.0,
            "max_abs_dr_off_band": float(dr.max()) if d.any() else 0.0,
            "fate_agreement_all": float((sub["outcome"] == ref["outcome"]).mean())}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "kt_trace_kernel_c2_1280x720_traffic.txt")
COPY_BYTES = 1 << 28


def traffic_from_profile(kernel_key):
    """HBM bytes per launch of the trace kernel recorded in the committed PMC summary."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch") and kernel_key in line:
                return float(line.split()[-1])
    except OSError:
        pass
    return None


def copy_calibration():
    """A device-to-device copy of COPY_BYTES bytes (reads and writes each byte once)."""
    import torch
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "launches": 3}))


REF_BIN = os.path.join(ROOT, "oracle", "_ref", "kerr_tracer_ref")
REF_SAMPLE = (W, H)


def cpu_reference(threads):
    """The reference's own CPU renderer (main.cpp built from its sources by oracle/Makefile, IEEE
    flags, OpenMP over pixels) on the whole benchmark frame (~14 s on 16 host cores); its "Trace:"
    time.  Run as a child process before this process touches the GPU.  None when the binary is
    not there."""
    if not os.access(REF_BIN, os.X_OK):
        return None
    w, h = REF_SAMPLE
    f = FRAME
    cmd = [REF_BIN, "--custom-res", str(w), str(h), "--a", repr(f["a"]), "--charge", repr(f["Q"]),
           "--lambda", repr(f["Lambda"]), "--theta", repr(f["theta"]), "--r-obs", repr(f["r_obs"]),
           "--fov", repr(f["fov"]), "--disk-out", repr(f["disk_out"]), "--dopri5", "--ks", "--tol", repr(RTOL),
           "--geo-only"]
    try:
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=os.path.dirname(REF_BIN),
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
    except (OSError, subprocess.TimeoutExpired):
        return None
    m = re.search(r"Trace: ([0-9.]+)s", p.stdout)
    if p.returncode != 0 or not m:
        return None
    el = float(m.group(1))
    return {"value": w * h / el, "unit": "rays/s", "cores": threads, "kind": "reference",
            "sample": f"{w * h} rays: the reference CPU renderer (main.cpp, DOPRI5 tol 1e-9, its KS-to-BL "
                      f"fallback for Lambda != 0) on the whole {w}x{h} frame, {el:.1f} s trace time"}


def cpu_baseline(budget_s):
    """The oracle (FP64 CPU restatement, OpenMP over pixels) on this host's cores, on whole rows
    of the same frame taken in a strided order until the time budget is used."""
    from oracle import kto
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fo = kto.make_frame(W, H, chart="ks", integrator="dopri5", tol=RTOL, **FRAME)
    order = np.concatenate([np.arange(s, H, 16) for s in range(16)])
    done_x, done_y = [], []
    t0 = time.perf_counter()
    outs = []
    for k in range(0, H, 8):
        rows = order[k:k + 8]
        ys, xs = np.meshgrid(rows, np.arange(W), indexing="ij")
        px, py = xs.ravel().astype(np.int32), ys.ravel().astype(np.int32)
        outs.append(kto.trace_pixels(fo, px, py, n_threads=threads))
        done_x.append(px)
        done_y.append(py)
        if time.perf_counter() - t0 >= budget_s:
            break
    el = time.perf_counter() - t0
    px, py = np.concatenate(done_x), np.concatenate(done_y)
    return {"value": px.size / el, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{px.size} rays = {px.size // W} strided rows of the 1280x720 frame, oracle FP64 "
                      f"(reference DOPRI5 semantics, tol 1e-9, BL fallback as the reference runs it), {el:.1f} s"}, \
        np.concatenate(outs), px, py


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--only", choices=["both", "mixed", "fp64", "c3", "c5"], default="both")
    ap.add_argument("--copy-calibration", action="store_true")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported by this build "
              "(rays and frames shard across GPUs without exchange; see the sharding notes)", file=sys.stderr)
        sys.exit(2)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ref_cpu = cpu_reference(threads) if (not a.no_cpu and a.only == "both") else None
    import kerrtracecpp2_amd as kt
    ctx = kt.Context(0)
    if a.only in ("c3", "c5"):   # profiling runs of the bundle / adjoint kernels, no JSON contract
        r = c3_run(ctx, kt, a.steps) if a.only == "c3" else c5_run(ctx, kt, a.steps)
        print(json.dumps({"only": a.only, **r}))
        ctx.close()
        return
    if a.only != "both":   # profiling runs: one kernel variant, no JSON contract
        prec = kt.KT_PRECISION_MIXED if a.only == "mixed" else kt.KT_PRECISION_FP64
        tuned = autotune_c2(ctx, kt, prec)      # as the full run does; its probe renders come first
        wall, st, _, _ = gpu_run(ctx, kt, prec, a.steps, a.warmup)
        print(json.dumps({"only": a.only, "ms_per_step": 1e3 * wall / a.steps,
                          "trace_ms": float(np.mean([s["trace_ms"] for s in st])), "autotune": tuned}))
        ctx.close()
        return
    tune = {"mixed": autotune_c2(ctx, kt, kt.KT_PRECISION_MIXED), "fp64": autotune_c2(ctx, kt, kt.KT_PRECISION_FP64)}
    wall_m, st_m, geo_m, info = gpu_run(ctx, kt, kt.KT_PRECISION_MIXED, a.steps, a.warmup)
    wall_d, st_d, geo_d, _ = gpu_run(ctx, kt, kt.KT_PRECISION_FP64, a.steps, a.warmup)
    ms_m = 1e3 * wall_m / a.steps
    ms_d = 1e3 * wall_d / a.steps
    rays = W * H
    trace_ms_m = float(np.mean([s["trace_ms"] for s in st_m]))
    trace_ms_d = float(np.mean([s["trace_ms"] for s in st_d]))
    fl_m = flops_of(st_m[-1])
    fl_d = flops_of(st_d[-1])
    achieved = fl_m / (trace_ms_m * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": rays / (ms_m * 1e-3),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_m,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (deterministic camera pixel grid; no input data)",
        "config": {"workload": "KNdS a=0.9 Q=0.3 Lambda=1e-5 1280x720 r_obs=40 theta=85 fov=30 disk<=12, "
                               "KS chart, DOPRI5 rtol=1e-9 atol=1e-12, mixed precision",
                   "rays_per_step": rays, "chart": "ks", "integrator": "dopri5"},
        "fp64": {"rays_per_s": rays / (ms_d * 1e-3), "ms_per_step": ms_d, "trace_kernel_ms": trace_ms_d,
                 "steps_accepted": st_d[-1]["steps_accepted"], "steps_rejected": st_d[-1]["steps_rejected"],
                 "achieved_tflops": fl_d / (trace_ms_d * 1e-3) / 1e12, "traffic": traffic_from_profile("FP64")},
        "mixed_vs_fp64": ms_d / ms_m,
        "autotune_c2": tune,
        "mixed": {"trace_kernel_ms": trace_ms_m, "steps_accepted": st_m[-1]["steps_accepted"],
                  "steps_rejected": st_m[-1]["steps_rejected"], "rhs_fp32": st_m[-1]["rhs_fp32"],
                  "rhs_fp64": st_m[-1]["rhs_fp64"]},
        "frame": {"r_horizon": info["r_horizon"], "r_isco": info["r_isco"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_VALU_TFLOPS, "traffic": traffic_from_profile("MIXED"),
                     "kernel": "kt_trace_kernel<KS,DOPRI5,MIXED>", "flops_per_launch": fl_m},
    }
    if not a.no_c4:
        result["batched_c4"] = c4_run(kt, 3)
        result["adjoint_c5"] = c5_run(ctx, kt, 3)
        result["bundles_c3"] = c3_run(ctx, kt, 2)
    if not a.no_cpu:
        cpu, ref, px, py = cpu_baseline(a.cpu_budget)
        result["cpu_baseline"] = ref_cpu if ref_cpu is not None else cpu
        result["cpu_port"] = cpu
        result["accuracy_vs_oracle"] = {"mixed": accuracy(geo_m, ref, px, py, geo_d),
                                        "fp64": accuracy(geo_d, ref, px, py, geo_d)}
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
