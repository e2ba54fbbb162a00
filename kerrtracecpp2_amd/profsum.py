"""Summarise rocprofv3 output as short text for profiles/.

    python -m kerrtracecpp2_amd.profsum [--last N] <results.db> "<command>"  kernel-trace summary
    (--last N, before any mode: each kernel's last N dispatches only)
    python -m kerrtracecpp2_amd.profsum pmc <counter_collection.csv> ...      per-kernel counter means
    python -m kerrtracecpp2_amd.profsum traffic <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv>
        HBM bytes per launch of each kt_ kernel: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
        scaled by bytes_copied / counter on a device copy of known size.
"""
import csv
import sqlite3
import sys
from collections import defaultdict


def summary(db, command, last=0):
    """Per-kernel calls / total / average duration; last > 0: only each kernel's last `last`
    dispatches (the timed renders after an autotune's probe renders)."""
    c = sqlite3.connect(db)
    lines = [f"command: {command}", "", "kernel                                                        calls   total_us     avg_us"]
    if last > 0:
        per = defaultdict(list)
        for name, t0, t1 in c.execute("select name, start, end from kernels order by start"):
            per[name].append((t1 - t0) / 1e3)
        rows = [(n, v[-last:]) for n, v in per.items()]
        grand = sum(sum(v) for _, v in rows)
        for name, v in sorted(rows, key=lambda x: -sum(x[1])):
            short = name.replace("(anonymous namespace)::", "")
            lines.append(f"{short[:60]:60s} {len(v):6d} {sum(v):11.1f} {sum(v) / len(v):10.1f}  "
                         f"({100.0 * sum(v) / grand:.1f}%)")
        rows = []
    else:
        rows = c.execute("select * from top_kernels")
    for name, calls, tot, avg, pct in rows:
        short = name.replace("(anonymous namespace)::", "")
        lines.append(f"{short[:60]:60s} {calls:6d} {tot:11.1f} {avg:10.1f}  ({pct:.1f}%)")
    lines.append("")
    lines.append("per-kernel resources (first dispatch): name, grid, workgroup, vgpr, accum_vgpr, sgpr, lds, scratch")
    seen = set()
    for row in c.execute("select name, grid_x, workgroup_x, vgpr_count, accum_vgpr_count, sgpr_count, lds_size, "
                         "scratch_size from kernels order by start"):
        if row[0] in seen:
            continue
        seen.add(row[0])
        lines.append("  " + ", ".join(str(v).replace("(anonymous namespace)::", "") for v in row))
    return "\n".join(lines) + "\n"


def _short(name):
    return name.replace("(anonymous namespace)::", "")


LAST = 0   # > 0: counter means over each kernel's last LAST dispatches only


def pmc_means(path):
    """{(kernel, counter): (mean value per dispatch, dispatches)} from a counter_collection.csv
    (dispatches in file order, which is dispatch order)."""
    acc = defaultdict(list)
    with open(path) as f:
        for row in csv.DictReader(f):
            acc[(_short(row["Kernel_Name"]), row["Counter_Name"])].append(float(row["Counter_Value"]))
    if LAST > 0:
        acc = {k: v[-LAST:] for k, v in acc.items()}
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def pmc_summary(paths):
    lines = ["kernel                                                        counter              mean/dispatch  n"]
    for p in paths:
        for (k, c), (m, n) in sorted(pmc_means(p).items()):
            lines.append(f"{k[:60]:60s} {c:20s} {m:14.1f} {n:3d}")
    return "\n".join(lines) + "\n"


def _copy_kib(path, counter):
    vals = [m for (k, c), (m, n) in pmc_means(path).items() if c == counter and "copy" in k.lower()]
    if not vals:
        vals = [m for (k, c), (m, n) in pmc_means(path).items() if c == counter]
    return max(vals)


def traffic_summary(fetch, write, cfetch, cwrite, copy_bytes=1 << 28):
    f_copy, w_copy = _copy_kib(cfetch, "FETCH_SIZE"), _copy_kib(cwrite, "WRITE_SIZE")
    sf, sw = copy_bytes / f_copy, copy_bytes / w_copy
    lines = [f"calibration: {copy_bytes} B device copy reads FETCH_SIZE {f_copy:.1f} KiB, writes WRITE_SIZE "
             f"{w_copy:.1f} KiB -> bytes per counter unit: fetch {sf:.2f}, write {sw:.2f}", ""]
    fm, wm = pmc_means(fetch), pmc_means(write)
    for (k, c), (m, n) in sorted(fm.items()):
        if c != "FETCH_SIZE" or not k.startswith(("kt_", "void kt_")):
            continue
        w = wm.get((k, "WRITE_SIZE"), (float("nan"), 0))[0]
        lines.append(f"{k}: FETCH_SIZE {m:.1f} KiB, WRITE_SIZE {w:.1f} KiB per dispatch ({n} dispatches)")
        lines.append(f"traffic_bytes_per_launch {k.replace(' ', '')} {m * sf + w * sw:.0f}")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--last":
        LAST = int(sys.argv[2])
        del sys.argv[1:3]
    if sys.argv[1] == "pmc":
        sys.stdout.write(pmc_summary(sys.argv[2:]))
    elif sys.argv[1] == "traffic":
        sys.stdout.write(traffic_summary(*sys.argv[2:6]))
    else:
        sys.stdout.write(summary(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else "", LAST))
