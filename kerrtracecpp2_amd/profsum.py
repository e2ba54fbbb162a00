"""Summarise a rocprofv3 results database (kernel trace) as text for profiles/.

    python -m kerrtracecpp2_amd.profsum <results.db> "<command that produced it>"
"""
import sqlite3
import sys


def summary(db, command):
    c = sqlite3.connect(db)
    lines = [f"command: {command}", "", "kernel                                                        calls   total_us     avg_us"]
    for name, calls, tot, avg, pct in c.execute("select * from top_kernels"):
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


if __name__ == "__main__":
    sys.stdout.write(summary(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else ""))
