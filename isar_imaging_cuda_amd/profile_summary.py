"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) as a markdown table.

  python -m isar_imaging_cuda_amd.profile_summary <results.db> "<command that produced it>" [steps]
  python -m isar_imaging_cuda_amd.profile_summary --counters <results.db>   (rocprofv3 --pmc runs)

Prints per-kernel launches, total and average duration, share of GPU time and, when `steps` is
given, the time per bench step.
"""
from __future__ import annotations

import sqlite3
import sys


def summarise(db: str, command: str, steps: int | None = None) -> str:
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration), "
        "max(grid_x), max(grid_y), max(workgroup_x), max(lds_size), max(vgpr_count) "
        "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = [f"command: `{command}`", "", "durations in microseconds (rocprofv3 kernel trace)", "",
           "| kernel | launches | total us | avg us | min us | max us | share | grid | block | LDS B | VGPR |",
           "|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, n, tot, avg, mn, mx, gx, gy, wx, lds, vg in rows:
        short = name.replace("isar::", "").replace("(isar::Job)", "").replace("(isar::Job, int, int)", "")
        out.append(f"| {short} | {n} | {tot / 1e3:.1f} | {avg / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | "
                   f"{100 * tot / total:.1f}% | {gx}x{gy} | {wx} | {lds} | {vg} |")
    if steps:
        dev = sum(r[2] for r in rows if "rocclr" not in r[0])
        out += ["", f"device time per step (excluding runtime copy kernels): {dev / 1e3 / steps:.1f} us"]
    return "\n".join(out)


def counters(db: str) -> dict[str, dict[str, float]]:
    """Mean value per dispatch of every collected counter, by kernel, from a rocprofv3 --pmc
    database: {kernel: {"dispatches": n, "duration_us": mean, counter: mean, ...}}."""
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, dispatch_id, sum(value), max(duration) "
                       "from counters_collection group by kernel_name, counter_name, dispatch_id").fetchall()
    acc: dict[str, dict[str, list[float]]] = {}
    durs: dict[str, dict[int, float]] = {}
    for name, counter, dispatch, value, duration in rows:
        short = name.replace("isar::", "").replace("(isar::Job)", "").replace("(isar::Job, int, int)", "")
        acc.setdefault(short, {}).setdefault(counter, []).append(float(value))
        durs.setdefault(short, {})[dispatch] = float(duration or 0.0)
    out = {}
    for k, cs in acc.items():
        d = {c: sum(v) / len(v) for c, v in cs.items()}
        d["dispatches"] = float(max(len(v) for v in cs.values()))
        d["duration_us"] = sum(durs[k].values()) / max(len(durs[k]), 1) / 1e3
        out[k] = d
    return out


if __name__ == "__main__":
    if sys.argv[1] == "--counters":
        import json
        print(json.dumps(counters(sys.argv[2]), indent=1))
    else:
        print(summarise(sys.argv[1], sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else None))
