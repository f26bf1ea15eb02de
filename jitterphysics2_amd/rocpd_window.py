"""Per-kernel summary of the timed steps of a `rocprofv3 --kernel-trace` database written while the
native driver `jh_cli <scene> <warmup> <steps>` ran: the window starts after the solver launch that
precedes the last `steps` solver launches (warmup and autotune steps fall before it).

    python -m jitterphysics2_amd.rocpd_window <results.db> <steps>
"""
from __future__ import annotations

import sqlite3
import sys


def summary(path: str, steps: int, top: int = 16) -> str:
    db = sqlite3.connect(path)
    rows = db.execute("select name, start, end, duration, grid_x, workgroup_x, vgpr_count, accum_vgpr_count, "
                      "lds_size from kernels order by start").fetchall()
    solves = [r for r in rows if r[0].startswith("jh::k_solve")]
    if len(solves) <= steps:
        raise SystemExit(f"{len(solves)} solver launches, need more than {steps}")
    cut = solves[-steps - 1][2]
    timed = [r for r in rows if r[1] > cut]
    agg: dict = {}
    for r in timed:
        a = agg.setdefault(r[0].split("(")[0], [0, 0.0, r[4] // max(r[5], 1), r[5], r[6], r[7], r[8]])
        a[0] += 1
        a[1] += r[3]
    busy = sum(a[1] for a in agg.values())
    span = timed[-1][2] - timed[0][1]
    out = [f"timed window: {len(timed)} dispatches, GPU busy {busy / 1e6:.2f} ms of {span / 1e6:.2f} ms wall "
           f"({steps} steps)",
           f"{'kernel':34s} {'calls':>6s} {'avg_us':>10s} {'per_step_ms':>12s} {'share':>7s} {'wgs':>6s} {'wg':>5s} "
           f"{'vgpr':>5s} {'agpr':>5s} {'lds':>6s}"]
    for n, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"{n:34s} {a[0]:6d} {a[1] / a[0] / 1e3:10.2f} {a[1] / steps / 1e6:12.3f} {100 * a[1] / busy:6.1f}% "
                   f"{a[2]:6d} {a[3]:5d} {a[4]:5d} {a[5]:5d} {a[6]:6d}")
    return "\n".join(out)


if __name__ == "__main__":
    print(summary(sys.argv[1], int(sys.argv[2])))
