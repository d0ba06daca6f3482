"""Summarise rocprofv3 rocpd databases (kernel trace and PMC runs) as text.

  python3 -m vxp.profsum trace  <results.db> [top]
  python3 -m vxp.profsum pmc    <results.db> <COUNTER> [name-filter]
  python3 -m vxp.profsum last   <results.db> <kernel-substring> <N> [COUNTER]
                               (the last N dispatches of one kernel: average
                                duration, or average counter value per dispatch)
  python3 -m vxp.profsum section <results.db> <marker-kernel> <skip> [top]
                               (kernels dispatched from the (skip+1)-th dispatch
                                of the marker kernel on: one section of a run)
  python3 vxp/profsum.py copycal      (the calibration program: five 256 MiB
                                       device copies, profiled like the bench)
  python3 vxp/profsum.py scene <c1|c3> <frames>
                               (a program to profile: the 4096-body scene or the
                                512-world batch for the first `frames` frames,
                                printing the per-frame means of the pair,
                                manifold, candidate and checked-voxel counts)
"""
import sqlite3
import sys


def trace(db, top=15):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage "
                       "from top_kernels order by total_duration desc limit ?", (top,)).fetchall()
    out = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for name, n, tot, avg, pct in rows:
        short = name.split("(")[0].replace("void ", "")
        out.append(f"| {short[:70]} | {n} | {tot / 1e3:.3f} | {avg:.1f} | {pct:.1f} |")
    return "\n".join(out)


def pmc(db, counter, filt=""):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, count(*), avg(value) from counters_collection "
                       "where counter_name=? group by kernel_name order by sum(value) desc",
                       (counter,)).fetchall()
    out = [f"| kernel | dispatches | avg {counter} (KiB) |", "|---|---|---|"]
    for name, n, avg in rows:
        if filt and filt not in name:
            continue
        out.append(f"| {name.split('(')[0].replace('void ', '')[:70]} | {n} | {avg:.1f} |")
    return "\n".join(out)


def section(db, marker, skip, top=30):
    """Per-kernel totals over the dispatches from the (skip+1)-th dispatch of the
    kernel whose name contains `marker` on (trace run), with the count of marker
    dispatches in the section (steps, for k_step_begin)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, duration from kernels order by start").fetchall()
    seen, t0 = 0, None
    for name, start, _ in rows:
        if marker in name:
            seen += 1
            if seen == skip + 1:
                t0 = start
                break
    if t0 is None:
        return f"fewer than {skip + 1} dispatches of {marker}"
    agg = {}
    steps = 0
    for name, start, dur in rows:
        if start < t0:
            continue
        short = name.split("(")[0].replace("void ", "")
        a = agg.setdefault(short, [0, 0.0])
        a[0] += 1
        a[1] += dur
        if marker in name:
            steps += 1
    total = sum(v[1] for v in agg.values())
    out = [f"{steps} dispatches of {marker} in the section", "",
           "| kernel | calls | total ms | avg us | per step us | % |", "|---|---|---|---|---|---|"]
    for k, (n, d) in sorted(agg.items(), key=lambda x: -x[1][1])[:top]:
        out.append(f"| {k[:70]} | {n} | {d / 1e6:.3f} | {d / n / 1e3:.1f} | "
                   f"{d / max(steps, 1) / 1e3:.1f} | {100 * d / total:.1f} |")
    return "\n".join(out)


def last(db, sub, n, counter=None):
    """Average over the last `n` dispatches (by dispatch id) of the kernels whose
    name contains `sub`: duration in us (trace run) or `counter` (PMC run)."""
    con = sqlite3.connect(db)
    if counter is None:
        rows = con.execute("select dispatch_id, duration from kernels where name like ? "
                           "order by dispatch_id desc limit ?", (f"%{sub}%", n)).fetchall()
        vals = [r[1] / 1e3 for r in rows]
        what = "duration (us)"
    else:
        rows = con.execute("select dispatch_id, sum(value) from counters_collection "
                           "where kernel_name like ? and counter_name = ? group by dispatch_id "
                           "order by dispatch_id desc limit ?", (f"%{sub}%", counter, n)).fetchall()
        vals = [r[1] for r in rows]
        what = f"{counter} (KiB)"
    if not vals:
        return f"no dispatches of {sub}"
    ids = sorted(r[0] for r in rows)
    return (f"{sub}: last {len(vals)} dispatches (ids {ids[0]}..{ids[-1]}): average {what} = "
            f"{sum(vals) / len(vals):.1f}, min {min(vals):.1f}, max {max(vals):.1f}")


def copycal():
    import torch
    n = 64 * 1024 * 1024  # floats: 256 MiB
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied", n * 4, "bytes x5", flush=True)


def scene(name, frames):
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import vxp
    from vxp import scenes
    sc = {"c1": scenes.c1_large, "c3": scenes.batched_worlds}[name]()
    s = vxp.from_scene(sc, lib=vxp.product_lib())
    keys = ("broad_phase_pairs", "narrow_phase_tested", "manifolds", "contact_candidates",
            "voxels_checked")
    acc = dict.fromkeys(keys, 0)
    for _ in range(frames):
        s.update(sc.dt)
        for st in s.stats_all():
            for k in keys:
                acc[k] += getattr(st, k)
    print(f"{sc.name}: frames 0..{frames - 1}, per-frame means:",
          {k: round(v / frames, 1) for k, v in acc.items()}, flush=True)
    s.close()


if __name__ == "__main__":
    if sys.argv[1] == "copycal":
        copycal()
    elif sys.argv[1] == "scene":
        scene(sys.argv[2], int(sys.argv[3]))
    elif sys.argv[1] == "last":
        print(last(sys.argv[2], sys.argv[3], int(sys.argv[4]),
                   sys.argv[5] if len(sys.argv) > 5 else None))
    elif sys.argv[1] == "section":
        print(section(sys.argv[2], sys.argv[3], int(sys.argv[4]),
                      int(sys.argv[5]) if len(sys.argv) > 5 else 30))
    elif sys.argv[1] == "trace":
        print(trace(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 15))
    else:
        print(pmc(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else ""))
