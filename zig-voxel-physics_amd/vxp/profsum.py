"""Summarise rocprofv3 rocpd databases (kernel trace and PMC runs) as text.

  python3 -m vxp.profsum trace  <results.db> [top]
  python3 -m vxp.profsum pmc    <results.db> <COUNTER> [name-filter]
  python3 vxp/profsum.py copycal      (the calibration program: five 256 MiB
                                       device copies, profiled like the bench)
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
        if "rocprim" in short:
            short = "rocprim::" + short.split("wrapped_")[-1].split("<")[0] if "wrapped_" in short \
                else "rocprim (" + short.split("::")[-1][:40] + ")"
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


def copycal():
    import torch
    n = 64 * 1024 * 1024  # floats: 256 MiB
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied", n * 4, "bytes x5", flush=True)


if __name__ == "__main__":
    if sys.argv[1] == "copycal":
        copycal()
    elif sys.argv[1] == "trace":
        print(trace(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 15))
    else:
        print(pmc(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else ""))
