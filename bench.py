#!/usr/bin/env python3
"""Benchmark of the MI355X step pipeline.

One "step" is one update(dt) frame of a world (4 solver substeps).  The metric
is body-substeps/s: awake dynamic bodies at the start of every executed step x
substeps, divided by the wall time of the update() calls, with the worlds
already resident in HBM.  `value` is the 4096-body scene on one GPU.  Beside
it: "batched" steps 512 independent worlds of uneven size together and
compares them with the plain single-world path looped over the same worlds;
"rocks" reports voxel-pair tests/s and contacts/s on the 256-rock scene;
"serving" drives the continuous-batching loop over 2048 resident worlds with
Poisson request arrivals and reports p50/p99 request latency per offered rate;
"c0" times the 64-cube scene; "cpu_baseline" times the CPU restatement on the
last frames of the 4096-body window, and each section carries the CPU
restatement on its own workload ("cpu").  "recorded" lines repeat the 4096-body
window, the batch and the serving loop with the steps replayed from recorded
HIP graphs (vxp_graph_enable; same results, fewer host round trips).  Prints
one JSON line.

The 4096-body scene's trajectory is 600 frames; the timed frames are its last
--steps frames (frames 600 - steps .. 599), after the earlier ones run untimed,
so a short run times the settled pile, not the start of the drop.

  python bench.py [--gpus 1] [--steps 590] [--warmup 10] [--batched-steps 290]
                  [--c2-steps 390] [--serve-rates 2000,5000,10000,20000]
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "zig-voxel-physics_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def solver_bytes(st, substeps, iters):
    """Algorithmic HBM bytes of the solve-units kernel (k_solve_units) for one
    world-step: the substeps, the warm start, the first position pass and the
    impulse store of ContactSolver.solve,
    S*[N*208 + K*120 + M*184 + (I+1)*(K*112 + M*176)] + (K*72 + M*176)
    + min(J, 1)*(K*40 + M*192) + K*20, with N awake bodies, K contact points and
    M manifolds in the solve, S substeps, I velocity iterations, J executed
    position iterations (FP32 SoA, each datum counted once per mandated pass).
    The rest of the solver's bytes, (J - 1)*(K*40 + M*192) + N*144, belong to
    k_solve_pos and k_solve_final.  A frame of several executed steps launches
    the kernel once per step: each counts with the frame's mean N and the last
    step's K, M and J (the stats keep the last step's), so the bytes stay
    matched to the launch count."""
    n_steps = st.steps_executed
    if not n_steps:
        return 0
    N = st.awake_bodies_stepped // n_steps
    K, M = st.solver_points, st.solver_manifolds
    S, I, J = substeps, iters, st.position_iterations_run
    if M == 0:
        return n_steps * S * N * 208
    return n_steps * (S * (N * 208 + K * 120 + M * 184 + (I + 1) * (K * 112 + M * 176))
                      + (K * 72 + M * 176) + min(J, 1) * (K * 40 + M * 192) + K * 20)


def run(sys_, sc, warmup, steps, profile=0):
    """Step `steps` frames after `warmup` untimed ones; per-frame wall time of
    update() and per-frame counts.  `profile`: the library's timing mode over
    the timed frames (0 off, 1 every phase, 2 only the solve units: two HIP
    events per step)."""
    lib = sys_.lib
    cfg = sc.config
    for _ in range(warmup):
        sys_.update(sc.dt)
    if profile:
        lib.profile_enable(sys_.ptr, profile)
    frame_ms, frame_bs, alg, cands = [], [], 0, 0
    for _ in range(steps):
        t = time.perf_counter()
        sys_.update(sc.dt)  # synchronous: returns after the frame's kernels finish
        frame_ms.append((time.perf_counter() - t) * 1e3)
        bs = 0
        for st in sys_.stats_all():
            bs += st.awake_bodies_stepped * cfg["substeps"]
            cands += st.contact_candidates
            alg += solver_bytes(st, cfg["substeps"], cfg["velocity_iterations"])
        frame_bs.append(bs)
    kern = {}
    if profile:
        import vxp
        for i, name in enumerate(vxp._abi.KERNEL_NAMES):
            ms, n = C.c_double(), C.c_uint64()
            lib.profile_get(sys_.ptr, i, C.byref(ms), C.byref(n))
            kern[name] = (ms.value, n.value)
        lib.profile_enable(sys_.ptr, 0)
    wall = sum(frame_ms) / 1e3
    return dict(value=sum(frame_bs) / wall, wall=wall, frame_ms=frame_ms, frame_bs=frame_bs,
                kern=kern, alg_solver_bytes=alg, body_substeps=sum(frame_bs), candidates=cands,
                tuning=sys_.tuning())


def cpu_run(sys_, sc, pre, frames, seconds, extra=()):
    """The CPU restatement: `pre` untimed frames, then up to `frames` timed ones
    (fewer if `seconds` of CPU time run out).  Returns per-frame times and
    body-substeps plus the sums of the `extra` stats."""
    cfg = sc.config
    t0 = time.perf_counter()
    for _ in range(pre):
        sys_.update(sc.dt)
    pre_s = time.perf_counter() - t0
    ms, bs, sums = [], [], {k: 0 for k in extra}
    while len(ms) < frames and sum(ms) / 1e3 < seconds:
        t = time.perf_counter()
        sys_.update(sc.dt)
        ms.append((time.perf_counter() - t) * 1e3)
        st = sys_.stats_all()
        bs.append(sum(x.awake_bodies_stepped for x in st) * cfg["substeps"])
        for k in extra:
            sums[k] += sum(getattr(x, k) for x in st)
    return dict(frame_ms=ms, frame_bs=bs, wall=sum(ms) / 1e3, sums=sums, pre_s=pre_s)


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def pct(xs, q):
    s = sorted(xs)
    return s[min(len(s) - 1, int(q * len(s)))]


def graph_line(r, sys_, steps):
    """Timing of a run whose steps replay recorded graphs, with its graph counters."""
    return {"value": r["value"], "ms_per_step": r["wall"] * 1e3 / steps,
            "p50_step_ms": statistics.median(r["frame_ms"]), "p99_step_ms": pct(r["frame_ms"], 0.99),
            "graph": sys_.graphInfo()}


def run_serving(lib, args, recorded=False):
    """Requests "advance world W by k frames" arrive as a Poisson process at each
    offered rate for --serve-seconds (W uniform over the resident worlds, k
    uniform in {1, 2, 4, 8}); the loop submits every request whose arrival time
    has passed, then ticks.  Latency is completion minus arrival; throughput is
    the body-substeps of the requests completed within the window per second."""
    import random
    from vxp import scenes, serve
    t0 = time.perf_counter()
    sc = scenes.serving_pool(n_worlds=args.serve_worlds)
    t_gen = time.perf_counter() - t0
    srv = serve.from_scene(sc, slots=args.serve_slots, budget_bytes=200 * 10**9, lib=lib)
    srv.set_recorded(recorded)
    t_build = time.perf_counter() - t0 - t_gen
    nw = len(sc.worlds)
    rng = random.Random(46)
    # warm-up: one request per slot, drained
    for i in range(args.serve_slots):
        srv.submit(rng.randrange(nw), 1, 1 << 62 | i)
    while srv.info()["completed"] < args.serve_slots:
        srv.tick()
    srv.poll()
    rates, tag = [], 0
    for lam in [float(x) for x in args.serve_rates.split(",") if x]:
        arrivals, t = [], 0.0
        while True:
            t += rng.expovariate(lam)
            if t >= args.serve_seconds:
                break
            arrivals.append((t, rng.randrange(nw), rng.choice((1, 2, 4, 8))))
        start = time.monotonic()
        i, ticks = 0, 0
        while True:
            now = time.monotonic()
            if now - start >= args.serve_seconds:
                break
            while i < len(arrivals) and start + arrivals[i][0] <= now:
                a, w, k = arrivals[i]
                srv.submit(w, k, tag, start + a)
                tag += 1
                i += 1
            if srv.info()["occupied"] or srv.info()["queued"]:
                srv.tick()
                ticks += 1
            elif i < len(arrivals):
                time.sleep(max(0.0, start + arrivals[i][0] - time.monotonic()))
        wall = time.monotonic() - start
        done = srv.poll()
        done = done[done["tag"] < (1 << 62)]
        lat = (done["done_s"] - done["submit_s"]) * 1e3
        info1 = srv.info()
        log(f"serving: {lam:g} req/s offered, {len(done)} completed")
        rates.append({
            "offered_rps": lam, "submitted": i, "completed": int(len(done)),
            "served_rps": len(done) / wall,
            "body_substeps_per_s": float(done["body_substeps"].sum()) / wall,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "ticks": ticks, "ms_per_tick": wall * 1e3 / max(ticks, 1),
            "queued_at_end": info1["queued"], "occupied_at_end": info1["occupied"]})
        # drain what is left before the next rate (not timed)
        while srv.info()["occupied"] or srv.info()["queued"]:
            srv.tick()
        srv.poll()
    info = srv.info()
    ginfo = srv.graph_info()
    srv.close()
    return {
        "steps_recorded": recorded, "graph": ginfo,
        "workload": f"{sc.name}: {nw} resident worlds, {sc.n_dynamic} dynamic bodies (log-uniform "
                    f"8..2048 per world), {args.serve_slots} batch slots, Poisson arrivals for "
                    f"{args.serve_seconds:g} s per rate, k uniform in {{1,2,4,8}}",
        "unit": "body-substeps/s", "rates": rates,
        "resident_bytes": info["resident_bytes"], "budget_bytes": info["budget_bytes"],
        "setup_s": {"generate": round(t_gen, 1), "load": round(t_build, 1)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=590,
                    help="timed frames: the last STEPS frames of the 600-frame 4096-body trajectory")
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed frames before the timed ones (at least 600 - STEPS)")
    ap.add_argument("--no-breakdown", action="store_true",
                    help="skip the second, untimed pass of the scene with every phase timed (kernel_ms)")
    ap.add_argument("--batched-steps", type=int, default=290)
    ap.add_argument("--looped-worlds", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU time of each timed CPU sample")
    ap.add_argument("--cpu-frames", type=int, default=30,
                    help="frames of the 4096-body CPU sample (the last ones of the timed window)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-c0", action="store_true")
    ap.add_argument("--no-recorded", action="store_true",
                    help="skip the recorded-step (HIP graph) runs beside the eager ones")
    ap.add_argument("--c2-steps", type=int, default=390)
    ap.add_argument("--serve-worlds", type=int, default=2048)
    ap.add_argument("--serve-slots", type=int, default=256)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-rates", default="2000,5000,10000,20000")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(worlds are independent and would shard across GPUs without communication, "
              "one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import vxp
    from vxp import scenes

    lib = vxp.product_lib()
    threads = min(16, os.cpu_count() or 1)
    cpu_model, nproc = host_cpu()
    # ---- plain single scene: 4096 bodies, the last `steps` of its 600 frames ----
    sc = scenes.c1_large()
    pre = max(args.warmup, sc.frames - args.steps)
    first, last = pre, pre + args.steps - 1
    log(f"4096-body scene: frames {first}..{last} timed after {pre} untimed")
    s1 = vxp.from_scene(sc, lib=lib)
    # the timed frames time only the solve units (the roofline's kernel), with
    # HIP events on their stream; timing every phase would add two events per
    # phase and keep the island labels off their side stream
    plain = run(s1, sc, pre, args.steps, profile=vxp._abi.PROFILE_SOLVE_UNITS)
    final = s1.stats()
    s1.close()
    brk = None
    if not args.no_breakdown:  # the same frames again on a fresh copy, every phase timed
        s1 = vxp.from_scene(sc, lib=lib)
        brk = run(s1, sc, pre, args.steps, profile=vxp._abi.PROFILE_ALL)
        s1.close()
    rec1 = None
    if not args.no_recorded:  # the same frames replayed from recorded graphs (no kernel timing)
        s1 = vxp.from_scene(sc, lib=lib)
        s1.setRecorded(True)
        r = run(s1, sc, pre, args.steps)
        rec1 = graph_line(r, s1, args.steps)
        s1.close()
    kern = plain["kern"]
    bk = brk["kern"] if brk else {}
    dom = max(bk, key=lambda k: bk[k][0]) if bk else "solve_units"
    solve_ms, solve_n = kern["solve_units"]
    avg_solve_s = solve_ms / max(solve_n, 1) / 1e3
    alg_per_launch = plain["alg_solver_bytes"] / max(solve_n, 1)
    achieved = alg_per_launch / avg_solve_s / 1e9 if solve_n else 0.0
    # HBM bytes per launch from the PMC runs of the same frame window (null when
    # no run covers this window)
    traffic = None
    tf = os.path.join(REPO, "profiles", "k_solve_units_c1_traffic.json")
    if os.path.exists(tf):
        for wnd in json.load(open(tf)).get("windows", []):
            if wnd["frames"] == f"{first}..{last}":
                traffic = wnd["hbm_bytes_per_launch"]

    # ---- the 64-cube scene (the small configuration), frames warmup..299 -------
    c0 = None
    if not args.no_c0:
        log("64-cube scene")
        s0c = scenes.c0_small()
        s0 = vxp.from_scene(s0c, lib=lib)
        r0 = run(s0, s0c, args.warmup, s0c.frames - args.warmup)
        s0.close()
        c0 = {"workload": f"{s0c.name}: 64 cubes + ground, frames {args.warmup}..{s0c.frames - 1}",
              "value": r0["value"], "unit": "body-substeps/s",
              "ms_per_step": r0["wall"] * 1e3 / len(r0["frame_ms"]),
              "p50_step_ms": statistics.median(r0["frame_ms"]),
              "p99_step_ms": pct(r0["frame_ms"], 0.99)}
        if not args.no_recorded:  # the same frames replayed from recorded graphs
            s0 = vxp.from_scene(s0c, lib=lib)
            s0.setRecorded(True)
            r = run(s0, s0c, args.warmup, s0c.frames - args.warmup)
            c0["recorded"] = graph_line(r, s0, s0c.frames - args.warmup)
            s0.close()
        if not args.no_cpu:
            o0 = vxp.from_scene(s0c, lib=vxp.oracle_lib(), cfg_over=dict(cpu_threads=threads))
            rc0 = cpu_run(o0, s0c, args.warmup, s0c.frames - args.warmup, args.cpu_seconds)
            o0.close()
            n = len(rc0["frame_ms"])
            c0["cpu"] = {"value": sum(rc0["frame_bs"]) / rc0["wall"], "cores": threads,
                         "frames": f"{args.warmup}..{args.warmup + n - 1}",
                         "p50_step_ms": statistics.median(rc0["frame_ms"]),
                         "p99_step_ms": pct(rc0["frame_ms"], 0.99)}

    # ---- batched: 512 worlds ------------------------------------------------------
    batched = None
    if not args.no_batched:
        log("512 batched worlds")
        sb = scenes.batched_worlds()
        sysb = vxp.from_scene(sb, lib=lib)
        rb = run(sysb, sb, args.warmup, args.batched_steps)
        sysb.close()
        recb = None
        if not args.no_recorded:
            sysb = vxp.from_scene(sb, lib=lib)
            sysb.setRecorded(True)
            r = run(sysb, sb, args.warmup, args.batched_steps)
            recb = graph_line(r, sysb, args.batched_steps)
            sysb.close()
        # the plain single-world path looped over a sample of the same worlds,
        # over the same frames
        nl = min(args.looped_worlds, len(sb.worlds))
        idx = [int(i * len(sb.worlds) / nl) for i in range(nl)]
        singles = [vxp.from_scene(sb, lib=lib, worlds=[w]) for w in idx]
        lf = min(args.batched_steps, 60)
        for s in singles:
            for _ in range(args.warmup):
                s.update(sb.dt)
        loop_t, loop_bs = 0.0, 0
        for _ in range(lf):
            for s in singles:
                t = time.perf_counter()
                s.update(sb.dt)
                loop_t += time.perf_counter() - t
                loop_bs += s.stats_all()[0].awake_bodies_stepped * sb.config["substeps"]
        for s in singles:
            s.close()
        batched = {
            "value": rb["value"], "unit": "body-substeps/s",
            "workload": f"{sb.name}: {len(sb.worlds)} worlds, {sb.n_dynamic} dynamic bodies "
                        f"(log-uniform 8..1024 per world), frames {args.warmup}.."
                        f"{args.warmup + args.batched_steps - 1}",
            "ms_per_step": rb["wall"] * 1e3 / args.batched_steps,
            "nominal_value": sb.n_dynamic * sb.config["substeps"] * args.batched_steps / rb["wall"],
            "contact_candidates_per_step": rb["candidates"] / args.batched_steps,
            "p50_step_ms": statistics.median(rb["frame_ms"]),
            "p99_step_ms": pct(rb["frame_ms"], 0.99),
            "tuning": rb["tuning"],
            "recorded": recb,
            "plain_looped_value": loop_bs / loop_t if loop_t else None,
            "plain_looped_sample": f"{nl} of the {len(sb.worlds)} worlds stepped one system at a "
                                   f"time, frames {args.warmup}..{args.warmup + lf - 1}",
        }
        if not args.no_cpu:
            # the reference has no world batching: one system per world, looped
            # over the same sample of worlds, frame by frame
            ors = [vxp.from_scene(sb, lib=vxp.oracle_lib(), worlds=[w],
                                  cfg_over=dict(cpu_threads=threads)) for w in idx]
            for o in ors:
                for _ in range(args.warmup):
                    o.update(sb.dt)
            ct, cbs, cf = 0.0, 0, 0
            while cf < lf and ct < args.cpu_seconds:
                for o in ors:
                    t = time.perf_counter()
                    o.update(sb.dt)
                    ct += time.perf_counter() - t
                    cbs += o.stats_all()[0].awake_bodies_stepped * sb.config["substeps"]
                cf += 1
            for o in ors:
                o.close()
            batched["cpu"] = {"value": cbs / ct, "cores": threads,
                              "sample": f"CPU restatement looped over the same {nl} worlds "
                                        f"(one system each), frames {args.warmup}.."
                                        f"{args.warmup + cf - 1}"}

    # ---- narrow-phase-heavy scene: 256 rocks -------------------------------------
    rocks = None
    if args.c2_steps:
        log("256-rock scene")
        sr = scenes.c2_rocks()
        s2 = vxp.from_scene(sr, lib=lib)
        for _ in range(args.warmup):
            s2.update(sr.dt)
        per = []  # per frame: (ms, voxels_checked, contacts, candidates, body-substeps)
        for _ in range(args.c2_steps):
            t = time.perf_counter()
            s2.update(sr.dt)
            ms = (time.perf_counter() - t) * 1e3
            st = s2.stats()
            per.append((ms, st["voxels_checked"], st["contacts_generated"], st["contact_candidates"],
                        st["awake_bodies_stepped"] * sr.config["substeps"]))
        s2.close()
        t_tot = sum(p[0] for p in per) / 1e3
        rocks = {"workload": f"{sr.name}: 256 rocks of 32^3 voxels (0.1 m), dt 1/120, "
                             f"{sr.config['substeps']} substeps, {sr.config['velocity_iterations']} "
                             f"iterations, frames {args.warmup}..{args.warmup + args.c2_steps - 1}",
                 "voxel_pair_tests_per_s": sum(p[1] for p in per) / t_tot,
                 "contacts_per_s": sum(p[2] for p in per) / t_tot,
                 "contact_candidates_per_s": sum(p[3] for p in per) / t_tot,
                 "body_substeps_per_s": sum(p[4] for p in per) / t_tot,
                 "ms_per_step": t_tot * 1e3 / args.c2_steps}
        if not args.no_cpu:
            o2 = vxp.from_scene(sr, lib=vxp.oracle_lib(), cfg_over=dict(cpu_threads=threads))
            rc2 = cpu_run(o2, sr, args.warmup, args.c2_steps, args.cpu_seconds,
                          extra=("voxels_checked", "contacts_generated"))
            o2.close()
            n = len(rc2["frame_ms"])
            g_t = sum(p[0] for p in per[:n]) / 1e3
            rocks["cpu"] = {
                "voxel_pair_tests_per_s": rc2["sums"]["voxels_checked"] / rc2["wall"],
                "contacts_per_s": rc2["sums"]["contacts_generated"] / rc2["wall"],
                "body_substeps_per_s": sum(rc2["frame_bs"]) / rc2["wall"], "cores": threads,
                "frames": f"{args.warmup}..{args.warmup + n - 1}",
                "gpu_on_same_frames": {
                    "voxel_pair_tests_per_s": sum(p[1] for p in per[:n]) / g_t,
                    "contacts_per_s": sum(p[2] for p in per[:n]) / g_t}}

    # ---- serving: Poisson requests over 2048 resident worlds -------------------
    serving = None
    if not args.no_serving:
        log("serving")
        serving = run_serving(lib, args)
        if not args.no_recorded:
            log("serving, recorded steps")
            serving["recorded"] = run_serving(lib, args, recorded=True)

    # ---- CPU baseline: the CPU restatement on the last frames of the timed window
    cpu = None
    if not args.no_cpu:
        k = min(args.cpu_frames, args.steps)
        c_first = last - k + 1
        log(f"CPU baseline: frames {c_first}..{last} of the 4096-body scene after {c_first} untimed")
        ora = vxp.from_scene(sc, lib=vxp.oracle_lib(), cfg_over=dict(cpu_threads=threads))
        rc = cpu_run(ora, sc, c_first, k, 1e9)
        ora.close()
        g_ms = sum(plain["frame_ms"][-k:]) / 1e3
        g_bs = sum(plain["frame_bs"][-k:])
        cpu = {"value": sum(rc["frame_bs"]) / rc["wall"], "unit": "body-substeps/s",
               "cores": threads, "kind": "port",
               "sample": f"CPU restatement (oracle/, {threads} narrow-phase threads, "
                         f"{cpu_model}, {nproc} logical CPUs) on frames {c_first}..{last} of the "
                         f"4096-body scene ({rc['wall']:.1f} s of CPU time after an untimed "
                         f"{rc['pre_s']:.0f} s run of frames 0..{c_first - 1}); the GPU on the "
                         f"same frames: {g_bs / g_ms:.0f} body-substeps/s",
               "p50_step_ms": statistics.median(rc["frame_ms"]),
               "p99_step_ms": pct(rc["frame_ms"], 0.99)}

    out = {
        "metric": "body-substeps/s (4096-body scene; 512 batched worlds); p50/p99 step latency",
        "value": plain["value"],
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "untimed_frames": pre,
        "ms_per_step": plain["wall"] * 1e3 / args.steps,
        # every dynamic body counted every frame, asleep or not
        "nominal_value": sc.n_dynamic * sc.config["substeps"] * args.steps / plain["wall"],
        "p50_step_ms": statistics.median(plain["frame_ms"]),
        "p99_step_ms": pct(plain["frame_ms"], 0.99),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, zig-voxel-physics_amd/vxp/scenes.py)",
        "config": {"workload": f"{sc.name}: {sc.n_dynamic} dynamic bodies + static ground, "
                               f"dt {sc.dt:.5f}, {sc.config['substeps']} substeps, "
                               f"{sc.config['velocity_iterations']} iterations, frames "
                               f"{first}..{last} of {sc.frames} timed (frames 0..{pre - 1} run "
                               f"untimed first)",
                   "worlds": 1},
        "roofline": {"bound": "hbm", "kernel": "k_solve_units", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "avg_launch_ms": avg_solve_s * 1e3, "launches": solve_n,
                     "alg_bytes_per_launch": alg_per_launch, "dominant_kernel": dom,
                     "frames": f"{first}..{last}"},
        "kernel_ms": {k: round(v[0], 3) for k, v in bk.items()},
        "kernel_ms_note": ("the same frames on a second copy of the scene with every phase "
                           "timed (HIP events per phase; not the timed run)") if bk else None,
        "tuning": plain["tuning"],
        "recorded": rec1,
        "final_stats": {k: final[k] for k in ("manifolds", "colors", "sleeping_bodies",
                                              "voxels_checked", "broad_phase_pairs")},
        "host_cpu": {"model": cpu_model, "logical_cpus": nproc},
        "c0": c0,
        "batched": batched,
        "rocks": rocks,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
