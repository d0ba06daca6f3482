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
"cpu_baseline" times the CPU restatement on the 4096-body scene.  Prints one
JSON line.

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
    """Step `steps` frames after `warmup`; per-frame wall time of update().
    `profile`: the library's timing mode over those frames (0 off, 1 every
    phase, 2 only the solve units: two HIP events per step)."""
    lib = sys_.lib
    cfg = sc.config
    for _ in range(warmup):
        sys_.update(sc.dt)
    if profile:
        lib.profile_enable(sys_.ptr, profile)
    frame_ms, body_substeps, alg, cands = [], 0, 0, 0
    for _ in range(steps):
        t = time.perf_counter()
        sys_.update(sc.dt)  # synchronous: returns after the frame's kernels finish
        frame_ms.append((time.perf_counter() - t) * 1e3)
        for st in sys_.stats_all():
            body_substeps += st.awake_bodies_stepped * cfg["substeps"]
            cands += st.contact_candidates
            alg += solver_bytes(st, cfg["substeps"], cfg["velocity_iterations"])
    kern = {}
    if profile:
        import vxp
        for i, name in enumerate(vxp._abi.KERNEL_NAMES):
            ms, n = C.c_double(), C.c_uint64()
            lib.profile_get(sys_.ptr, i, C.byref(ms), C.byref(n))
            kern[name] = (ms.value, n.value)
        lib.profile_enable(sys_.ptr, 0)
    wall = sum(frame_ms) / 1e3
    return dict(value=body_substeps / wall, wall=wall, frame_ms=frame_ms, kern=kern,
                alg_solver_bytes=alg, body_substeps=body_substeps, candidates=cands,
                tuning=sys_.tuning())


def pct(xs, q):
    s = sorted(xs)
    return s[min(len(s) - 1, int(q * len(s)))]


def run_serving(lib, args):
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
    srv.close()
    return {
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
    ap.add_argument("--steps", type=int, default=590)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-breakdown", action="store_true",
                    help="skip the second, untimed pass of the scene with every phase timed (kernel_ms)")
    ap.add_argument("--batched-steps", type=int, default=290)
    ap.add_argument("--looped-worlds", type=int, default=48)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--c2-steps", type=int, default=390)
    ap.add_argument("--serve-worlds", type=int, default=2048)
    ap.add_argument("--serve-slots", type=int, default=256)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-rates", default="2000,5000,10000,20000")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(worlds shard across GPUs without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import vxp
    from vxp import scenes

    lib = vxp.product_lib()
    # ---- plain single scene: 4096 bodies ----------------------------------------
    log("4096-body scene")
    sc = scenes.c1_large()
    s1 = vxp.from_scene(sc, lib=lib)
    # the timed frames time only the solve units (the roofline's kernel), with
    # HIP events on their stream; timing every phase would add two events per
    # phase and keep the island labels off their side stream
    plain = run(s1, sc, args.warmup, args.steps, profile=vxp._abi.PROFILE_SOLVE_UNITS)
    final = s1.stats()
    s1.close()
    brk = None
    if not args.no_breakdown:  # the same frames again on a fresh copy, every phase timed
        s1 = vxp.from_scene(sc, lib=lib)
        brk = run(s1, sc, args.warmup, args.steps, profile=vxp._abi.PROFILE_ALL)
        s1.close()
    kern = plain["kern"]
    bk = brk["kern"] if brk else {}
    dom = max(bk, key=lambda k: bk[k][0]) if bk else "solve_units"
    solve_ms, solve_n = kern["solve_units"]
    avg_solve_s = solve_ms / max(solve_n, 1) / 1e3
    alg_per_launch = plain["alg_solver_bytes"] / max(solve_n, 1)
    achieved = alg_per_launch / avg_solve_s / 1e9 if solve_n else 0.0
    traffic = None
    tf = os.path.join(REPO, "profiles", "k_solve_units_c1_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")

    # ---- batched: 512 worlds ------------------------------------------------------
    batched = None
    if not args.no_batched:
        log("512 batched worlds")
        sb = scenes.batched_worlds()
        sysb = vxp.from_scene(sb, lib=lib)
        rb = run(sysb, sb, args.warmup, args.batched_steps)
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
            "plain_looped_value": loop_bs / loop_t if loop_t else None,
            "plain_looped_sample": f"{nl} of the {len(sb.worlds)} worlds stepped one system at a "
                                   f"time, frames {args.warmup}..{args.warmup + lf - 1}",
        }

    # ---- narrow-phase-heavy scene: 256 rocks -------------------------------------
    rocks = None
    if args.c2_steps:
        log("256-rock scene")
        sr = scenes.c2_rocks()
        s2 = vxp.from_scene(sr, lib=lib)
        for _ in range(args.warmup):
            s2.update(sr.dt)
        vchk = cont = bsub = cand = 0
        t_tot = 0.0
        for _ in range(args.c2_steps):
            t = time.perf_counter()
            s2.update(sr.dt)
            t_tot += time.perf_counter() - t
            st = s2.stats()
            vchk += st["voxels_checked"]
            cont += st["contacts_generated"]
            cand += st["contact_candidates"]
            bsub += st["awake_bodies_stepped"] * sr.config["substeps"]
        s2.close()
        rocks = {"workload": f"{sr.name}: 256 rocks of 32^3 voxels (0.1 m), dt 1/120, "
                             f"{sr.config['substeps']} substeps, {sr.config['velocity_iterations']} "
                             f"iterations, frames {args.warmup}..{args.warmup + args.c2_steps - 1}",
                 "voxel_pair_tests_per_s": vchk / t_tot, "contacts_per_s": cont / t_tot,
                 "contact_candidates_per_s": cand / t_tot,
                 "body_substeps_per_s": bsub / t_tot, "ms_per_step": t_tot * 1e3 / args.c2_steps}

    # ---- serving: Poisson requests over 2048 resident worlds -------------------
    serving = None
    if not args.no_serving:
        log("serving")
        serving = run_serving(lib, args)

    # ---- CPU baseline: the CPU restatement on the same scene, bounded sample ----
    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        threads = min(16, os.cpu_count() or 1)
        ora = vxp.from_scene(sc, lib=vxp.oracle_lib(), cfg_over=dict(cpu_threads=threads))
        gpu2 = vxp.from_scene(sc, lib=lib)
        cb, gb, frames, tc, tg = 0, 0, 0, 0.0, 0.0
        while tc < args.cpu_seconds and frames < args.warmup + args.steps:
            t = time.perf_counter()
            ora.update(sc.dt)
            tc += time.perf_counter() - t
            t = time.perf_counter()
            gpu2.update(sc.dt)
            tg += time.perf_counter() - t
            cb += ora.stats()["awake_bodies_stepped"] * sc.config["substeps"]
            gb += gpu2.stats()["awake_bodies_stepped"] * sc.config["substeps"]
            frames += 1
        cpu = {"value": cb / tc, "unit": "body-substeps/s", "cores": threads, "kind": "port",
               "sample": f"CPU restatement (oracle/, {threads} narrow-phase threads) on frames "
                         f"0..{frames - 1} of the 4096-body scene ({tc:.1f} s of CPU time); "
                         f"the GPU on the same frames: {gb / tg:.0f} body-substeps/s"}
        ora.close()
        gpu2.close()

    out = {
        "metric": "body-substeps/s (4096-body scene; 512 batched worlds); p50/p99 step latency",
        "value": plain["value"],
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
                               f"{args.warmup}..{args.warmup + args.steps - 1}",
                   "worlds": 1},
        "roofline": {"bound": "hbm", "kernel": "k_solve_units", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "avg_launch_ms": avg_solve_s * 1e3, "launches": solve_n,
                     "alg_bytes_per_launch": alg_per_launch, "dominant_kernel": dom},
        "kernel_ms": {k: round(v[0], 3) for k, v in bk.items()},
        "kernel_ms_note": ("the same frames on a second copy of the scene with every phase "
                           "timed (HIP events per phase; not the timed run)") if bk else None,
        "tuning": plain["tuning"],
        "final_stats": {k: final[k] for k in ("manifolds", "colors", "sleeping_bodies",
                                              "voxels_checked", "broad_phase_pairs")},
        "batched": batched,
        "rocks": rocks,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
