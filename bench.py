#!/usr/bin/env python3
"""Benchmark of the MI355X step pipeline on the 4096-body scene.

One "step" is one update(dt) frame of the whole world (4 solver substeps).
The metric is body-substeps/s: awake dynamic bodies at the start of every
executed step x substeps, divided by wall time, with the world already resident
in HBM.  Prints one JSON line (see the keys below).

  python bench.py [--gpus 1] [--steps 590] [--warmup 10] [--config c1]
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "zig-voxel-physics_amd"))

HBM_PEAK_GBS = 8000.0


def solver_bytes(st, cfg):
    """Algorithmic HBM bytes of one solver launch (one world-step):
    S*[N*208 + K*120 + M*184 + (I+1)*(K*112 + M*176)] + (K*72 + M*176)
    + J*(K*40 + M*192) + N*144 + K*20, with N awake bodies, K contact points and
    M manifolds in the solve, S substeps, I velocity iterations, J executed
    position iterations (FP32 SoA, each datum counted once per mandated pass)."""
    N = st["awake_bodies_stepped"] // max(st["steps_executed"], 1)
    K, M = st["solver_points"], st["solver_manifolds"]
    S, I, J = cfg["substeps"], cfg["velocity_iterations"], st["position_iterations_run"]
    if M == 0:
        return S * N * 208 + N * 144
    return (S * (N * 208 + K * 120 + M * 184 + (I + 1) * (K * 112 + M * 176)) + (K * 72 + M * 176)
            + J * (K * 40 + M * 192) + N * 144 + K * 20)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=590)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c2"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(worlds shard across GPUs without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import vxp
    from vxp import scenes

    gen = {"c0": scenes.c0_small, "c1": scenes.c1_large, "c2": scenes.c2_rocks}[args.config]
    sc = gen()
    lib = vxp.product_lib()
    sys_ = vxp.from_scene(sc, lib=lib)
    cfg = dict(sc.config)
    substeps = cfg["substeps"]

    for _ in range(args.warmup):
        sys_.update(sc.dt)
    lib.profile_enable(sys_.ptr, 1)
    frame_ms, body_substeps, alg_solver = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        sys_.update(sc.dt)  # synchronous: returns after the frame's kernels finish
        frame_ms.append((time.perf_counter() - t) * 1e3)
        st = sys_.stats()
        body_substeps += st["awake_bodies_stepped"] * substeps
        alg_solver += solver_bytes(st, {"substeps": substeps,
                                        "velocity_iterations": cfg["velocity_iterations"]}) \
            if st["steps_executed"] else 0
    wall = time.perf_counter() - t0
    kern = {}
    for i, name in enumerate(vxp._abi.KERNEL_NAMES):
        ms, n = C.c_double(), C.c_uint64()
        lib.profile_get(sys_.ptr, i, C.byref(ms), C.byref(n))
        kern[name] = (ms.value, n.value)
    lib.profile_enable(sys_.ptr, 0)
    final = sys_.stats()
    sys_.close()

    dom = max(kern, key=lambda k: kern[k][0])
    solve_ms, solve_n = kern["solve"]
    avg_solve_s = solve_ms / max(solve_n, 1) / 1e3
    achieved = (alg_solver / max(solve_n, 1)) / avg_solve_s / 1e9 if solve_n else 0.0

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        ora = vxp.from_scene(sc, lib=vxp.oracle_lib(), cfg_over=dict(cpu_threads=threads))
        gpu2 = vxp.from_scene(sc, lib=lib)
        cb, frames, tc, tg = 0, 0, 0.0, 0.0
        while tc < args.cpu_seconds and frames < args.warmup + args.steps:
            t = time.perf_counter()
            ora.update(sc.dt)
            tc += time.perf_counter() - t
            t = time.perf_counter()
            gpu2.update(sc.dt)
            tg += time.perf_counter() - t
            cb += ora.stats()["awake_bodies_stepped"] * substeps
            frames += 1
        cpu = {"value": cb / tc, "unit": "body-substeps/s", "cores": threads, "kind": "port",
               "sample": f"CPU restatement (oracle/, {threads} narrow-phase threads) on frames "
                         f"0..{frames - 1} of the same scene ({tc:.1f} s); the GPU on the same "
                         f"frames: {cb / tg:.0f} body-substeps/s"}
        ora.close()
        gpu2.close()

    out = {
        "metric": "body-substeps/s (4096-body scene); p50/p99 step latency",
        "value": body_substeps / wall,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "p50_step_ms": statistics.median(frame_ms),
        "p99_step_ms": sorted(frame_ms)[min(len(frame_ms) - 1, int(0.99 * len(frame_ms)))],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, zig-voxel-physics_amd/vxp/scenes.py)",
        "config": {"workload": f"{sc.name}: {sc.n_dynamic} dynamic bodies + static ground, "
                               f"dt {sc.dt:.5f}, {substeps} substeps, "
                               f"{cfg['velocity_iterations']} iterations, frames "
                               f"{args.warmup}..{args.warmup + args.steps - 1}",
                   "worlds": 1},
        "roofline": {"bound": "hbm", "kernel": "k_solve", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None,
                     "avg_launch_ms": avg_solve_s * 1e3, "launches": solve_n,
                     "alg_bytes_per_launch": alg_solver / max(solve_n, 1),
                     "dominant_kernel": dom},
        "kernel_ms": {k: round(v[0], 3) for k, v in kern.items()},
        "final_stats": {k: final[k] for k in ("manifolds", "colors", "sleeping_bodies",
                                              "voxels_checked", "broad_phase_pairs")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
