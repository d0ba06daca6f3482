"""Parity of the HIP pipeline (libvxp.so, called through the C ABI) with the
CPU restatement on identical seeded scenes.

The bar is exact agreement, frame after frame, of every artefact of the step:
the broad-phase pair set, every manifold field (contacts, feature ids, solver
colours, warm-start impulses and effective masses), every body's pose,
velocities and sleep state, island labels and the step statistics.  Floats are
compared with IEEE equality (NaN == NaN), i.e. exact up to the sign of zero,
which the reference itself leaves unspecified (Zig @min/@max lower to
minnum/maxnum).  Tolerances of 1e-5 relative on contacts and 1e-4 on body state
are therefore met with margin 0.
"""
import os

import numpy as np
import pytest

import vxp
from vxp import _abi, compare, scenes
from vxp.scenes import SplitMix64

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _lockstep(product, oracle, scene, frames, check_every=1, cfg_over=None, worlds=None):
    over = dict(cfg_over or {})
    print(f"lockstep {scene.name} frames={frames} over={over}", flush=True)
    f, msg, _ = compare.lockstep(scene, product, oracle, frames, cfg_over=over, worlds=worlds,
                                 check_every=check_every)
    assert f is None, f"frame {f}: {msg}"


def test_c0_every_frame(product, oracle):
    _lockstep(product, oracle, scenes.c0_small(), 300)


@pytest.mark.parametrize("over", [dict(use_brick_pair_cache=0), dict(enable_warm_starting=0, warm_starting=0),
                                  dict(enable_sleeping=0), dict(enable_coloring=1),
                                  dict(use_brick_sat=0, velocity_iterations=2, position_iterations=4),
                                  dict(brick_pair_cache_size=64), dict(enable_friction=0)])
def test_c0_config_knobs(product, oracle, over):
    _lockstep(product, oracle, scenes.c0_small(), 90, check_every=3, cfg_over=over)


def test_c1_4096_bodies(product, oracle):
    _lockstep(product, oracle, scenes.c1_large(), 120, check_every=20,
              cfg_over=dict(cpu_threads=THREADS))


def test_c2_rocks(product, oracle):
    _lockstep(product, oracle, scenes.c2_rocks(), 40, check_every=10,
              cfg_over=dict(cpu_threads=THREADS))


def test_batched_worlds_match_oracle_and_single(product, oracle):
    sc = scenes.batched_worlds(n_worlds=6, lo=8, hi=96, blob_pool=8)
    _lockstep(product, oracle, sc, 120, check_every=20, cfg_over=dict(cpu_threads=THREADS))
    # every world stepped alone equals the same world inside the batch
    batch = vxp.from_scene(sc, lib=product)
    singles = [vxp.from_scene(sc, lib=product, worlds=[w]) for w in range(len(sc.worlds))]
    for _ in range(60):
        batch.update(sc.dt)
        for s in singles:
            s.update(sc.dt)
    for w, s in enumerate(singles):
        assert compare.diff_bodies(batch.bodies(w), s.bodies(0)) is None
        assert compare.diff_manifolds(batch.manifolds(w), s.manifolds(0)) is None


def test_destroy_create_and_wake(product, oracle):
    sc = scenes.c0_small()
    a = vxp.from_scene(sc, lib=product)
    b = vxp.from_scene(sc, lib=oracle)
    for fr in range(200):
        for s in (a, b):
            s.update(sc.dt)
            if fr == 60:
                s.destroyBody(s.handles[0][5])
                s.destroyBody(s.handles[0][9])
            if fr == 61:
                s.handles[0][5] = s.createDynamicBody((0.3, 4.0, 0.2), (0, 0, 0, 1), 1, 1.0)
            if fr == 150:  # wake a sleeping body with a kick
                s.setState(s.handles[0][3], linear_velocity=(0.0, 3.0, 1.0))
        if fr % 10 == 0 or fr in (60, 61, 150, 151):
            assert compare.diff_bodies(a.bodies(), b.bodies()) is None, fr
            assert compare.diff_manifolds(a.manifolds(), b.manifolds()) is None, fr
            assert compare.diff_stats(a.stats(), b.stats()) is None, fr


def test_product_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_abi, "PRODUCT_SO", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_abi, "_cache", {})
    with pytest.raises(vxp.VxpError):
        vxp.VoxelPhysicsSystem()


def test_mass_gravity_solver_config_and_kinematic(product, oracle):
    """createDynamicBodyWithMass, setGravity, setSolverConfig and a kinematic
    body driven through the pile, in lockstep with the restatement."""
    sc = scenes.c0_small()
    a = vxp.from_scene(sc, lib=product, cfg_over=dict(max_bodies=68))
    b = vxp.from_scene(sc, lib=oracle, cfg_over=dict(max_bodies=68))
    inv_i = (3.0, 0.1, 0.0, 0.1, 2.5, 0.0, 0.0, 0.0, 4.0)
    for fr in range(160):
        for s in (a, b):
            if fr == 20:
                s.setGravity((0.0, -4.0, 1.5))
            if fr == 35:
                s.setSolverConfig(velocity_iterations=3, hertz=20.0, position_iterations=4)
            if fr == 45:
                s.extra = s.createDynamicBodyWithMass((0.4, 3.0, -0.3), (0.0, 0.0, 0.0, 1.0), 1,
                                                      0.7, inv_i)
            if fr == 60:
                s.kin = s.createKinematicBody((-4.5, 0.2, 0.0), (0.0, 0.0, 0.0, 1.0), 1)
                s.setState(s.kin, linear_velocity=(2.0, 0.0, 0.3))
            s.update(sc.dt)
        if fr % 5 == 0 or fr in (20, 35, 45, 60, 61):
            assert compare.diff_bodies(a.bodies(), b.bodies()) is None, fr
            assert compare.diff_manifolds(a.manifolds(), b.manifolds()) is None, fr
            assert compare.diff_stats(a.stats(), b.stats()) is None, fr


@pytest.mark.parametrize("order", [_abi.PAIR_ORDER_KEY, _abi.PAIR_ORDER_TREE])
def test_empty_collider_body_in_scene(product, oracle, order):
    """A body whose collider has no voxels (src/system.zig:1139-1158) falls under
    gravity; its broad-phase box is NaN (transform of the empty AABB) and must
    not disturb the pair set of the other bodies (nor, in the tree order, the
    tree's descent and queries)."""
    sc = scenes.c0_small()
    runs = []
    for lib in (product, oracle):
        s = vxp.from_scene(sc, lib=lib, cfg_over=dict(max_bodies=67, pair_order=order))
        cid = s.addCollider(voxel_size=0.1, brick_pos=np.zeros((0, 3), np.int32),
                            voxels=np.zeros((0, 512), np.uint8))
        h = s.createDynamicBody((0.2, 5.0, 0.1), (0.0, 0.0, 0.0, 1.0), cid, 1.0)
        runs.append((s, h))
    (a, ha), (b, hb) = runs
    for fr in range(90):
        a.update(sc.dt)
        b.update(sc.dt)
        if fr % 10 == 0:
            assert compare.diff_bodies(a.bodies(), b.bodies()) is None, fr
            assert np.array_equal(a.pairs(), b.pairs()), fr
            assert compare.diff_manifolds(a.manifolds(), b.manifolds()) is None, fr
    assert a.getPosition(ha)[1] < 5.0


def test_tuning_choices_are_bit_identical(product):
    # Every tuning knob changes only a launch shape: forced extremes of the
    # lane-pair colour limit (every colour on single lanes / every colour of up
    # to 1024 manifolds on lane pairs), both solver workgroup sizes, both broad
    # phases, every probe workgroup size and XCD runs of 1, 8 and 64, and the
    # online tuner going through a whole round, give the same bits.
    sc = scenes.batched_worlds(n_worlds=6, lo=8, hi=1024, blob_pool=8)
    runs = []
    shapes = ((0, 512, "sweep", 256, 64), (1024, 256, "lbvh", 64, 8), (0, 256, "sweep", 128, 1),
              (None, 0, "auto", 0, 0))
    for forced, threads, bp, pt, run in shapes:
        s = vxp.from_scene(sc, lib=product)
        if forced is not None:
            s.setTuning(forced, threads, bp, pt, run)
        print(f"tuning forced={forced} unit_threads={threads} broad={bp} probe={pt}/{run}",
              flush=True)
        for _ in range(60):
            s.update(sc.dt)
        t = s.tuning()
        assert t["fixed"] == (forced is not None)
        if forced is not None:
            assert t["pair_color_max"] == forced and t["windows"] == 0
            assert (t["unit_threads"], t["broad_phase"], t["probe_threads"],
                    t["probe_xcd_run"]) == (threads, bp, pt, run)
        else:
            # one window per knob: 15 + 6 + 6 + 9 + 9 frames
            assert t["windows"] >= 5
            assert t["unit_threads"] in (256, 512) and t["broad_phase"] in ("sweep", "lbvh")
            assert t["probe_threads"] in (64, 128, 256) and t["probe_xcd_run"] in (8, 64, 512)
        runs.append(([s.bodies(w) for w in range(len(sc.worlds))],
                     [s.manifolds(w) for w in range(len(sc.worlds))],
                     [s.stats(w) for w in range(len(sc.worlds))]))
        s.close()
    for other in runs[1:]:
        for w in range(len(sc.worlds)):
            assert compare.diff_bodies(other[0][w], runs[0][0][w]) is None, w
            assert compare.diff_manifolds(other[1][w], runs[0][1][w]) is None, w
            assert compare.diff_stats(other[2][w], runs[0][2][w]) is None, w


def test_tuning_rejects_values_outside_the_knobs(product):
    s = vxp.from_scene(scenes.c0_small(frames=1), lib=product)
    for kw in (dict(unit_threads=384), dict(probe_threads=96), dict(probe_xcd_run=48),
               dict(probe_xcd_run=8192)):
        with pytest.raises(vxp.VxpError):
            s.setTuning(None, **kw)
    s.setTuning(None, probe_threads=128, probe_xcd_run=16)
    t = s.tuning()
    assert (t["probe_threads"], t["probe_xcd_run"], t["fixed"]) == (128, 16, False)
    s.close()


def test_body_mutators_between_frames(product, oracle):
    # every reference mutator (applyForce ... sleep) applied between frames to
    # both libraries; bodies, manifolds and stats equal after every frame, and
    # getWorldAABB / findBodyById agree
    sc = scenes.c0_small()
    a = vxp.from_scene(sc, lib=product)
    b = vxp.from_scene(sc, lib=oracle)
    rng = SplitMix64(11)

    def v3(lo=-2.0, hi=2.0):
        return tuple(rng.uniform(lo, hi) for _ in range(3))

    schedule = {}
    for f in range(5, 150, 10):
        ops = []
        for _ in range(6):
            h = 1 + rng.randint(0, 63)
            op = rng.randint(0, 11)
            pos = v3(-4, 4)
            if op == _abi.OP_SET_ROTATION:
                q = np.array(rng.quat(), dtype=np.float32)
                ops.append((h, op, tuple(float(x) for x in q), None))
            elif op in (_abi.OP_APPLY_FORCE_AT_POINT, _abi.OP_APPLY_IMPULSE_AT_POINT):
                ops.append((h, op, v3(), pos))
            elif op == _abi.OP_SET_POSITION:
                ops.append((h, op, (pos[0], rng.uniform(1.0, 4.0), pos[2]), None))
            else:
                ops.append((h, op, v3(), None))
        schedule[f] = ops
    ids = a.bodies()["id"]  # slot order
    print(f"mutators: {sum(len(v) for v in schedule.values())} ops over 150 frames", flush=True)
    for f in range(150):
        for h, op, v, p in schedule.get(f, []):
            for s in (a, b):
                hh = s.findBodyById(int(ids[h]))
                assert hh is not None
                s._op(hh, op, v, p)
        a.update(sc.dt)
        b.update(sc.dt)
        msg = (compare.diff_bodies(a.bodies(), b.bodies()) or
               compare.diff_manifolds(a.manifolds(), b.manifolds()) or
               compare.diff_stats(a.stats(), b.stats()))
        assert msg is None, f"frame {f}: {msg}"
    for h in range(1, 65):
        ha = a.findBodyById(int(ids[h]))
        hb = b.findBodyById(int(ids[h]))
        assert (ha.index, ha.generation) == (hb.index, hb.generation)
        assert a.getWorldAABB(ha) == b.getWorldAABB(hb)
    a.close()
    b.close()


def test_profile_modes_change_no_result(product):
    # Timing every phase keeps the island labels on the main stream; timing
    # only the solve units (or nothing) runs them beside the solver on a second
    # stream.  Every mode gives the same bits, including island labels and
    # island statistics, and the solve-units mode times nothing else.
    import ctypes as C
    sc = scenes.c0_small(n_side=4, frames=150)
    frames = 150
    runs = []
    for mode in (_abi.PROFILE_OFF, _abi.PROFILE_ALL, _abi.PROFILE_SOLVE_UNITS):
        print(f"profile mode {mode}", flush=True)
        s = vxp.from_scene(sc, lib=product)
        assert s.lib.profile_enable(s.ptr, 7) == _abi.VXP_ERR_ARGUMENT
        assert s.lib.profile_enable(s.ptr, mode) == _abi.VXP_OK
        labels = []
        for _ in range(frames):
            s.update(sc.dt)
            labels.append(s.islands())
        launches = []
        for i in range(len(_abi.KERNEL_NAMES)):
            ms, n = C.c_double(), C.c_uint64()
            s.lib.profile_get(s.ptr, i, C.byref(ms), C.byref(n))
            launches.append(n.value)
        su = _abi.KERNEL_NAMES.index("solve_units")
        if mode == _abi.PROFILE_SOLVE_UNITS:
            assert launches[su] > 0 and sum(launches) == launches[su]
        elif mode == _abi.PROFILE_ALL:
            assert all(launches[i] > 0 for i in (su, _abi.KERNEL_NAMES.index("islands")))
        else:
            assert sum(launches) == 0
        runs.append((s.bodies(), s.manifolds(), s.stats(), s.getIslandStats(), labels))
        s.close()
    for other in runs[1:]:
        assert compare.diff_bodies(other[0], runs[0][0]) is None
        assert compare.diff_manifolds(other[1], runs[0][1]) is None
        assert compare.diff_stats(other[2], runs[0][2]) is None
        assert other[3] == runs[0][3]
        for a, b in zip(other[4], runs[0][4]):
            assert all(np.array_equal(x, y) for x, y in zip(a, b))
