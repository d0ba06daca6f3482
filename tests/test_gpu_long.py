"""Parity of the HIP pipeline with the CPU restatement at the sizes the bench
runs, and on the paths the short lockstep tests leave out.

- the 4096-body scene through all 600 frames and the 256-rock scene through
  all 400 (the piles the bench times form late);
- 64 batched worlds of up to 1024 bodies (batches of 64 or more worlds solve
  in 256-thread workgroups);
- one world of 8200 bodies (above 8192 slots the sweep order is sorted
  device-wide, and the solver's and union-find's working sets leave LDS);
- update(dt) with 0, 1, 2 and more steps per call, and the spiral-of-death
  clamp (src/system.zig:508-534);
- writeBack's guards: NaN pose reset, 50 m/s and 30 rad/s clamps
  (src/solver.zig:246-296);
- MaxBodiesReached and slot reuse (src/system.zig:475-499).

Every comparison is exact (IEEE equality, NaN == NaN), including every
statistic (accumulated_time among them).
"""
import os

import numpy as np
import pytest

import vxp
from vxp import _abi, compare, scenes

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _lockstep(product, oracle, scene, frames, check_every=1, cfg_over=None, dts=None):
    over = dict(cfg_over or {})
    print(f"lockstep {scene.name} frames={frames} every={check_every} over={over} dts={dts}",
          flush=True)

    def log(f):
        if (f + 1) % 100 == 0:
            print(f"  frame {f + 1} equal", flush=True)

    f, msg, _ = compare.lockstep(scene, product, oracle, frames, cfg_over=over,
                                 check_every=check_every, dts=dts, log=log)
    assert f is None, f"frame {f}: {msg}"


@pytest.mark.timeout(900)
def test_c1_all_600_frames(product, oracle):
    _lockstep(product, oracle, scenes.c1_large(), 600, check_every=50,
              cfg_over=dict(cpu_threads=THREADS))


@pytest.mark.timeout(600)
def test_c2_all_400_frames(product, oracle):
    _lockstep(product, oracle, scenes.c2_rocks(), 400, check_every=50,
              cfg_over=dict(cpu_threads=THREADS))


@pytest.mark.timeout(600)
def test_64_worlds_up_to_1024_bodies(product, oracle):
    sc = scenes.batched_worlds(n_worlds=64, lo=8, hi=1024, blob_pool=32, name="c3_64")
    s = vxp.from_scene(sc, lib=product)
    assert s.tuning()["unit_threads"] == 256  # the batched bench's solve path
    s.close()
    _lockstep(product, oracle, sc, 90, check_every=30, cfg_over=dict(cpu_threads=THREADS))


@pytest.mark.timeout(600)
def test_one_world_of_8200_bodies(product, oracle):
    sc = scenes.c1_large(n_bodies=8200)
    _lockstep(product, oracle, sc, 30, check_every=10, cfg_over=dict(cpu_threads=THREADS))


def test_update_catch_up_and_spiral_clamp(product, oracle):
    # update(dt) with dt of 0.4, 2.5, 9 and 20 fixed steps (0, 2, 8 steps and
    # the clamp of the accumulator to one fixed step), 1, 0.3 and 0
    h = 1 / 60
    dts = [0.4 * h, 2.5 * h, 9.0 * h, 1.0 * h, 0.3 * h, 20.0 * h, 0.0, 1.0 * h]
    _lockstep(product, oracle, scenes.c0_small(), 96, dts=dts)


def test_update_catch_up_batched(product, oracle):
    sc = scenes.batched_worlds(n_worlds=4, lo=8, hi=64, blob_pool=4, name="c3_4")
    h = sc.dt
    _lockstep(product, oracle, sc, 40, cfg_over=dict(cpu_threads=THREADS),
              dts=[2.5 * h, 0.4 * h, 1.0 * h, 12.0 * h])


def test_writeback_guards(product, oracle):
    """NaN position, NaN rotation, 300 m/s and 100 rad/s set between frames on
    bodies inside the 64-cube pile; both libraries reset and clamp alike."""
    sc = scenes.c0_small()
    a = vxp.from_scene(sc, lib=product)
    b = vxp.from_scene(sc, lib=oracle)
    nan = float("nan")
    plan = {
        30: [(3, dict(position=(nan, 1.0, 0.0)))],
        40: [(5, dict(linear_velocity=(300.0, 10.0, -40.0))),
             (7, dict(angular_velocity=(0.0, 100.0, 20.0)))],
        50: [(9, dict(rotation=(0.0, nan, 0.0, 1.0)))],
        60: [(11, dict(linear_velocity=(0.0, -70.0, 0.0), angular_velocity=(45.0, 0.0, 0.0)))],
    }
    ids = a.bodies()["id"]
    for f in range(90):
        for slot, kw in plan.get(f, []):
            for s in (a, b):
                s.setState(s.findBodyById(int(ids[slot])), **kw)
        a.update(sc.dt)
        b.update(sc.dt)
        msg = (compare.diff_bodies(a.bodies(), b.bodies()) or
               compare.diff_manifolds(a.manifolds(), b.manifolds()) or
               compare.diff_stats(a.stats(), b.stats()))
        assert msg is None, f"frame {f}: {msg}"
        if f == 30:  # the NaN pose was reset to (0, 2, 0) with the identity rotation
            bb = a.bodies()[3]
            assert list(bb["position"]) == [0.0, 2.0, 0.0]
            assert list(bb["rotation"]) == [0.0, 0.0, 0.0, 1.0]
        if f == 40:
            lv = a.bodies()[5]["linear_velocity"]
            av = a.bodies()[7]["angular_velocity"]
            assert np.linalg.norm(lv) <= 50.0 * (1 + 1e-6)
            assert np.linalg.norm(av) <= 30.0 * (1 + 1e-6)
    a.close()
    b.close()


def test_max_bodies_reached_and_slot_reuse(product, oracle):
    runs = []
    for lib in (product, oracle):
        s = vxp.VoxelPhysicsSystem(lib=lib, max_bodies=5)
        cid = s.addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                            voxels=np.full((1, 512), vxp.INSIDE, np.uint8))
        hs = [s.createDynamicBody((0.0, 1.0 + 2.0 * i, 0.0), (0, 0, 0, 1), cid, 1.0)
              for i in range(5)]
        with pytest.raises(vxp.VxpError):
            s.createDynamicBody((0.0, 20.0, 0.0), (0, 0, 0, 1), cid, 1.0)
        for _ in range(10):
            s.update(1 / 60)
        assert s.destroyBody(hs[2]) == _abi.VXP_OK
        assert s.destroyBody(hs[2]) == _abi.VXP_ERR_INVALID_HANDLE
        h = s.createDynamicBody((0.5, 3.0, 0.0), (0, 0, 0, 1), cid, 1.0)
        assert h.index == hs[2].index and h.generation != hs[2].generation
        for _ in range(20):
            s.update(1 / 60)
        runs.append((s.bodies(), s.manifolds(), s.stats()))
        s.close()
    (ba, ma, sa), (bb, mb, sb) = runs
    assert compare.diff_bodies(ba, bb) is None
    assert compare.diff_manifolds(ma, mb) is None
    assert compare.diff_stats(sa, sb) is None


def test_wide_id_space_uses_64_bit_pair_keys(product, oracle):
    # 70 000 body slots: pair keys need 2 x 17 id bits, beyond 32, so the pair
    # sort runs on 64-bit keys; the sweep order is sorted device-wide (over
    # 8192 slots per world)
    sc = scenes.c0_small(frames=60)
    _lockstep(product, oracle, sc, 60, check_every=10, cfg_over=dict(max_bodies=70000))
