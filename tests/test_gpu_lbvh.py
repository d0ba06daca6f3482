"""The LBVH broad phase (a Karras radix tree per world over the Morton codes of
the fat-box centres, one stackless query per leaf) finds the pair set of
BroadPhase.findOverlappingPairs (src/broad_phase.zig:488-526): every frame's
pairs, and everything downstream of them, equal the CPU restatement's and the
sort-and-sweep's.

Covered: the 64-cube scene every frame; the 4096-body scene through its 600
frames (the pile of its last frames); 64 batched worlds of up to 1024 bodies;
worlds of 1 to 3 bodies (a tree of one leaf, of two, of none besides the
ground); one world of 8200 bodies (above 8192 slots the Morton order is
sorted device-wide); 70 000 body slots (64-bit pair keys); bodies destroyed
and created between frames (leaves come and go).  Every comparison is exact.
"""
import os

import numpy as np
import pytest

import vxp
from vxp import compare, scenes

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
LBVH = dict(broad_phase="lbvh")


def _lockstep(product, oracle, scene, frames, check_every=1, cfg_over=None):
    over = dict(cfg_over or {})
    print(f"lbvh lockstep {scene.name} frames={frames} every={check_every} over={over}",
          flush=True)
    f, msg, (sa, _) = compare.lockstep(scene, product, oracle, frames, cfg_over=over,
                                       check_every=check_every, tune_a=LBVH)
    assert f is None, f"frame {f}: {msg}"
    assert sa.tuning()["broad_phase"] == "lbvh"


def test_lbvh_c0_every_frame(product, oracle):
    _lockstep(product, oracle, scenes.c0_small(frames=120), 120)


@pytest.mark.timeout(900)
def test_lbvh_c1_all_600_frames(product, oracle):
    _lockstep(product, oracle, scenes.c1_large(), 600, check_every=100,
              cfg_over=dict(cpu_threads=THREADS))


@pytest.mark.timeout(600)
def test_lbvh_64_worlds_up_to_1024_bodies(product, oracle):
    sc = scenes.batched_worlds(n_worlds=64, lo=8, hi=1024, blob_pool=32, name="c3_64")
    _lockstep(product, oracle, sc, 60, check_every=20, cfg_over=dict(cpu_threads=THREADS))


def test_lbvh_tiny_worlds(product, oracle):
    sc = scenes.batched_worlds(n_worlds=24, lo=1, hi=4, blob_pool=4, name="c3_tiny")
    _lockstep(product, oracle, sc, 40)


@pytest.mark.timeout(600)
def test_lbvh_one_world_of_8200_bodies(product, oracle):
    _lockstep(product, oracle, scenes.c1_large(n_bodies=8200), 20, check_every=10,
              cfg_over=dict(cpu_threads=THREADS))


def test_lbvh_wide_id_space(product, oracle):
    _lockstep(product, oracle, scenes.c0_small(frames=40), 40, check_every=10,
              cfg_over=dict(max_bodies=70000))


def test_lbvh_equals_sweep_with_bodies_coming_and_going(product):
    """Two product systems, one per broad phase, under the same destroy /
    create sequence: pairs, manifolds and bodies equal every frame."""
    sc = scenes.c0_small(frames=80)
    sys_ = [vxp.from_scene(sc, lib=product) for _ in range(2)]
    sys_[0].setTuning(broad_phase="sweep")
    sys_[1].setTuning(broad_phase="lbvh")
    rng = np.random.default_rng(7)
    cid = sys_[0].addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                              voxels=np.full((1, 512), vxp.INSIDE, np.uint8))
    assert sys_[1].addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                               voxels=np.full((1, 512), vxp.INSIDE, np.uint8)) == cid
    for f in range(80):
        if f % 10 == 5:
            bs = sys_[0].bodies()
            ids = bs["id"][(bs["active"] == 1) & (bs["body_type"] == vxp.DYNAMIC)]
            victims = rng.choice(len(ids), 3, replace=False)
            for s in sys_:
                for v in victims:
                    s.destroyBody(s.findBodyById(int(ids[v])))
                for k in range(2):
                    s.createDynamicBody((0.3 * k, 3.0 + f * 0.01, 0.2 * k), (0, 0, 0, 1), cid, 1.0)
        for s in sys_:
            s.update(sc.dt)
        a, b = sys_
        pa, pb = a.pairs(0), b.pairs(0)
        assert len(pa) == len(pb) and (pa == pb).all(), f"frame {f}: pairs differ"
        msg = (compare.diff_bodies(a.bodies(), b.bodies()) or
               compare.diff_manifolds(a.manifolds(), b.manifolds()) or
               compare.diff_stats(a.stats(), b.stats()))
        assert msg is None, f"frame {f}: {msg}"
    for s in sys_:
        s.close()


def test_lbvh_rejects_unknown_broad_phase(product):
    s = vxp.from_scene(scenes.c0_small(frames=1), lib=product)
    with pytest.raises(ValueError):
        s.setTuning(broad_phase="octree")
    s.close()
