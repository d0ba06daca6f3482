"""The thread-pool narrow phase (enable_threading, src/system.zig:618-638,
src/narrow_phase.zig:1538-1632) on the GPU against the CPU restatement, frame
for frame and exact: the 256-body list (300 resting cubes: those past it fall
through the ground), the read-only cache of the parallel branch, the
MIN_PARALLEL_PAIRS boundary (7 and 8 pairs), worlds that switch branch as
their bodies land and sleep (the 64-cube scene), one thread, batched worlds
on either side of 256 bodies, and the reference's own configuration on the
4096-body scene (512-pair cap, tree pair order, thread pool).
"""
import os

import pytest

from vxp import _abi, compare, scenes

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
POOL = dict(enable_threading=1, thread_count=4)


def _lockstep(product, oracle, scene, frames, check_every=1, **over):
    print(f"threaded lockstep {scene.name} frames={frames} over={over}", flush=True)
    f, msg, _ = compare.lockstep(scene, product, oracle, frames, cfg_over=over,
                                 check_every=check_every)
    assert f is None, f"frame {f}: {msg}"


def test_body_list_of_256(product, oracle):
    _lockstep(product, oracle, scenes.resting_cubes(300), 40, **POOL)


def test_parallel_branch_read_only_cache(product, oracle):
    _lockstep(product, oracle, scenes.resting_cubes(64), 60, enable_sleeping=0, **POOL)


@pytest.mark.parametrize("n", [7, 8])
def test_min_parallel_pairs_boundary(product, oracle, n):
    _lockstep(product, oracle, scenes.resting_cubes(n), 40, enable_sleeping=0, **POOL)


def test_branch_changes_as_bodies_land_and_sleep(product, oracle):
    _lockstep(product, oracle, scenes.c0_small(frames=240), 240, **POOL)


def test_one_thread(product, oracle):
    _lockstep(product, oracle, scenes.c0_small(frames=90), 90, enable_threading=1, thread_count=1)


def test_batched_worlds_either_side_of_256(product, oracle):
    sc = scenes.batched_worlds(n_worlds=12, lo=100, hi=600, blob_pool=8, name="c3_thr")
    _lockstep(product, oracle, sc, 60, check_every=5, cpu_threads=THREADS, **POOL)


@pytest.mark.timeout(600)
def test_reference_configuration_4096_bodies(product, oracle):
    _lockstep(product, oracle, scenes.c1_large(), 120, check_every=10, pair_cap=512,
              pair_order=_abi.PAIR_ORDER_TREE, enable_coloring=1, **POOL)
