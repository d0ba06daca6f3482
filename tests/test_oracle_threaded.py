"""Pins of the CPU restatement's thread-pool narrow phase (enable_threading,
src/system.zig:618-638, src/narrow_phase.zig:1538-1632), before it checks the
GPU.  Each test compares it with the restatement's sequential narrow phase on
a case where the reference's two paths must agree, or where a named part of
the thread-pool path (the 256-body list, the read-only cache, the
MIN_PARALLEL_PAIRS branch, pairs_tested before the body lookup) must show.
"""
import numpy as np
import pytest

import vxp
from vxp import compare, scenes


def _run(oracle, sc, frames, **over):
    s = vxp.from_scene(sc, lib=oracle, cfg_over=over)
    per_frame = []
    for _ in range(frames):
        s.update(sc.dt)
        per_frame.append((s.bodies(), s.manifolds(), s.stats()))
    s.close()
    return per_frame


def _same(a, b, stats=True):
    for f, ((ba, ma, sa), (bb, mb, sb)) in enumerate(zip(a, b)):
        msg = compare.diff_bodies(ba, bb) or compare.diff_manifolds(ma, mb)
        if not msg and stats:
            msg = compare.diff_stats(sa, sb)
        assert msg is None, f"frame {f}: {msg}"


def test_one_thread_is_the_sequential_narrow_phase(oracle):
    # thread_count 1: processPairsParallel's sequential branch calls processPair
    # on every pair of listed bodies; 65 bodies are all listed
    sc = scenes.c0_small(frames=120)
    _same(_run(oracle, sc, 120, enable_threading=1, thread_count=1), _run(oracle, sc, 120))


def test_parallel_branch_reads_the_cache_and_never_writes_it(oracle):
    # 64 resting cubes: 64 cube-ground pairs in every frame, so the parallel
    # branch runs every frame.  Its cache stays empty: no hit ever, one miss
    # per cache-eligible pair, and the motion equals the sequential narrow
    # phase without a cache.  The sequential narrow phase with the cache hits
    # once cubes rest (a slip that stored misses would do the same).
    # (sleeping off: sleeping pairs would leave fewer than 8 in the narrow phase)
    sc = scenes.resting_cubes(64, spacing=1.0)
    thr = _run(oracle, sc, 60, enable_threading=1, thread_count=4, enable_sleeping=0)
    off = _run(oracle, sc, 60, use_brick_pair_cache=0, enable_sleeping=0)
    seq = _run(oracle, sc, 60, enable_sleeping=0)
    _same(thr, off, stats=False)
    for f, (_, _, st) in enumerate(thr):
        assert st["broad_phase_pairs"] == st["narrow_phase_tested"] == 64, f
        assert st["cache_hits"] == 0, f
    # once the cubes touch, every pair passes SAT and looks the cache up
    assert all(st["cache_misses"] == 64 for _, _, st in thr[10:])
    assert all(st["cache_hits"] == 0 for _, _, st in thr)
    assert sum(st["cache_hits"] for _, _, st in seq) > 0


def test_fewer_than_eight_pairs_take_the_sequential_branch(oracle):
    # 5 resting cubes: 5 pairs < MIN_PARALLEL_PAIRS, so processPair runs with
    # its cache writes: the same as no thread pool at all
    sc = scenes.resting_cubes(5, spacing=1.0)
    seq = _run(oracle, sc, 60)
    _same(_run(oracle, sc, 60, enable_threading=1, thread_count=4), seq)
    assert sum(st["cache_hits"] for _, _, st in seq) > 0


def test_only_the_first_256_active_bodies_collide(oracle):
    # 300 resting cubes in slots 1..300 behind the ground in slot 0: the body
    # list holds slots 0..255, so cubes in slots 256..300 have no contact and
    # fall through the ground, the others rest; every pair that reaches the
    # narrow phase counts as tested (parallel branch, counted before lookup)
    sc = scenes.resting_cubes(300, spacing=1.0)
    frames = _run(oracle, sc, 30, enable_threading=1, thread_count=4)
    bodies, mans, st = frames[-1]
    ids = bodies["id"]
    listed = set(int(i) for i in ids[:256])
    assert len(mans) > 0
    for m in mans:
        assert m.body_a_id in listed and m.body_b_id in listed
    y = bodies["position"][:, 1]
    assert (y[256:] < 0.0).all()          # fell through
    assert (y[1:256] > 0.25).all()        # rest on the ground
    assert st["narrow_phase_tested"] == st["broad_phase_pairs"] - st["pairs_skipped_sleeping"]
    # without the thread pool every cube rests
    y_seq = _run(oracle, sc, 30)[-1][0]["position"][:, 1]
    assert (y_seq[1:] > 0.25).all()


@pytest.mark.parametrize("n,parallel", [(7, False), (8, True)])
def test_min_parallel_pairs_is_eight(oracle, n, parallel):
    # MIN_PARALLEL_PAIRS = 8 (src/narrow_phase.zig:41,1546): 8 resting cubes
    # take the parallel branch (the cache is never written, so no hit), 7 the
    # sequential one (hits once the cubes rest); sleeping off keeps n pairs
    sc = scenes.resting_cubes(n, spacing=1.0)
    thr = _run(oracle, sc, 40, enable_threading=1, thread_count=2, enable_sleeping=0)
    hits = sum(st["cache_hits"] for _, _, st in thr)
    assert (hits == 0) == parallel
    assert all(st["broad_phase_pairs"] == n for _, _, st in thr)


def test_thread_count_zero_is_the_host_cpu_count(oracle):
    # thread_count 0 sizes the pool by the host's CPU count (2 or more on any
    # host this runs on): the parallel branch, as with an explicit 2
    sc = scenes.resting_cubes(8, spacing=1.0)
    auto = _run(oracle, sc, 40, enable_threading=1, thread_count=0, enable_sleeping=0)
    _same(auto, _run(oracle, sc, 40, enable_threading=1, thread_count=2, enable_sleeping=0))
    assert sum(st["cache_hits"] for _, _, st in auto) == 0
