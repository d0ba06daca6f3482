"""Recorded steps (vxp_graph_enable: the step replayed from three HIP graphs
sized for capacities, with device-side counts) equal the eager steps bit for
bit: pairs, manifolds, bodies, island labels and every statistic, frame by
frame, through capacity growth (a step past a capacity is rolled back and
replayed with larger graphs), several steps per update() call, tuning windows
(which run eagerly) and the serving loop's admissions and retirements.
"""
import numpy as np
import pytest

import vxp
from vxp import compare, scenes, serve
from vxp.scenes import SplitMix64

pytestmark = pytest.mark.gpu
GiB = 1 << 30


def _recorded(s):
    s.setRecorded(True)


def _lockstep(product, scene, frames, check_every, **kw):
    print(f"recorded vs eager: {scene.name} frames={frames} every={check_every}", flush=True)
    f, msg, (a, b) = compare.lockstep(scene, product, product, frames, check_every=check_every,
                                      prep_a=_recorded, **kw)
    info = a.graphInfo()
    print(f"  graph info {info}", flush=True)
    assert f is None, f"frame {f}: {msg}"
    a.close()
    b.close()
    return info


def test_c0_every_frame(product):
    info = _lockstep(product, scenes.c0_small(), 300, 1)
    assert info["recorded_steps"] > 200


def test_batched_worlds(product):
    sc = scenes.batched_worlds(n_worlds=64, lo=8, hi=1024, blob_pool=32, name="c3_64")
    info = _lockstep(product, sc, 200, 20)
    assert info["recorded_steps"] > 100


def test_c1_capacity_growth(product):
    # the column collapses into piles: pairs, work items and candidates grow
    # past the capacities taken from the first recorded steps
    info = _lockstep(product, scenes.c1_large(), 300, 50)
    assert info["recorded_steps"] > 200 and info["redos"] >= 1 and info["recordings"] >= 3


def test_several_steps_per_update(product):
    sc = scenes.batched_worlds(n_worlds=16, lo=8, hi=256, blob_pool=16, name="c3_16")
    info = _lockstep(product, sc, 120, 10, dts=[sc.dt * 0.4, sc.dt * 2.5, sc.dt, sc.dt * 3])
    assert info["recorded_steps"] > 80


def test_candidate_capacity_growth(product):
    # tuning fixed (no tuning windows): recording starts after the first frame,
    # while the bodies still fall, and the contact candidates grow past their
    # capacity mid-step (segment 3 recorded again, nothing rolled back)
    sc = scenes.batched_worlds(n_worlds=64, lo=8, hi=1024, blob_pool=32, name="c3_64")
    fixed = dict(pair_color_max=128, unit_threads=256, broad_phase="sweep", probe_threads=64,
                 probe_xcd_run=64)
    info = _lockstep(product, sc, 150, 25, tune_a=fixed)
    assert info["recorded_steps"] >= 140 and info["cap_candidates"] > 8192
    assert info["recordings"] >= 3


def test_serving_recorded_equals_eager(product):
    sc = scenes.serving_pool(n_worlds=10, lo=8, hi=64, workers=1)
    rng = SplitMix64(11)
    # enough frames that the batch's first tuning round (whose steps run
    # eagerly) ends and recorded steps follow
    reqs = [(w, [4, 8][rng.randint(0, 1)]) for w in range(len(sc.worlds)) for _ in range(8)]
    out = []
    for rec in (False, True):
        srv = serve.from_scene(sc, slots=4, budget_bytes=8 * GiB, lib=product)
        srv.set_recorded(rec)
        for tag, (w, k) in enumerate(reqs):
            srv.submit(w, k, tag)
            if tag % 4 == 3:
                srv.tick()
        for _ in range(10000):
            if srv.info()["completed"] == len(reqs):
                break
            srv.tick()
        assert srv.info()["completed"] == len(reqs)
        if rec:
            gi = srv.graph_info()
            print(f"  server graph info {gi}", flush=True)
            assert gi["recorded_steps"] > 0
        out.append([(srv.bodies(w), srv.manifold_array(w), srv.stats(w))
                    for w in range(len(sc.worlds))])
        srv.close()
    for w in range(len(sc.worlds)):
        (ba, ma, sa), (bb, mb, sb) = out[0][w], out[1][w]
        assert np.array_equal(ba, bb), f"world {w}: bodies differ"
        assert np.array_equal(ma, mb), f"world {w}: manifolds differ"
        assert compare.diff_stats(sa, sb) is None, f"world {w}"
