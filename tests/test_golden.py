"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the CPU restatement): the restatement keeps reproducing them (CPU tests),
and the HIP pipeline reproduces the same states exactly (GPU tests)."""
import os

import numpy as np
import pytest

import vxp
from vxp import _abi, compare, scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAT_KEYS = ("broad_phase_pairs", "narrow_phase_tested", "narrow_phase_colliding",
             "contacts_generated", "voxels_checked", "cache_hits", "cache_misses", "manifolds",
             "solver_manifolds", "colors", "island_count", "largest_island_size",
             "sleeping_bodies", "position_iterations_run", "awake_bodies_stepped", "solver_points")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _check_world(sysobj, bodies_raw, manifolds_raw, w=0):
    got = sysobj.bodies(w)
    want = np.frombuffer(bodies_raw.tobytes(), dtype=got.dtype)
    assert compare.diff_bodies(got, want) is None
    mans = [_abi.Manifold.from_buffer_copy(r.tobytes()) for r in manifolds_raw]
    assert compare.diff_manifolds(sysobj.manifolds(w), mans) is None


def _stats_vec(st):
    return np.array([st[k] for k in STAT_KEYS], dtype=np.float64)


def _c0(lib):
    g = _load("c0_f150.npz")
    sc = scenes.c0_small()
    s = vxp.from_scene(sc, lib=lib)
    for _ in range(int(g["frames"])):
        s.update(sc.dt)
    _check_world(s, g["bodies"], g["manifolds"])
    assert np.array_equal(_stats_vec(s.stats()), g["stats"])
    s.close()


def _c2(lib, threads):
    g = _load("c2_f20.npz")
    sc = scenes.c2_rocks()
    s = vxp.from_scene(sc, lib=lib, cfg_over=dict(cpu_threads=threads))
    for _ in range(int(g["frames"])):
        s.update(sc.dt)
    assert np.array_equal(s.pairs(0), g["pairs"])
    assert np.array_equal(_stats_vec(s.stats()), g["stats"])
    s.close()


def _c3(lib):
    g = _load("c3w4_f60.npz")
    sc = scenes.batched_worlds(n_worlds=4, lo=8, hi=64, blob_pool=8)
    s = vxp.from_scene(sc, lib=lib)
    for _ in range(int(g["frames"])):
        s.update(sc.dt)
    for w in range(4):
        _check_world(s, g[f"bodies{w}"], g[f"manifolds{w}"], w)
    s.close()


def test_restatement_reproduces_c0(oracle):
    _c0(oracle)


def test_restatement_reproduces_c2_pairs(oracle):
    _c2(oracle, min(8, os.cpu_count() or 1))


def test_restatement_reproduces_batched(oracle):
    _c3(oracle)


@pytest.mark.gpu
def test_gpu_reproduces_c0(product):
    _c0(product)


@pytest.mark.gpu
def test_gpu_reproduces_c2_pairs(product):
    _c2(product, 0)


@pytest.mark.gpu
def test_gpu_reproduces_batched(product):
    _c3(product)
