"""Serving loop (vxp_server_*): continuous batching over resident worlds.

Each world's result is the same alone and in the mix: worlds advanced by
interleaved requests while other worlds share every launch end bit-identical
(bodies, manifolds, counters) to the same world advanced alone by the plain
single-world path for the same number of frames, which the parity suite pins
to the CPU restatement.  Also: FIFO per world, request bookkeeping, and
admission by the HBM budget.
"""
import numpy as np
import pytest

import vxp
from vxp import _abi, compare, scenes, serve
from vxp.scenes import SplitMix64

pytestmark = pytest.mark.gpu
GiB = 1 << 30


def _alone(product, sc, w, frames):
    s = vxp.from_scene(sc, lib=product, worlds=[w])
    for _ in range(frames):
        s.update(sc.dt)
    out = (s.bodies(0), s.manifold_array(0), s.stats(0))
    s.close()
    return out


def _raw(structured):
    return np.ascontiguousarray(structured).view(np.uint8).reshape(len(structured), -1)


def test_alone_equals_in_the_mix(product):
    sc = scenes.serving_pool(n_worlds=10, lo=8, hi=64, workers=1)
    srv = serve.from_scene(sc, slots=4, budget_bytes=8 * GiB, lib=product)
    rng = SplitMix64(7)
    reqs = [(w, [1, 2, 4, 8][rng.randint(0, 3)]) for w in range(len(sc.worlds)) for _ in range(3)]
    order = sorted(range(len(reqs)), key=lambda i: rng.next_u64())
    reqs = [reqs[i] for i in order]
    frames = [0] * len(sc.worlds)
    tags = {}
    # half now, the rest while the first ones are on the GPU
    for tag, (w, k) in enumerate(reqs[: len(reqs) // 2]):
        srv.submit(w, k, tag)
        tags[tag] = (w, k)
        frames[w] += k
    print(f"serving: {len(reqs)} requests over {len(sc.worlds)} worlds, 4 slots", flush=True)
    for _ in range(3):
        srv.tick()
    for tag, (w, k) in enumerate(reqs[len(reqs) // 2:], start=len(reqs) // 2):
        srv.submit(w, k, tag)
        tags[tag] = (w, k)
        frames[w] += k
    for _ in range(10000):
        if srv.info()["completed"] == len(reqs):
            break
        srv.tick()
    done = srv.poll()
    info = srv.info()
    assert info["completed"] == len(reqs) and len(done) == len(reqs), info
    assert info["occupied"] == 0 and info["queued"] == 0
    for d in done:
        w, k = tags[int(d["tag"])]
        assert d["world"] == w and d["frames"] == k
        assert d["submit_s"] <= d["admit_s"] <= d["done_s"]
    # requests of one world complete in submission order
    for w in range(len(sc.worlds)):
        mine = [int(d["tag"]) for d in done if d["world"] == w]
        assert mine == sorted(mine), (w, mine)
    for w in range(len(sc.worlds)):
        bodies, mans, stats = _alone(product, sc, w, frames[w])
        n = len(sc.worlds[w])
        got = srv.bodies(w)[:n]
        assert np.array_equal(got, _raw(bodies)[:n]), f"world {w}: bodies differ"
        assert np.array_equal(srv.manifold_array(w), mans), f"world {w}: manifolds differ"
        assert compare.diff_stats(srv.stats(w), stats) is None, f"world {w}"
    srv.close()


def test_matches_restatement_through_the_server(product, oracle):
    sc = scenes.serving_pool(n_worlds=3, lo=8, hi=40, workers=1)
    srv = serve.from_scene(sc, slots=2, budget_bytes=4 * GiB, lib=product)
    for tag, (w, k) in enumerate([(0, 8), (1, 4), (2, 8), (0, 4), (1, 8)]):
        srv.submit(w, k, tag)
    while srv.info()["completed"] < 5:
        srv.tick()
    frames = {0: 12, 1: 12, 2: 8}
    for w, f in frames.items():
        o = vxp.from_scene(sc, lib=oracle, worlds=[w])
        for _ in range(f):
            o.update(sc.dt)
        n = len(sc.worlds[w])
        ob = o.bodies(0)[:n]
        sb = np.frombuffer(srv.bodies(w)[:n].tobytes(), dtype=ob.dtype)
        assert compare.diff_bodies(sb, ob) is None, f"world {w}"
        sm = [_abi.Manifold.from_buffer_copy(row.tobytes()) for row in srv.manifold_array(w)]
        assert compare.diff_manifolds(sm, o.manifolds(0)) is None, f"world {w}"
        assert compare.diff_stats(srv.stats(w), o.stats(0)) is None
        o.close()
    srv.close()


def test_hbm_budget_admits_worlds(product):
    sc = scenes.serving_pool(n_worlds=2, lo=8, hi=16, workers=1)
    cfg = dict(sc.config, max_bodies=64)
    big = serve.Server(cfg, 1000, 4, 64 * GiB, lib=product)
    assert big.info()["capacity"] == 1000
    big.close()
    small = serve.Server(cfg, 1000, 4, 16 << 20, lib=product)
    cap = small.info()["capacity"]
    assert 0 < cap < 1000
    for _ in range(cap):
        small.create_world()
    with pytest.raises(vxp.VxpError, match="HBM budget"):
        small.create_world()
    assert small.info()["resident_bytes"] <= small.info()["budget_bytes"]
    small.close()
