"""Tree pair order and the pair cap (cap-faithful mode): the broad phase's
AABB trees (src/broad_phase.zig:64-529), their emission order, and the
512-entry filtered-pair array (src/system.zig:594-616).

CPU: the restatement's trees against structural invariants (links, heights,
boxes equal to the merge of the children, leaves = bodies, free list) after
creations, moves and destructions; the emitted order against a depth-first
walk of the exported trees (dynamic leaves in node-index order, child2 popped
before child1, self pairs to higher ids first, then the static tree); the
pair set and the dynamics against the key order (the order changes only
which brick-pair cache entries are hit or evicted, not any result); and the
cap against its definition (only the first 512 pairs in emission order that
are not both asleep are tested).  GPU: the HIP pipeline against the
restatement frame for frame, trees node for node, with and without the cap.
"""
import numpy as np
import pytest

import vxp
from vxp import _abi, compare, scenes

NULL = 0xFFFFFFFF
TREE = dict(pair_order=_abi.PAIR_ORDER_TREE)


def _check_tree(nodes, root, body_ids):
    live = nodes["height"] >= 0
    leaves = set()
    if root != NULL:
        assert live[root] and nodes["parent"][root] == NULL
    for i in np.nonzero(live)[0]:
        n = nodes[i]
        if n["child1"] == NULL:
            leaves.add(int(n["body_id"]))
            assert n["height"] == 0
            continue
        a, b = int(n["child1"]), int(n["child2"])
        assert nodes["parent"][a] == i and nodes["parent"][b] == i
        assert n["height"] == 1 + max(nodes["height"][a], nodes["height"][b])
        np.testing.assert_array_equal(n["min"], np.fmin(nodes["min"][a], nodes["min"][b]))
        np.testing.assert_array_equal(n["max"], np.fmax(nodes["max"][a], nodes["max"][b]))
    assert leaves == set(body_ids)
    if leaves:
        assert live.sum() == 2 * len(leaves) - 1
    # every live node is reachable from the root
    seen, stack = set(), [root] if root != NULL else []
    while stack:
        i = stack.pop()
        seen.add(i)
        if nodes["child1"][i] != NULL:
            stack += [int(nodes["child1"][i]), int(nodes["child2"][i])]
    assert seen == set(np.nonzero(live)[0].tolist())


def _walk(nodes, root, lo, hi):
    """Depth-first query, child2 popped first (src/broad_phase.zig:461-485)."""
    out, stack = [], [root] if root != NULL else []
    while stack:
        i = stack.pop()
        n = nodes[i]
        if not ((n["min"] <= hi).all() and (n["max"] >= lo).all()):
            continue
        if n["child1"] == NULL:
            out.append(int(n["body_id"]))
        else:
            stack += [int(n["child1"]), int(n["child2"])]
    return out


def _emission(s, w=0):
    dyn, droot = s.tree(1, w)
    sta, sroot = s.tree(0, w)
    out = []
    for q in range(2):
        for i in range(len(dyn)):
            n = dyn[i]
            if n["height"] < 0 or n["child1"] != NULL:
                continue
            me = int(n["body_id"])
            for o in _walk(dyn if q == 0 else sta, droot if q == 0 else sroot, n["min"], n["max"]):
                if q == 1 or o > me:
                    out.append((min(me, o) << 32) | max(me, o))
    return out


def _ids_by_tree(s, w=0):
    b = s.bodies(w)
    act = b[b["active"] == 1]
    st = act["body_type"] == _abi.STATIC
    return act["id"][st].tolist(), act["id"][~st].tolist()


def test_tree_invariants_and_emission_order(oracle):
    sc = scenes.c0_small(n_side=4, frames=160)
    s = vxp.from_scene(sc, lib=oracle, cfg_over=TREE)
    cube = [c.name for c in sc.colliders].index("cube3")
    for f in range(160):
        if f == 60:
            s.destroyBody(s.handles[0][5])
            s.destroyBody(s.handles[0][9])
        if f == 90:  # reuses freed slots and tree nodes
            s.createDynamicBody((0.3, 4.0, 0.2), (0, 0, 0, 1), cube, 1.0)
            s.createKinematicBody((1.5, 0.5, 1.5), (0, 0, 0, 1), cube)
        s.update(sc.dt)
        st_ids, dyn_ids = _ids_by_tree(s)
        for which, ids in ((0, st_ids), (1, dyn_ids)):
            nodes, root = s.tree(which)
            _check_tree(nodes, root, ids)
        assert s.pairs().tolist() == _emission(s), f"frame {f}"
    s.close()


def test_tree_order_changes_no_result(oracle):
    sc = scenes.c0_small(n_side=5, frames=150)
    a = vxp.from_scene(sc, lib=oracle)
    b = vxp.from_scene(sc, lib=oracle, cfg_over=TREE)
    differs = False
    for f in range(150):
        a.update(sc.dt)
        b.update(sc.dt)
        pa, pb = a.pairs(), b.pairs()
        assert sorted(pa.tolist()) == sorted(pb.tolist()), f"frame {f}"
        differs |= pa.tolist() != pb.tolist()
        assert compare.diff_bodies(a.bodies(), b.bodies()) is None, f"frame {f}"
        assert compare.diff_manifolds(a.manifolds(), b.manifolds()) is None, f"frame {f}"
    assert differs, "the tree order never differed from the key order"
    a.close()
    b.close()


def _cap_scene():
    return scenes.c1_large(n_bodies=300, ground_vox=96, frames=80)


@pytest.mark.parametrize("order", [_abi.PAIR_ORDER_KEY, _abi.PAIR_ORDER_TREE])
def test_pair_cap_keeps_the_first_awake_pairs(oracle, order):
    sc = _cap_scene()
    s = vxp.from_scene(sc, lib=oracle, cfg_over=dict(pair_order=order, pair_cap=512))
    over = 0
    for f in range(80):
        pre = s.bodies()
        asleep = {int(b["id"]) for b in pre if b["active"] and b["is_sleeping"]}
        s.update(sc.dt)
        st = s.stats()
        pairs = s.pairs().tolist()
        awake = [k for k in pairs if not ((k >> 32) in asleep and (k & 0xFFFFFFFF) in asleep)]
        kept = set(awake[:512])
        over += len(awake) > 512
        assert st["narrow_phase_tested"] == len(kept)
        for m in s.manifolds():
            k = (min(m.body_a_id, m.body_b_id) << 32) | max(m.body_a_id, m.body_b_id)
            assert k in kept or ((k >> 32) in asleep and (k & 0xFFFFFFFF) in asleep), f"frame {f}"
    assert over > 40
    s.close()


@pytest.mark.gpu
def test_tree_order_gpu_matches_restatement(product, oracle):
    sc = scenes.c0_small(n_side=4, frames=160)
    cube = [c.name for c in sc.colliders].index("cube3")
    sp = vxp.from_scene(sc, lib=product, cfg_over=TREE)
    so = vxp.from_scene(sc, lib=oracle, cfg_over=TREE)
    for f in range(160):
        for s in (sp, so):
            if f == 60:
                s.destroyBody(s.handles[0][5])
                s.destroyBody(s.handles[0][9])
            if f == 90:
                s.createDynamicBody((0.3, 4.0, 0.2), (0, 0, 0, 1), cube, 1.0)
                s.createKinematicBody((1.5, 0.5, 1.5), (0, 0, 0, 1), cube)
            s.update(sc.dt)
        assert sp.pairs().tolist() == so.pairs().tolist(), f"frame {f}: pair order"
        for which in (0, 1):
            (na, ra), (nb, rb) = sp.tree(which), so.tree(which)
            assert ra == rb and len(na) == len(nb), f"frame {f}: tree {which}"
            for k in ("min", "max"):  # IEEE equality: the sign of a zero may differ
                assert np.array_equal(na[k], nb[k], equal_nan=True), f"frame {f}: tree {which} {k}"
            for k in ("body_id", "parent", "child1", "child2", "height"):
                assert np.array_equal(na[k], nb[k]), f"frame {f}: tree {which} {k}"
        msg = compare.diff_bodies(sp.bodies(), so.bodies()) or \
            compare.diff_manifolds(sp.manifolds(), so.manifolds()) or \
            compare.diff_stats(sp.stats(), so.stats())
        assert msg is None, f"frame {f}: {msg}"
    sp.close()
    so.close()


@pytest.mark.gpu
@pytest.mark.parametrize("order", [_abi.PAIR_ORDER_KEY, _abi.PAIR_ORDER_TREE])
def test_pair_cap_gpu_matches_restatement(product, oracle, order):
    sc = _cap_scene()
    f, msg, _ = compare.lockstep(sc, product, oracle, 80,
                                 cfg_over=dict(pair_order=order, pair_cap=512))
    assert f is None, f"frame {f}: {msg}"


@pytest.mark.gpu
def test_tree_order_batched_worlds_with_cap(product, oracle):
    # per-world trees and per-world cap ranks side by side in one system
    sc = scenes.batched_worlds(n_worlds=6, lo=8, hi=96, blob_pool=8)
    f, msg, _ = compare.lockstep(sc, product, oracle, 90, check_every=15,
                                 cfg_over=dict(pair_order=_abi.PAIR_ORDER_TREE, pair_cap=48))
    assert f is None, f"frame {f}: {msg}"
