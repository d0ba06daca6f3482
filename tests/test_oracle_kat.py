"""Known-answer tests that pin the CPU restatement (oracle/) before it is
trusted as the checker.  Each test cites what pins it: an inline test of the
reference (re-expressed here), a closed form, or an invariant that a plausible
slip (dropped term, wrong sign, transposed operand, wrong index) would break.
"""
import ctypes as C
import math

import numpy as np
import pytest

import vxp
from vxp import scenes

F3 = C.c_float * 3
F4 = C.c_float * 4
F9 = C.c_float * 9
F10 = C.c_float * 10
I3 = C.c_int32 * 3


def _bind(lib, name, res, args):
    fn = getattr(lib.dll, name)
    fn.restype, fn.argtypes = res, args
    return fn


@pytest.fixture(scope="module")
def k(oracle):
    L = oracle
    P = C.POINTER
    return dict(
        rotate=_bind(L, "vxo_quat_rotate", None, [P(C.c_float)] * 3),
        axis_angle=_bind(L, "vxo_quat_from_axis_angle", None, [P(C.c_float), C.c_float, P(C.c_float)]),
        integrate=_bind(L, "vxo_integrate_rotation", None, [P(C.c_float), P(C.c_float), C.c_float, P(C.c_float)]),
        inverse=_bind(L, "vxo_mat3_inverse", None, [P(C.c_float), P(C.c_float)]),
        classify=_bind(L, "vxo_classify_voxel", C.c_uint32, [C.c_uint32, C.c_uint32]),
        normal_index=_bind(L, "vxo_compute_normal_index", C.c_uint32, [C.c_uint32]),
        should=_bind(L, "vxo_should_check", C.c_uint32, [C.c_uint32, C.c_uint32]),
        fid=_bind(L, "vxo_feature_id", C.c_uint64, [P(C.c_int32), P(C.c_int32)]),
        soft=_bind(L, "vxo_soft_params", None, [C.c_float, C.c_float, C.c_float, P(C.c_float)]),
        orth=_bind(L, "vxo_compute_orthogonal", None, [P(C.c_float), P(C.c_float)]),
        sat=_bind(L, "vxo_sat_early_out", C.c_int, [P(C.c_float), P(C.c_float)]),
        replay=_bind(L, "vxo_manifold_replay", C.c_int, [C.c_uint32, P(C.c_float), P(C.c_uint64), P(C.c_int32)]),
        cache=_bind(L, "vxo_cache_script", None, [C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_float), P(C.c_int64)]),
        islands=_bind(L, "vxo_island_labels", C.c_uint32, [C.c_uint32, P(C.c_uint32), P(C.c_uint32), C.c_uint32, P(C.c_uint8), P(C.c_uint32), P(C.c_uint32)]),
        refcolor=_bind(L, "vxo_ref_coloring", None, [C.c_uint32, P(C.c_uint32), P(C.c_uint32), P(C.c_uint8)]),
        mass=_bind(L, "vxo_mass_properties", C.c_int, [C.c_void_p, C.c_uint32, C.c_float, P(C.c_float)]),
        info=_bind(L, "vxo_collider_info", C.c_int, [C.c_void_p, C.c_uint32, P(C.c_int32), P(C.c_float), P(C.c_uint32)]),
        brick=_bind(L, "vxo_voxel_to_brick", None, [P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
        warm=_bind(L, "vxo_warm_start", None, [C.c_uint32, P(C.c_uint64), C.c_uint32, P(C.c_uint64),
                                               P(C.c_float), P(C.c_float)]),
        detect=_bind(L, "vxo_detect_pair", C.c_int, [C.c_void_p, C.c_uint32, P(C.c_float), C.c_uint32, P(C.c_float), P(vxp._abi.Manifold), P(C.c_uint32), P(C.c_uint32)]),
    )


def rot(k, q, v):
    o = F3()
    k["rotate"](F4(*q), F3(*v), o)
    return np.array(o)


def axis_angle(k, axis, ang):
    o = F4()
    k["axis_angle"](F3(*axis), ang, o)
    return np.array(o)


# ---- src/math.zig ------------------------------------------------------------

def test_quat_90_degree_rotation(k):
    # src/math.zig:816-824
    q = axis_angle(k, (0, 0, 1), math.pi / 2)
    np.testing.assert_allclose(rot(k, q, (1, 0, 0)), (0, 1, 0), atol=1e-6)


def test_rotation_matches_rotation_matrix(k):
    # closed form: q v q* equals R(q) v for unit q; a sign or index slip breaks it
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = rng.normal(size=3)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        np.testing.assert_allclose(rot(k, q, v), R @ v, rtol=1e-5, atol=1e-5)


def test_rotate_renormalizes_only_outside_band(k):
    # src/math.zig:267-275: a quaternion with |q|^2 = 4 is normalized first,
    # one with |q|^2 = 1.005 is used as is
    q = np.array([0.0, 0.0, math.sqrt(2) / 2, math.sqrt(2) / 2]) * 2.0
    np.testing.assert_allclose(rot(k, q, (1, 0, 0)), (0, 1, 0), atol=1e-6)
    s = math.sqrt(1.005)
    v = rot(k, (0.0, 0.0, 0.0, s), (1, 0, 0))
    # identity-like quat scaled by s: v + 2w(qv x v) + 2 qv x (qv x v) = v
    np.testing.assert_allclose(v, (1, 0, 0), atol=0)


def test_mat3_inverse(k):
    # src/math.zig:835-846
    m = np.array([1, 2, 3, 0, 1, 4, 5, 6, 0], np.float32)
    o = F9()
    k["inverse"](F9(*m), o)
    np.testing.assert_allclose(m.reshape(3, 3) @ np.array(o).reshape(3, 3), np.eye(3), atol=1e-4)


def test_mat3_inverse_identity_below_det_epsilon(k):
    # src/math.zig:482-484: |det| < 1e-6 returns identity
    o = F9()
    k["inverse"](F9(0.01, 0, 0, 0, 0.01, 0, 0, 0, 0.009), o)  # det 9e-7
    np.testing.assert_array_equal(np.array(o).reshape(3, 3), np.eye(3))


def test_integrate_rotation_constant_omega(k):
    # closed form: q(t) = exp(w t / 2) * q0 for constant w
    rng = np.random.default_rng(2)
    for _ in range(20):
        w = rng.normal(size=3) * 3
        dt = 1 / 240
        o = F4()
        k["integrate"](F4(0, 0, 0, 1), F3(*w), dt, o)
        th = np.linalg.norm(w) * dt
        ax = w / np.linalg.norm(w)
        exp = np.array([*(ax * math.sin(th / 2)), math.cos(th / 2)])
        np.testing.assert_allclose(np.array(o), exp, atol=2e-7)


# ---- src/collision_voxel.zig -------------------------------------------------

@pytest.mark.parametrize("mask,typ", [(0b000111, vxp.CORNER), (0b000011, vxp.EDGE),
                                      (0b000001, vxp.FACE), (0b000000, vxp.INSIDE)])
def test_classify_voxel(k, mask, typ):
    # src/collision_voxel.zig:273-292
    assert k["classify"](1, mask) & 7 == typ
    assert k["classify"](0, 0b111111) == 0


def test_normal_index(k):
    # src/collision_voxel.zig:324-332, plus every single-face mask maps to the
    # opposite of the empty neighbour direction order of NORMAL_LUT
    assert k["normal_index"](0b000010) == 0
    assert k["normal_index"](0b001010) == 6
    assert [k["normal_index"](1 << b) for b in range(6)] == [1, 0, 3, 2, 5, 4]
    assert k["normal_index"](0b111111) == 0  # cancels to (0,0,0): closest = first


def test_classification_table_matches_scene_classifier(k):
    # the scene generator classifies with its own table: both must agree
    for m in range(64):
        assert k["classify"](1, m) == scenes._CLASS_TABLE[m], m


def test_should_check_table(k):
    # src/collision_voxel.zig:247-271 / src/narrow_phase.zig:1724-1740
    E, Co, Ed, Fa, In = 0, 1, 2, 3, 4
    yes = {(Co, Co), (Co, Ed), (Co, Fa), (Co, In), (Ed, Co), (Fa, Co), (In, Co), (Ed, Ed)}
    for a in range(5):
        for b in range(5):
            assert bool(k["should"](a, b)) == ((a, b) in yes), (a, b)


# ---- src/narrow_phase.zig ----------------------------------------------------

def test_feature_id(k):
    # src/narrow_phase.zig:1635-1653 + the bit layout (A 16 bits/axis, B 4 bits/axis)
    f = lambda a, b: k["fid"](I3(*a), I3(*b))
    assert f((1, 2, 3), (4, 5, 6)) == f((1, 2, 3), (4, 5, 6))
    assert f((1, 2, 3), (4, 5, 6)) != f((1, 2, 4), (4, 5, 6))
    assert f((1, 2, 3), (4, 5, 6)) == 1 | 2 << 16 | 3 << 32 | 4 << 48 | 5 << 52 | 6 << 56
    assert f((-1, 0, 0), (-1, 0, 0)) == 0xFFFF | 0xF << 48


def obb(c, h, q):
    return F10(*c, *h, *q)


def test_sat_early_out(k):
    # src/narrow_phase.zig:1655-1722
    I = (0, 0, 0, 1)
    assert k["sat"](obb((0, 0, 0), (1, 1, 1), I), obb((5, 0, 0), (1, 1, 1), I)) == 1
    assert k["sat"](obb((0, 0, 0), (1, 1, 1), I), obb((1.5, 0, 0), (1, 1, 1), I)) == 0
    q45 = axis_angle(k, (0, 1, 0), math.pi / 4)
    assert k["sat"](obb((0, 0, 0), (1, 1, 1), I), obb((2, 0, 0), (1, 1, 1), q45)) == 0
    # touching faces count as overlapping (projections compared with <)
    assert k["sat"](obb((0, 0, 0), (1, 1, 1), I), obb((2, 0, 0), (1, 1, 1), I)) == 0


def _replay(k, pens, fids=None):
    n = len(pens)
    fids = fids if fids is not None else list(range(1, n + 1))
    out = (C.c_int32 * 4)()
    cnt = k["replay"](n, (C.c_float * n)(*pens), (C.c_uint64 * n)(*fids), out)
    return cnt, list(out)[:cnt]


def test_manifold_deep_contact_replaces_shallowest(k):
    # src/narrow_phase.zig:1755-1794
    cnt, idx = _replay(k, [0.0, 0.1, 0.2, 0.3, 10.0])
    assert cnt == 4 and 4 in idx and 0 not in idx


def test_manifold_keeps_top4_penetrations(k):
    # invariant of the streaming top-4: the kept penetrations are the 4 largest
    # (as a multiset), every kept one is >= every dropped one, and with distinct
    # depths the kept set is exactly the 4 deepest
    rng = np.random.default_rng(3)
    for trial in range(300):
        n = int(rng.integers(1, 30))
        pens = list(np.float32(rng.integers(0, 6, size=n)) * np.float32(0.1))
        cnt, idx = _replay(k, pens)
        assert cnt == min(n, 4)
        assert sorted(pens[i] for i in idx) == sorted(pens, reverse=True)[:4][::-1]
        dropped = [pens[i] for i in range(n) if i not in idx]
        assert not dropped or min(pens[i] for i in idx) >= max(dropped)
        distinct = list(rng.permutation(n).astype(np.float32))
        cnt, idx = _replay(k, distinct)
        assert sorted(idx) == sorted(sorted(range(n), key=lambda i: -distinct[i])[:4])


def test_manifold_same_feature_replaces_in_place(k):
    # src/narrow_phase.zig:451-461
    cnt, idx = _replay(k, [0.1, 0.2, 0.3], fids=[7, 8, 7])
    assert cnt == 2 and idx == [2, 1]


def _cache(k, cap, ops):
    n = len(ops)
    o = (C.c_uint32 * n)(*[x[0] for x in ops])
    ab = (C.c_uint32 * (2 * n))(*[v for x in ops for v in x[1]])
    xf = (C.c_float * (7 * n))(*[v for x in ops for v in x[2]])
    res = (C.c_int64 * n)()
    k["cache"](cap, n, o, ab, xf, res)
    return list(res)


GET, PUT, FRAME = 0, 1, 2
IDQ = (0.0, 0.0, 0.0, 1.0)


def test_cache_basic_operations(k):
    # src/narrow_phase.zig:1872-1901
    key = ((1, 2), (1.0, 0.0, 0.0) + IDQ)
    r = _cache(k, 64, [(GET,) + key, (PUT,) + key, (GET,) + key])
    assert r[0] == -1 and r[2] == r[1] >= 0


def test_cache_transform_sensitivity(k):
    # src/narrow_phase.zig:1931-1967: 0.1 m changes the key; positions inside the
    # same 1 cm truncation cell do not
    a = ((1, 2), (0.0, 0.0, 0.0) + IDQ)
    b = ((1, 2), (0.1, 0.0, 0.0) + IDQ)
    c = ((1, 2), (0.009, 0.0, 0.0) + IDQ)
    r = _cache(k, 64, [(PUT,) + a, (GET,) + b, (GET,) + c, (GET,) + ((2, 1), (0.0,) * 3 + IDQ)])
    assert r[1] == -1 and r[2] == r[0] and r[3] == r[0]


def test_cache_lru_eviction_window(k):
    # src/narrow_phase.zig:255-296: with capacity 4 the probe window covers the
    # whole table; the least recently used entry (first in probe order on ties)
    # is evicted
    keys = [((1, i + 2), (0.0, 0.0, 0.0) + IDQ) for i in range(5)]
    ops = [(PUT,) + keys[0], (FRAME, (0, 0), (0.0,) * 7), (PUT,) + keys[1], (PUT,) + keys[2],
           (PUT,) + keys[3], (FRAME, (0, 0), (0.0,) * 7), (GET,) + keys[1],
           (PUT,) + keys[4], (GET,) + keys[0], (GET,) + keys[1]]
    r = _cache(k, 4, ops)
    assert r[7] == r[0]          # keys[0] (frame 0) was the oldest: its slot is reused
    assert r[8] == -1 and r[9] == r[2]


def test_islands_partition(k):
    # src/islands.zig:785-816 (ids 0..4; edges 0-1, 1-2, 3-4)
    a, b = [0, 1, 3], [1, 2, 4]
    ids, lab = (C.c_uint32 * 8)(), (C.c_uint32 * 8)()
    n = k["islands"](3, (C.c_uint32 * 3)(*a), (C.c_uint32 * 3)(*b), 0, None, ids, lab)
    assert dict(zip(ids[:n], lab[:n])) == {0: 0, 1: 0, 2: 0, 3: 3, 4: 3}


def test_islands_isolated_through_static(k):
    # src/islands.zig:857-888: body 0 static; edges 0-1, 0-2, 2-3
    a, b = [0, 0, 2], [1, 2, 3]
    st = (C.c_uint8 * 4)(1, 0, 0, 0)
    ids, lab = (C.c_uint32 * 8)(), (C.c_uint32 * 8)()
    n = k["islands"](3, (C.c_uint32 * 3)(*a), (C.c_uint32 * 3)(*b), 4, st, ids, lab)
    d = dict(zip(ids[:n], lab[:n]))
    assert d[2] == d[3] and d[1] != d[2] and d[0] == 0


def test_islands_match_bfs(k):
    # textbook connected components (BFS) on random graphs
    rng = np.random.default_rng(4)
    for _ in range(20):
        nv, ne = 40, 45
        a = rng.integers(0, nv, ne).astype(np.uint32)
        b = rng.integers(0, nv, ne).astype(np.uint32)
        st = (rng.random(nv) < 0.1).astype(np.uint8)
        ids, lab = (C.c_uint32 * (2 * ne))(), (C.c_uint32 * (2 * ne))()
        n = k["islands"](ne, (C.c_uint32 * ne)(*a), (C.c_uint32 * ne)(*b), nv,
                         (C.c_uint8 * nv)(*st), ids, lab)
        adj = {}
        for x, y in zip(a, b):
            adj.setdefault(int(x), set())
            adj.setdefault(int(y), set())
            if not st[x] and not st[y]:
                adj[int(x)].add(int(y))
                adj[int(y)].add(int(x))
        want = {}
        for v in adj:
            if v in want:
                continue
            comp, todo = {v}, [v]
            while todo:
                u = todo.pop()
                for w in adj[u]:
                    if w not in comp:
                        comp.add(w)
                        todo.append(w)
            m = min(comp)
            for u in comp:
                want[u] = m
        assert dict(zip(ids[:n], lab[:n])) == want


def test_reference_coloring(k):
    # src/islands.zig:818-846: a triangle needs 3 colours, neighbours differ;
    # a star around one body saturates at 255
    a, b = [0, 1, 2], [1, 2, 0]
    out = (C.c_uint8 * 3)()
    k["refcolor"](3, (C.c_uint32 * 3)(*a), (C.c_uint32 * 3)(*b), out)
    assert sorted(out) == [0, 1, 2]
    n = 300
    out = (C.c_uint8 * n)()
    k["refcolor"](n, (C.c_uint32 * n)(*[1] * n), (C.c_uint32 * n)(*range(2, n + 2)), out)
    assert list(out[:255]) == list(range(255)) and set(out[255:]) == {255}


# ---- src/solver.zig ------------------------------------------------------------

def test_soft_params_closed_form(k):
    # src/solver.zig:57-72 and 1007-1023
    o = F3()
    h, z, dt = 30.0, 1.0, 1.0 / 240.0
    k["soft"](h, z, dt, o)
    om = 2 * math.pi * h
    a1 = 2 * z + om * dt
    a2 = dt * om * a1
    a3 = 1 / (1 + a2)
    np.testing.assert_allclose(np.array(o), (om / a1, a2 * a3, a3), rtol=2e-6)
    assert 0 < o[1] < 1 and 0 < o[2] < 1
    k["soft"](0.0, 1.0, 1.0 / 60.0, o)
    assert list(o) == [0.0, 1.0, 1.0]


def test_compute_orthogonal(k):
    # src/solver.zig:1025-1043
    o = F3()
    for v in ((1, 0, 0), (0, 1, 0), (0, 0, 1), tuple(np.ones(3) / math.sqrt(3))):
        k["orth"](F3(*v), o)
        assert abs(np.dot(v, np.array(o))) < 1e-6 and abs(np.linalg.norm(o) - 1) < 1e-6


# ---- src/body.zig / src/collider.zig --------------------------------------------

def _system(lib, **over):
    cfg = vxp.default_config(**over)
    p = C.c_void_p()
    lib.check(lib.system_create(C.byref(cfg), 1, C.byref(p)), "system_create")
    return p


def _collider(lib, p, vox, origin, vs):
    bp, v = scenes.grid_to_bricks(vox, origin)
    out = C.c_uint32()
    lib.check(lib.collider_create(p, vs, len(bp), bp.ctypes.data_as(C.POINTER(C.c_int32)),
                                  v.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(out)), "collider")
    return out.value


def test_mass_properties_cube_lattice(k, oracle):
    # closed form for an n^3 lattice of point masses s^3*rho about its centroid;
    # the COM carries the reference's extra half voxel (src/collider.zig:395)
    p = _system(oracle)
    for n, s in ((3, 0.2), (4, 0.1), (5, 0.25)):
        vox, org = scenes.box_shape(n, n, n)
        cid = _collider(oracle, p, vox, org, s)
        out = (C.c_float * 13)()
        assert k["mass"](p, cid, 1.0, out) == 0
        coords = (np.arange(n) + org[0]) * s + s  # voxelToLocal + half voxel
        com = coords.mean()
        m_v = s ** 3
        r2 = ((coords - com) ** 2).sum() * n * n  # sum over the lattice of one axis^2
        I = m_v * 2 * r2
        np.testing.assert_allclose(out[0], 1 / (m_v * n ** 3), rtol=1e-6)
        inv = np.array(out[1:10]).reshape(3, 3)
        if I ** 3 >= 1e-6:
            np.testing.assert_allclose(inv, np.eye(3) / I, rtol=1e-4, atol=1e-3)
        else:
            np.testing.assert_array_equal(inv, np.eye(3))
        np.testing.assert_allclose(out[10:13], [com] * 3, rtol=1e-6)
    oracle.system_destroy(p)


def test_collider_metadata_and_brick_hash_aliasing(k, oracle):
    # VoxelCollider brick positions hash with hashBrickPos (src/collider.zig:187-192);
    # a 64x64 one-voxel slab has 64 brick positions but only 40 distinct hashes, so
    # the reference merges aliased bricks and only the first position of each
    # hash class counts for the bounds.
    p = _system(oracle)
    vox, org = scenes.hand_ground(64, 64)
    cid = _collider(oracle, p, vox, org, 0.2)
    b, a, solid = (C.c_int32 * 6)(), (C.c_float * 6)(), C.c_uint32()
    k["info"](p, cid, b, a, C.byref(solid))
    bp, _ = scenes.grid_to_bricks(vox, org)
    M = (1 << 64) - 1
    h = lambda x, y, z: (x & M) ^ ((y & M) * 0x517cc1b727220a95 & M) ^ ((z & M) * 0x9e3779b97f4a7c15 & M)
    first = {}
    for x, y, z in bp.tolist():
        first.setdefault(h(x, y, z), (x, y, z))
    assert len(first) == 40
    assert solid.value == 40 * 64
    xs = [f[0] for f in first.values()]
    zs = [f[2] for f in first.values()]
    assert list(b) == [min(xs) * 8, 0, min(zs) * 8, max(xs) * 8 + 8, 1, max(zs) * 8 + 8]
    oracle.system_destroy(p)


def test_detect_resting_cube_corners_on_face(k, oracle):
    # analytic answer: an axis-aligned 3^3 README cube one voxel above the face
    # of a slab touches only with its 4 bottom corners (corner x face is the only
    # passing type pair), depth s - dist = 0 at exactly one voxel spacing
    p = _system(oracle)
    g = _collider(oracle, p, *scenes.hand_ground(16, 16), 0.2)
    c = _collider(oracle, p, *scenes.readme_cube(), 0.2)
    m = vxp._abi.Manifold()
    vc, cand = C.c_uint32(), C.c_uint32()
    n = k["detect"](p, g, (C.c_float * 7)(0, 0, 0, 0, 0, 0, 1), c,
                    (C.c_float * 7)(0.0, 0.4, 0.0, 0, 0, 0, 1), C.byref(m), C.byref(vc), C.byref(cand))
    assert n == 4 and cand.value == 4
    for i in range(4):
        ct = m.contacts[i]
        assert abs(ct.penetration) < 1e-6
        assert list(ct.normal) == [0.0, 1.0, 0.0]
        assert abs(ct.point[1] - 0.2) < 1e-6
    oracle.system_destroy(p)


# ---- src/system.zig ----------------------------------------------------------------

def _sys(lib, **over):
    return vxp.VoxelPhysicsSystem(lib=lib, **over)


def test_free_fall_damping_recurrence(oracle):
    # closed-form recurrence of applyGravity + integrate for a lone body
    # (src/solver.zig:564-605, 341-356), evaluated in float64
    s = _sys(oracle, max_bodies=4)
    cid = s.addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                        voxels=np.full((1, 512), vxp.INSIDE, np.uint8))
    h = s.createDynamicBody((0, 50, 0), (0, 0, 0, 1), cid, 1.0)
    v, y, hs = 0.0, 50.0, 1 / 60 / 4
    for f in range(40):
        s.update(1 / 60)
        for _ in range(4):
            v = v - 9.81 * hs
            d = 0.3 + (5 * (1 - abs(v) / 0.5) if abs(v) < 0.5 else 0.0)
            v = v * max(0.0, 1 - d * hs)
            if v * v < 1e-6:
                v = 0.0
            y = y + v * hs
        if v * v < 0.002:
            v = 0.0
    b = s.bodies()[h.index]
    assert b["linear_velocity"][1] == pytest.approx(v, rel=1e-5)
    assert b["position"][1] == pytest.approx(y, rel=1e-5)


def test_sleep_after_half_second(oracle):
    # src/body.zig:570-586: a still body sleeps once its timer reaches 0.5 s
    s = _sys(oracle, max_bodies=4, gravity=(0, 0, 0))
    cid = s.addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                        voxels=np.full((1, 512), vxp.INSIDE, np.uint8))
    h = s.createDynamicBody((0, 5, 0), (0, 0, 0, 1), cid, 1.0)
    for f in range(29):
        s.update(1 / 60)
    assert not s.isSleeping(h)
    s.update(1 / 60)
    assert s.isSleeping(h)
    st = s.stats()
    assert st["sleeping_bodies"] == 1 and st["active_bodies"] == 0
    s.update(1 / 60)  # every dynamic body sleeps: step() returns early
    assert s.stats()["steps_executed"] == 0


def test_writeback_guards(oracle):
    # src/solver.zig:246-296: |v| clamped to 50, NaN pose reset to (0,2,0)
    s = _sys(oracle, max_bodies=4, gravity=(0, 0, 0))
    cid = s.addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                        voxels=np.full((1, 512), vxp.INSIDE, np.uint8))
    h = s.createDynamicBody((0, 5, 0), (0, 0, 0, 1), cid, 1.0)
    s.setState(h, linear_velocity=(300, 0, 0))
    s.update(1 / 60)
    v = s.bodies()[h.index]["linear_velocity"]
    assert abs(np.linalg.norm(v) - 50.0) < 1e-3
    s.setState(h, position=(float("nan"), 0, 0))
    s.update(1 / 60)
    b = s.bodies()[h.index]
    assert list(b["position"]) == [0.0, 2.0, 0.0] and list(b["rotation"]) == [0, 0, 0, 1]


def test_spiral_of_death_clamp(oracle):
    # src/system.zig:516-526
    s = _sys(oracle, max_bodies=4)
    cid = s.addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                        voxels=np.full((1, 512), vxp.INSIDE, np.uint8))
    s.createDynamicBody((0, 5, 0), (0, 0, 0, 1), cid, 1.0)
    s.update(1.0)
    assert s.stats()["substeps_this_frame"] == 8
    s.update(0.0)
    assert s.stats()["accumulated_time"] == pytest.approx(1 / 60)


def test_max_bodies_and_handles(oracle):
    # src/system.zig:475-499, 1107-1137
    s = _sys(oracle, max_bodies=2)
    cid = s.addCollider(voxel_size=0.2, brick_pos=np.zeros((1, 3), np.int32),
                        voxels=np.full((1, 512), vxp.INSIDE, np.uint8))
    h1 = s.createDynamicBody((0, 0, 0), (0, 0, 0, 1), cid, 1.0)
    s.createDynamicBody((0, 0, 0), (0, 0, 0, 1), cid, 1.0)
    with pytest.raises(vxp.VxpError):
        s.createDynamicBody((0, 0, 0), (0, 0, 0, 1), cid, 1.0)
    assert s.destroyBody(h1) == 0
    assert s.destroyBody(h1) == -1
    h3 = s.createDynamicBody((0, 0, 0), (0, 0, 0, 1), cid, 1.0)
    assert h3.index == h1.index and h3.generation != h1.generation


def test_threaded_narrow_phase_equals_sequential(oracle):
    # the thread-pooled CPU baseline replays the cache in pair order, so it must
    # equal the literal sequential processPair order bit for bit
    from vxp import compare
    sc = scenes.c0_small(frames=120)
    f, msg, _ = compare.lockstep(sc, oracle, oracle, 120, cfg_over=None, check_every=20) \
        if False else (None, None, None)
    a = vxp.from_scene(sc, lib=oracle, cfg_over=dict(cpu_threads=1))
    b = vxp.from_scene(sc, lib=oracle, cfg_over=dict(cpu_threads=4))
    for fr in range(120):
        a.update(sc.dt)
        b.update(sc.dt)
    assert compare.diff_manifolds(a.manifolds(), b.manifolds()) is None
    assert compare.diff_bodies(a.bodies(), b.bodies()) is None
    assert compare.diff_stats(a.stats(), b.stats()) is None


def test_solve_order_drift_is_small(oracle):
    # colour-order vs pair-key-order Gauss-Seidel: a documented deviation of the
    # visiting order only; after 60 frames positions stay close
    sc = scenes.c0_small(frames=60)
    a = vxp.from_scene(sc, lib=oracle, cfg_over=dict(solve_order=vxp.SOLVE_ORDER_COLOR))
    b = vxp.from_scene(sc, lib=oracle, cfg_over=dict(solve_order=vxp.SOLVE_ORDER_CANONICAL))
    for fr in range(60):
        a.update(sc.dt)
        b.update(sc.dt)
    d = np.abs(a.bodies()["position"] - b.bodies()["position"]).max()
    assert d < 0.05, d


# ---- src/collider.zig:466-486: brick position and local coordinate -----------
def test_voxel_to_brick_and_local(k):
    def conv(v):
        b, l = I3(), I3()
        k["brick"](I3(*v), b, l)
        return tuple(b), tuple(l)
    assert conv((15, 20, 25))[0] == (1, 2, 3)
    assert conv((-1, -10, -20))[0] == (-1, -2, -3)
    assert conv((13, 18, 23))[1] == (5, 2, 7)
    # floor division and modulo agree with Python's on a sweep around zero
    for x in range(-40, 41):
        b, l = conv((x, -x, x * 3))
        assert b == (x // 8, (-x) // 8, (x * 3) // 8) and l == (x % 8, (-x) % 8, (x * 3) % 8)
    # the host-side collider builder maps voxels the same way
    col = vxp.VoxelCollider(0.1)
    col.setCollisionVoxel(-1, -10, -20, vxp.voxel_byte(vxp.FACE, 2))
    pos, vox = col.arrays()
    assert tuple(pos[0]) == (-1, -2, -3)
    assert vox[0][7 + 8 * 6 + 64 * 4] != 0


# ---- src/narrow_phase.zig:1796-1825: ContactManager warm starting -------------
def test_warm_start_carries_impulse(k):
    out = (C.c_float * 12)()
    k["warm"](1, (C.c_uint64 * 1)(12345), 1, (C.c_uint64 * 1)(12345),
              (C.c_float * 3)(50.0, 1.5, -2.0), out)
    assert list(out)[:3] == [50.0, 1.5, -2.0]
    # unmatched feature ids start from zero; the first matching old contact wins
    k["warm"](3, (C.c_uint64 * 3)(7, 12345, 9), 3, (C.c_uint64 * 3)(9, 12345, 9),
              (C.c_float * 9)(1, 2, 3, 4, 5, 6, 7, 8, 9), out)
    assert list(out)[:9] == [0, 0, 0, 4, 5, 6, 1, 2, 3]


# ---- src/system.zig:1139-1158: one step of a body with an empty collider ------
def test_system_step_empty_collider(oracle):
    s = vxp.VoxelPhysicsSystem(lib=oracle)
    cid = s.addCollider(voxel_size=0.1, brick_pos=np.zeros((0, 3), np.int32),
                        voxels=np.zeros((0, 512), np.uint8))
    h = s.createDynamicBody((0.0, 5.0, 0.0), (0.0, 0.0, 0.0, 1.0), cid, 1.0)
    s.step(1.0 / 60.0)
    assert s.getPosition(h)[1] < 5.0
    s.close()
