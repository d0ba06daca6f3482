"""Pins of the CPU restatement's broad-phase, narrow-phase and solver functions
that the reference's own inline tests leave open.  Each test names the rule it
pins and the slip it would catch; the expected values come from geometry,
closed forms, brute force or invariants computed here in float64 (numpy), not
from the restatement.

Rules (reference files):
- fat AABB: src/broad_phase.zig:150-212 (inflate 0.1 on insert, the +0.4
  containment skip, one-sided extension by 2 v)
- contacts: src/narrow_phase.zig:644-698 (overlap region), 701-709 (type
  filter), 796-857 (neighbourhood probes, voxels_checked), 860-922
  (generateContact: midpoint, penetration, normal rules, anchors)
- solver: src/solver.zig:608-699 (effective masses, bias), 732-838 (normal and
  friction rows: clamps), 860-923 (position NGS), and the resting balance of
  the whole TGS-soft step
- solver colouring: greedy colours in ascending (prio(key), key) order over
  manifolds that share an awake dynamic body (the GPU's colour order)
"""
import ctypes as C
import math

import numpy as np
import pytest

import vxp
from vxp import _abi, scenes

# NORMAL_LUT of src/collision_voxel.zig:36-67 (a table kept alike: voxel bytes
# index it)
LUT = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
       (1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0), (1, 0, 1), (1, 0, -1), (-1, 0, 1),
       (-1, 0, -1), (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1),
       (1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1), (-1, 1, 1), (-1, 1, -1), (-1, -1, 1),
       (-1, -1, -1)]
IDENT = (0.0, 0.0, 0.0, 1.0)
S = 0.2  # voxel size of the contact cases: contact if centre distance < S + 0.02


def lut(i):
    v = np.array(LUT[i], float)
    return v / np.linalg.norm(v)


def qmat(q):
    x, y, z, w = np.asarray(q, float) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def axis_angle(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    return tuple(np.append(a * math.sin(ang / 2), math.cos(ang / 2)))


@pytest.fixture(scope="module")
def det(oracle):
    fn = oracle.dll.vxo_detect_pair
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_uint32,
                   C.POINTER(C.c_float), C.POINTER(_abi.Manifold), C.POINTER(C.c_uint32),
                   C.POINTER(C.c_uint32)]
    return fn


class Pair:
    """Two colliders in one restatement system and detectVoxelCollisions between
    them at given poses (pos xyz + quaternion xyzw)."""

    def __init__(self, oracle, det, a, b):
        self.s = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=4)
        self.det = det
        self.ca = self.s.addCollider(voxel_size=a[0], brick_pos=a[1], voxels=a[2])
        self.cb = self.s.addCollider(voxel_size=b[0], brick_pos=b[1], voxels=b[2])

    def run(self, pa, qa, pb, qb):
        m, vc, cand = _abi.Manifold(), C.c_uint32(), C.c_uint32()
        xa = (C.c_float * 7)(*pa, *qa)
        xb = (C.c_float * 7)(*pb, *qb)
        n = self.det(self.s.ptr, self.ca, xa, self.cb, xb, C.byref(m), C.byref(vc), C.byref(cand))
        return n, m, vc.value, cand.value


def one_voxel(vtype, normal, vs=S):
    vox = np.zeros((1, 1, 1), np.uint8)
    vox[0, 0, 0] = vxp.voxel_byte(vtype, normal)
    bp, vx = scenes.grid_to_bricks(vox, (0, 0, 0))
    return vs, bp, vx


def centre(pos, q, coord=(0, 0, 0), vs=S):
    # voxelToLocal (voxel centre) through transformPoint
    return qmat(q) @ ((np.asarray(coord, float) + 0.5) * vs) + np.asarray(pos, float)


def place(q, target, coord=(0, 0, 0), vs=S):
    """Body position that puts voxel `coord`'s centre at `target`."""
    return tuple(np.asarray(target, float) - qmat(q) @ ((np.asarray(coord, float) + 0.5) * vs))


def contact(m, i=0):
    c = m.contacts[i]
    return (np.array(c.point, float), np.array(c.normal, float), float(c.penetration),
            np.array(c.local_anchor_a, float), np.array(c.local_anchor_b, float))


def single(oracle, det, ta, na, tb, nb, qa, qb, offset):
    """A's voxel centre at (1, 2, 3), B's voxel centre at (1, 2, 3) + offset."""
    p = Pair(oracle, det, one_voxel(ta, na), one_voxel(tb, nb))
    ca = np.array([1.0, 2.0, 3.0])
    pa = place(qa, ca)
    pb = place(qb, ca + np.asarray(offset, float))
    return p.run(pa, qa, pb, qb), (pa, pb)


# ---- generateContact normal and penetration rules (src/narrow_phase.zig:886-906)

def test_face_voxel_of_a_gives_its_rotated_lut_normal(oracle, det):
    # a face voxel of A wins over the direction (slip: the direction or B's normal)
    qa = axis_angle((0, 0, 1), math.pi / 2)
    (n, m, vc, cand), _ = single(oracle, det, vxp.FACE, 0, vxp.CORNER, 1, qa, IDENT, (0.03, 0.12, 0.05))
    assert n == 1 and vc == 1 and cand == 1
    _, normal, pen, _, _ = contact(m)
    np.testing.assert_allclose(normal, qmat(qa) @ lut(0), atol=1e-6)
    np.testing.assert_allclose(pen, S - np.linalg.norm((0.03, 0.12, 0.05)), atol=1e-6)


def test_face_voxel_of_b_gives_its_negated_lut_normal(oracle, det):
    qb = axis_angle((1, 0, 0), 0.7)
    (n, m, _, _), _ = single(oracle, det, vxp.CORNER, 0, vxp.FACE, 3, IDENT, qb, (0.02, -0.15, 0.01))
    assert n == 1
    np.testing.assert_allclose(contact(m)[1], -(qmat(qb) @ lut(3)), atol=1e-6)


@pytest.mark.parametrize("na,nb,want", [
    (1, 0, "dir"),    # neither LUT normal within 0.3 of the direction: diff / dist
    (6, 0, "a"),      # A's normal (1,1,0)/sqrt2: dot 0.82 > 0.3, A's normal (not the direction)
    (1, 9, "-b"),     # -B's normal (1,1,0)/sqrt2 likewise
    (1, 24, "-b"),    # -B's normal (1,1,-1)/sqrt3: dot 1/3 > 0.3
    (6, 24, "a"),     # both pass: A's is tested first
    (7, 0, "dir"),    # A's normal (1,-1,0)/sqrt2 . (1,1,1)/sqrt3 = 0 <= 0.3
])
def test_corner_pairs_choose_direction_or_lut_normal(oracle, det, na, nb, want):
    off = np.array([0.1, 0.1, 0.1])
    (n, m, _, _), _ = single(oracle, det, vxp.CORNER, na, vxp.CORNER, nb, IDENT, IDENT, off)
    assert n == 1
    normal = contact(m)[1]
    expect = {"dir": off / np.linalg.norm(off), "a": lut(na), "-b": -lut(nb)}[want]
    np.testing.assert_allclose(normal, expect, atol=1e-6)


def test_coincident_centres_use_a_normal(oracle, det):
    # dist <= MIN_AXIS_LENGTH (0.001): A's normal, penetration = voxel size
    (n, m, _, _), _ = single(oracle, det, vxp.CORNER, 4, vxp.CORNER, 0, IDENT, IDENT, (0.0, 0.0, 0.0))
    assert n == 1
    _, normal, pen, _, _ = contact(m)
    np.testing.assert_allclose(normal, lut(4), atol=0)
    assert pen == pytest.approx(S, abs=1e-7)


@pytest.mark.parametrize("ta,tb", [(vxp.CORNER, vxp.INSIDE), (vxp.INSIDE, vxp.CORNER)])
def test_inside_voxel_penetration_is_at_least_voxel_size(oracle, det, ta, tb):
    # pen = max(s - dist, s) when either voxel is inside (slip: s - dist)
    (n, m, _, _), _ = single(oracle, det, ta, 1, tb, 0, IDENT, IDENT, (0.12, 0.0, 0.05))
    assert n == 1
    assert contact(m)[2] == pytest.approx(S, abs=1e-7)


@pytest.mark.parametrize("ta,tb,hit", [
    (vxp.EDGE, vxp.EDGE, True), (vxp.EDGE, vxp.FACE, False), (vxp.FACE, vxp.FACE, False),
    (vxp.FACE, vxp.INSIDE, False), (vxp.CORNER, vxp.FACE, True), (vxp.INSIDE, vxp.CORNER, True)])
def test_type_filter_counts_the_probe_but_not_the_contact(oracle, det, ta, tb, hit):
    # shouldCheckCollision (src/narrow_phase.zig:701-709); voxels_checked counts
    # every non-empty B voxel probed, before the filter (src/narrow_phase.zig:825)
    qb = axis_angle((0, 0, 1), math.pi / 4)  # edge x edge at 45 degrees
    (n, m, vc, cand), _ = single(oracle, det, ta, 6, tb, 0, IDENT, qb, (0.1, 0.1, 0.0))
    assert vc == 1
    assert (n == 1) == hit and cand == (1 if hit else 0)


def test_out_of_range_pair_makes_no_contact(oracle, det):
    # speculative distance: dist^2 < s^2 + 0.02^2 + 2 s 0.02, i.e. dist < s + 0.02
    for d, hit in ((0.2195, True), (0.2205, False)):
        (n, _, vc, _), _ = single(oracle, det, vxp.CORNER, 0, vxp.CORNER, 0, IDENT, IDENT, (d, 0.0, 0.0))
        assert vc == 1 and (n == 1) == hit, d


def test_contact_point_anchors_and_depth_on_random_poses(oracle, det):
    # midpoint of the two voxel centres; anchors = inverseTransformPoint of it
    # in each body; penetration = s - |diff|; feature id of the voxel pair
    rng = np.random.default_rng(7)
    fid = oracle.dll.vxo_feature_id
    fid.restype = C.c_uint64
    fid.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    for _ in range(20):
        qa = tuple(rng.normal(size=4))
        qb = tuple(rng.normal(size=4))
        qa = tuple(np.array(qa) / np.linalg.norm(qa))
        qb = tuple(np.array(qb) / np.linalg.norm(qb))
        off = rng.normal(size=3)
        off *= rng.uniform(0.05, 0.19) / np.linalg.norm(off)
        (n, m, _, _), (pa, pb) = single(oracle, det, vxp.CORNER, 0, vxp.CORNER, 0, qa, qb, off)
        assert n == 1
        point, _, pen, la, lb = contact(m)
        ca, cb = centre(pa, qa), centre(pb, qb)
        mid = 0.5 * (ca + cb)
        np.testing.assert_allclose(point, mid, atol=2e-6)
        np.testing.assert_allclose(la, qmat(qa).T @ (mid - np.array(pa)), atol=2e-6)
        np.testing.assert_allclose(lb, qmat(qb).T @ (mid - np.array(pb)), atol=2e-6)
        assert pen == pytest.approx(S - np.linalg.norm(cb - ca), abs=2e-6)
        assert m.contacts[0].feature_id == fid((C.c_int32 * 3)(0, 0, 0), (C.c_int32 * 3)(0, 0, 0))


# ---- overlap region and voxels_checked (src/narrow_phase.zig:644-698,722-857)

def _dense(vs, bp, vx):
    """Non-empty voxels of a collider as integer coordinates plus a lookup."""
    vox = {}
    for (bx, by, bz), bytes_ in zip(bp, vx):
        for i in np.nonzero(bytes_)[0]:
            x, y, z = i % 8, (i // 8) % 8, i // 64
            vox[(bx * 8 + x, by * 8 + y, bz * 8 + z)] = int(bytes_[i])
    return vox


def _region(vs_a, vox_a, pa, qa, vs_b, vox_b, pb, qb):
    """getOverlapRegion from the geometry: the corners of B's local box in A's
    frame, widened by one A voxel, clipped to A's box, to voxel indices."""
    kb = np.array(list(vox_b.keys()))
    lo_b, hi_b = kb.min(0) * vs_b, (kb.max(0) + 1) * vs_b
    corners = np.array([[x, y, z] for x in (lo_b[0], hi_b[0]) for y in (lo_b[1], hi_b[1])
                        for z in (lo_b[2], hi_b[2])])
    world = corners @ qmat(qb).T + np.array(pb)
    rel = (world - np.array(pa)) @ qmat(qa)
    ka = np.array(list(vox_a.keys()))
    lo_a, hi_a = ka.min(0) * vs_a, (ka.max(0) + 1) * vs_a
    tmin = np.clip(rel.min(0) - vs_a, lo_a, hi_a)
    tmax = np.clip(rel.max(0) + vs_a, lo_a, hi_a)
    if np.any(tmin >= tmax):
        return None
    mn = np.maximum(np.floor(tmin / vs_a).astype(int), ka.min(0))
    mx = np.minimum(np.floor(tmax / vs_a).astype(int), ka.max(0) + 1)
    return mn, mx, (tmin, tmax)


def _brute_checked(vs_a, vox_a, pa, qa, vs_b, vox_b, pb, qb, region):
    mn, mx, _ = region
    kb = np.array(list(vox_b.keys()))
    bmin, bmax = kb.min(0), kb.max(0) + 1
    n, margin = 0, 1.0
    for (x, y, z) in vox_a:
        if not (mn[0] <= x <= mx[0] and mn[1] <= y <= mx[1] and mn[2] <= z <= mx[2]):
            continue
        w = qmat(qa) @ ((np.array([x, y, z]) + 0.5) * vs_a) + np.array(pa)
        lb = qmat(qb).T @ (w - np.array(pb)) / vs_b
        frac = lb - np.floor(lb)
        margin = min(margin, float(np.min(np.minimum(frac, 1 - frac))))
        c = np.floor(lb).astype(int)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    q = (c[0] + dx, c[1] + dy, c[2] + dz)
                    if all(bmin[k] <= q[k] < bmax[k] for k in range(3)) and q in vox_b:
                        n += 1
    return n, margin


def test_voxels_checked_equals_brute_force_for_rotated_cubes_over_a_slab(oracle, det):
    """A = 16x1x16 slab, B = the README 3^3 cube at random orientations resting
    across it: voxels_checked = the number of non-empty B voxels in the 3x3x3
    neighbourhoods of the slab voxels inside the overlap region, counted here
    voxel by voxel (slips caught: region not widened by a voxel, rotation by q
    instead of its conjugate, exclusive instead of inclusive region loop,
    probes outside B's bounds counted)."""
    gvox, gorg = scenes.hand_ground(16, 16)
    ga = scenes.grid_to_bricks(gvox, gorg)
    cvox, corg = scenes.readme_cube()
    cb = scenes.grid_to_bricks(cvox, corg)
    p = Pair(oracle, det, (S, *ga), (S, *cb))
    vox_a, vox_b = _dense(S, *ga), _dense(S, *cb)
    rng = np.random.default_rng(3)
    pa, qa = (0.0, -0.1, 0.0), axis_angle((0.1, 1.0, -0.05), 0.5)  # slab yawed and tilted
    done = 0
    for _ in range(100):
        if done == 12:
            break
        qb = rng.normal(size=4)
        qb = tuple(qb / np.linalg.norm(qb))
        pb = (rng.uniform(-0.6, 0.6), rng.uniform(0.15, 0.35), rng.uniform(-0.6, 0.6))
        reg = _region(S, vox_a, pa, qa, S, vox_b, pb, qb)
        if reg is None:
            continue
        want, margin = _brute_checked(S, vox_a, pa, qa, S, vox_b, pb, qb, reg)
        if margin < 1e-3:  # a centre on a cell boundary: float32 may floor differently
            continue
        n, m, vc, cand = p.run(pa, qa, pb, qb)
        assert vc == want, (qb, pb)
        done += 1
    assert done == 12


# ---- fat AABB (src/broad_phase.zig:150-212, src/system.zig:570-583) ---------

def _leaf_box(s, body_id):
    nodes, _ = s.tree(1)
    for n in nodes:
        if n["height"] == 0 and n["body_id"] == body_id:
            return np.array(n["min"], np.float32), np.array(n["max"], np.float32)
    raise AssertionError("leaf not found")


def test_fat_aabb_insert_skip_and_velocity_extension(oracle):
    """The dynamic tree's leaf box is the body's fat AABB (tree pair order).
    Insert: world AABB inflated by 0.1.  Update: skipped while the box is
    contained and the fat box lies inside the AABB inflated by 0.5; otherwise
    the AABB inflated by 0.1 with the side in the direction of v moved by 2 v
    (slips caught: extension on both sides, by v instead of 2 v, the skip
    ignored, a 0.4 margin in place of 0.1 * 4 around the fat box)."""
    f32 = np.float32
    s = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=4, pair_order=_abi.PAIR_ORDER_TREE,
                               gravity=(0.0, 0.0, 0.0), enable_sleeping=0)
    cvox, corg = scenes.readme_cube()
    c = s.addCollider(voxel_size=S, **dict(zip(("brick_pos", "voxels"), scenes.grid_to_bricks(cvox, corg))))
    h = s.createDynamicBody((0.3, 2.0, -0.2), axis_angle((1, 2, 3), 0.4), c, 1.0)
    bid = int(s.bodies()["id"][h.index])

    def aabb():
        lo_, hi_ = s.getWorldAABB(h)
        return np.array(lo_, np.float32), np.array(hi_, np.float32)

    lo, hi = aabb()
    flo, fhi = _leaf_box(s, bid)
    np.testing.assert_array_equal(flo, lo - f32(0.1))
    np.testing.assert_array_equal(fhi, hi + f32(0.1))
    # box still contained and the fat box within AABB + 0.5: the update is
    # skipped, so the velocity does not extend the box
    s.setState(h, linear_velocity=(0.3, 0.0, 0.0))
    s.update(1 / 60)
    np.testing.assert_array_equal(_leaf_box(s, bid)[0], flo)
    np.testing.assert_array_equal(_leaf_box(s, bid)[1], fhi)
    # moved by 0.15 in x (out of the fat box): refit with 2 v extension
    v = np.array([1.5, -2.0, 0.25], np.float32)
    s.setState(h, position=(0.45, 2.0, -0.2), linear_velocity=tuple(float(x) for x in v))
    lo, hi = aabb()  # the pose the broad phase sees (before the solve moves it)
    s.update(1 / 60)
    flo, fhi = _leaf_box(s, bid)
    d = v * f32(2.0)
    want_lo = lo - f32(0.1)
    want_hi = hi + f32(0.1)
    for k in range(3):
        if d[k] < 0:
            want_lo[k] = want_lo[k] + d[k]
        else:
            want_hi[k] = want_hi[k] + d[k]
    np.testing.assert_array_equal(flo, want_lo)
    np.testing.assert_array_equal(fhi, want_hi)
    # box contained in the (large) fat box, fat box not within AABB + 0.5:
    # refit to the tight fat box
    s.setState(h, linear_velocity=(0.0, 0.0, 0.0))
    lo, hi = aabb()
    assert np.all(lo >= flo) and np.all(hi <= fhi)
    s.update(1 / 60)
    flo, fhi = _leaf_box(s, bid)
    np.testing.assert_array_equal(flo, lo - f32(0.1))
    np.testing.assert_array_equal(fhi, hi + f32(0.1))
    s.close()


# ---- solver (src/solver.zig:608-699,732-838,860-923) -------------------------

def _probe_system(oracle, d, vs=0.5, im=3.0, inv_i=None, qb=IDENT, gravity=(0.0, 0.0, 0.0),
                  hertz=0.0, **cfg):
    """A static 3x1x3 slab of face voxels (normal +y) and one dynamic body of a
    single corner voxel whose centre is `vs - d` above the centre of the slab's
    middle voxel: exactly one contact (the neighbours are farther than vs + 0.02),
    penetration d, normal +y.  No gravity; hertz 0 (the velocity rows add no
    bias at rest), so only the position passes move the body."""
    s = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=4, gravity=gravity, hertz=hertz, **cfg)
    slab = np.zeros((3, 1, 3), np.uint8)
    slab[:, 0, :] = vxp.voxel_byte(vxp.FACE, 2)
    g = s.addCollider(voxel_size=vs, **dict(zip(("brick_pos", "voxels"),
                                                 scenes.grid_to_bricks(slab, (-1, 0, -1)))))
    s.createStaticBody((0.0, 0.0, 0.0), IDENT, g)
    cv = np.full((1, 1, 1), vxp.voxel_byte(vxp.CORNER, 3), np.uint8)
    c = s.addCollider(voxel_size=vs, **dict(zip(("brick_pos", "voxels"), scenes.grid_to_bricks(cv, (0, 0, 0)))))
    target = np.array([0.5 * vs, 0.5 * vs + vs - d, 0.5 * vs])
    pb = place(qb, target, vs=vs)
    if inv_i is None:
        inv_i = np.array([[2.0, 0.3, 0.1], [0.3, 1.5, -0.2], [0.1, -0.2, 2.5]])
    h = s.createDynamicBodyWithMass(pb, qb, c, im, tuple(np.asarray(inv_i, float).ravel()))
    return s, h, np.array(pb), np.asarray(inv_i, float)


def _k_normal(point, pb, qb, im, inv_i, n):
    """im + (r x n) . I_world^-1 (r x n) with the reference's world inverse
    inertia M I^-1 M^T, M = Quat.toMat3 (src/math.zig:350-371, src/body.zig
    getWorldInvInertia): toMat3's rows are the columns of the rotation that
    Quat.rotate applies, so I_world^-1 = R^T I^-1 R for that rotation R."""
    r = point - pb
    M = qmat(qb).T
    rn = np.cross(r, n)
    return im + rn @ (M @ inv_i @ M.T) @ rn


@pytest.mark.parametrize("d,capped", [(0.05, False), (0.3, True)])
def test_position_pass_closed_form(oracle, d, capped):
    """One position pass on one contact: correction = max(0.2 (sep + slop),
    -0.04) with sep = -d, impulse -correction / k along n, k = m_a^-1 + m_b^-1
    + (r x n) . I_world^-1 (r x n); the body moves by impulse / m (slips caught:
    slop sign, the 0.04 cap, a dropped angular term, the world inertia rotated
    the textbook way instead of the reference's)."""
    qb = axis_angle((0.3, 1.0, 0.2), 0.6)
    s, h, pb, inv_i = _probe_system(oracle, d, qb=qb, position_iterations=1)
    s.update(1 / 60)
    m = s.manifolds()
    assert len(m) == 1 and m[0].contact_count == 1
    point = np.array(m[0].contacts[0].point, float)
    n = np.array([0.0, 1.0, 0.0])
    k = _k_normal(point, pb, qb, 3.0, inv_i, n)
    corr = max(0.2 * (-d + 0.005), -0.04)
    assert (corr == -0.04) == capped
    y = s.bodies()[h.index]["position"]
    np.testing.assert_allclose(y, pb + n * (-corr / k) * 3.0, atol=2e-6)
    st = s.stats()
    assert st["position_iterations_run"] == 1
    assert st["max_penetration"] == pytest.approx(d, abs=1e-6)
    s.close()


@pytest.mark.parametrize("d,passes", [(0.004, 1), (0.012, 1), (0.05, 2)])
def test_position_pass_slop_and_early_exit(oracle, d, passes):
    """Below the slop (0.005) a contact is not corrected; the passes stop after
    the first whose minimum separation is >= -3 slop (src/solver.zig:551-552,
    922): 0.012 stops after one pass, 0.05 runs both."""
    s, h, pb, _ = _probe_system(oracle, d, position_iterations=2)
    s.update(1 / 60)
    assert s.stats()["position_iterations_run"] == passes
    y = s.bodies()[h.index]["position"]
    if d < 0.005:
        np.testing.assert_array_equal(y, pb.astype(np.float32))
    else:
        assert y[1] > pb[1]
    s.close()


def test_effective_masses_closed_form(oracle):
    """prepareConstraints' effective masses (src/solver.zig:630-656): normal
    1 / k, tangents along computeOrthogonal(n) and n x t1, with a full
    (non-diagonal) inverse inertia rotated to world (slips caught: transposed
    rotation, a dropped body term, the cross product in the wrong order)."""
    qb = axis_angle((1.0, -0.4, 0.7), 1.1)
    s, h, pb, inv_i = _probe_system(oracle, 0.004, qb=qb)  # below the slop: the body stays put
    s.update(1 / 60)
    c = s.manifolds()[0].contacts[0]
    point = np.array(c.point, float)
    n = np.array([0.0, 1.0, 0.0])
    t1 = np.array([-1.0, 0.0, 0.0])  # computeOrthogonal((0,1,0)) (src/solver.zig:967-981)
    t2 = np.cross(n, t1)
    assert c.normal_mass == pytest.approx(1.0 / _k_normal(point, pb, qb, 3.0, inv_i, n), rel=1e-6)
    assert c.tangent_mass[0] == pytest.approx(1.0 / _k_normal(point, pb, qb, 3.0, inv_i, t1), rel=1e-6)
    assert c.tangent_mass[1] == pytest.approx(1.0 / _k_normal(point, pb, qb, 3.0, inv_i, t2), rel=1e-6)
    s.close()


def test_resting_cube_carries_its_weight(oracle):
    """A README cube resting on the ground: the normal impulses the step
    accumulates over its substeps balance gravity over the frame, sum = m g dt
    (within 3 %), and the cube does not sink past the slop."""
    s = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=4)
    gv, go = scenes.hand_ground(8, 8)
    g = s.addCollider(voxel_size=S, **dict(zip(("brick_pos", "voxels"), scenes.grid_to_bricks(gv, go))))
    s.createStaticBody((0.0, -0.1, 0.0), IDENT, g)
    cv, co = scenes.readme_cube()
    c = s.addCollider(voxel_size=S, **dict(zip(("brick_pos", "voxels"), scenes.grid_to_bricks(cv, co))))
    h = s.createDynamicBody((0.0, 0.5, 0.0), IDENT, c, 1.0)
    inv_m = s.colliderMassProperties(c, 1.0)[0]
    sums = []
    for f in range(42):
        s.update(1 / 60)
        if f >= 25:
            assert not s.isSleeping(h)
            m = s.manifolds()[0]
            sums.append(sum(m.contacts[k].normal_impulse for k in range(m.contact_count)))
    assert np.mean(sums) == pytest.approx(9.81 / 60 / inv_m, rel=0.03)
    s.close()


def test_impulse_clamps_hold_in_a_pile(oracle):
    """Every stored contact of the 64-cube scene over 120 frames: normal impulse
    >= 0 and each accumulated friction impulse within +-friction x normal
    impulse (src/solver.zig:768,805-806,826-827)."""
    sc = scenes.c0_small(frames=120)
    s = vxp.from_scene(sc, lib=oracle)
    clamped = 0
    for f in range(120):
        s.update(sc.dt)
        for m in s.manifolds():
            for k in range(m.contact_count):
                c = m.contacts[k]
                assert c.normal_impulse >= 0.0
                lim = m.friction * c.normal_impulse
                for t in c.tangent_impulse:
                    assert abs(t) <= lim * (1 + 1e-6) + 1e-12
                    clamped += abs(t) >= lim * (1 - 1e-6) and lim > 0
    assert clamped > 0  # the clamp is reached (sliding contacts exist)
    s.close()


@pytest.mark.parametrize("v,row", [((1.0, 0.0, 0.0), 0), ((0.0, 0.0, 1.0), 1)])
def test_friction_rows_along_the_reference_tangents(oracle, v, row):
    """The tangents of a +y normal are t1 = computeOrthogonal(n) = (-1, 0, 0)
    and t2 = n x t1 = (0, 0, 1) (src/solver.zig:659-662,967-981): sliding along
    +x builds a positive impulse on row 0 (it pushes along -x = +t1) and none on
    row 1; sliding along +z a negative impulse on row 1 (slips caught: t1 x n
    for t2, a tangent of the wrong sign)."""
    s, h, pb, _ = _probe_system(oracle, 0.002, gravity=(0.0, -9.81, 0.0), hertz=30.0)
    for _ in range(3):
        s.update(1 / 60)
    s.setState(h, linear_velocity=v)
    s.update(1 / 60)
    c = s.manifolds()[0].contacts[0]
    ti = list(c.tangent_impulse)
    assert c.normal_impulse > 0
    expect = 1.0 if row == 0 else -1.0
    assert np.sign(ti[row]) == expect and abs(ti[row]) > 0.5 * 0.6 * c.normal_impulse
    assert abs(ti[1 - row]) < 0.2 * abs(ti[row])
    s.close()


@pytest.mark.parametrize("vy,differs", [(-0.8, False), (-3.0, True)])
def test_restitution_applies_only_below_minus_one(oracle, vy, differs):
    """bias = max(bias, -e vn) only when vn < -1 m/s (src/solver.zig:669-673):
    an approach at 0.8 m/s gives the same trajectory with restitution on or
    off; at 3 m/s (penetrating, so the soft bias is below -e vn) it differs."""
    runs = []
    for on in (1, 0):
        s, h, pb, _ = _probe_system(oracle, 0.002, enable_restitution=on, default_restitution=0.5)
        s.setState(h, linear_velocity=(0.0, vy, 0.0))
        for _ in range(3):
            s.update(1 / 60)
        runs.append(s.bodies()[h.index])
        s.close()
    same = (np.array_equal(runs[0]["position"], runs[1]["position"]) and
            np.array_equal(runs[0]["linear_velocity"], runs[1]["linear_velocity"]))
    assert same != differs


def _prio(key):
    # the solver colouring's priority: a 64-bit finalizer mix of the pair key
    h = key
    h ^= h >> 33
    h = (h * 0xff51afd7ed558ccd) & (2**64 - 1)
    h ^= h >> 33
    h = (h * 0xc4ceb9fe1a85ec53) & (2**64 - 1)
    h ^= h >> 33
    return h >> 32


@pytest.mark.parametrize("scene", ["c0", "batched"])
def test_solver_colouring_is_the_greedy_colouring(oracle, scene):
    """Colours of the solved manifolds, frame by frame: recomputed here by the
    textbook greedy rule (smallest colour not taken by an earlier manifold that
    shares an awake dynamic body, manifolds in ascending (prio, key) order) and
    equal; valid (no two manifolds of one colour share an awake dynamic body);
    and exactly the manifolds with an awake body are solved (src/system.zig:
    687-702: both asleep, static counting as asleep, are skipped)."""
    if scene == "c0":
        sc = scenes.c0_small(frames=150)
        worlds, frames = [0], 150
    else:
        sc = scenes.batched_worlds(n_worlds=3, lo=40, hi=200, blob_pool=4, name="c3_3")
        worlds, frames = [0, 1, 2], 80
    s = vxp.from_scene(sc, lib=oracle)
    checked = 0
    for f in range(frames):
        pre = [s.bodies(w) for w in worlds]
        s.update(sc.dt)
        for wi, w in enumerate(worlds):
            b = pre[wi]
            awake = {int(i) for i, t, sl, a in zip(b["id"], b["body_type"], b["is_sleeping"], b["active"])
                     if a and t == vxp.DYNAMIC and not sl}
            ms = s.manifolds(w)
            solved = [m for m in ms if m.body_a_id in awake or m.body_b_id in awake]
            assert all(m.color != 0xFFFFFFFF for m in solved)
            assert all(m.color == 0xFFFFFFFF for m in ms if m not in solved)
            order = sorted(solved, key=lambda m: (_prio(m.body_a_id << 32 | m.body_b_id),
                                                  m.body_a_id << 32 | m.body_b_id))
            taken = {}
            for m in order:
                used = set()
                for bid in (m.body_a_id, m.body_b_id):
                    if bid in awake:
                        used |= taken.get(bid, set())
                c = 0
                while c in used:
                    c += 1
                assert m.color == c, (f, w, m.body_a_id, m.body_b_id)
                for bid in (m.body_a_id, m.body_b_id):
                    if bid in awake:
                        taken.setdefault(bid, set()).add(c)
            checked += len(order)
    assert checked > 1000
    s.close()
