"""Collider build from solid occupancy: classification (src/collision_voxel.zig:
194-235), recomputeMetadata (src/collider.zig:90-133,283-333) and
computeMassProperties (src/body.zig:462-519).

CPU: the restatement's solid path (vxo_collider_create_solid) gives the same
stored bricks, metadata and mass properties as classifying the same grid with
the numpy classifier of the scene generator (an independent restatement of
classifyFromNeighbors, itself pinned by the classification known answers) and
storing the bytes.  GPU: the device path (k_collider_classify, k_collider_meta,
k_collider_mass) against both.
"""
import numpy as np
import pytest

import vxp
from vxp import scenes


def shapes():
    rng = np.random.default_rng(5)
    blob = rng.random((19, 13, 17)) < 0.55
    shell = np.ones((10, 10, 10), dtype=bool)
    shell[2:8, 2:8, 2:8] = False
    slab = np.ones((20, 1, 9), dtype=bool)
    line = np.ones((1, 1, 12), dtype=bool)
    cube = np.ones((3, 3, 3), dtype=bool)
    return [("blob", blob, (-7, -3, -9)), ("shell", shell, (0, 0, 0)), ("slab", slab, (-10, 0, -4)),
            ("line", line, (3, -1, -6)), ("cube", cube, (-1, -1, -1))]


def classified_bricks(solid, origin, positions):
    """Voxel bytes of the numpy classifier for the given brick positions."""
    cls = scenes.classify_grid(solid)
    v = np.arange(512)
    local = np.stack([v & 7, (v >> 3) & 7, v >> 6], axis=1)
    g = np.asarray(positions)[:, None, :] * 8 + local[None] - np.asarray(origin)
    inside = ((g >= 0) & (g < np.array(solid.shape))).all(axis=2)
    out = np.zeros(inside.shape, dtype=np.uint8)
    gi = g[inside]
    out[inside] = cls[gi[:, 0], gi[:, 1], gi[:, 2]]
    return out


def build_both(s, solid, origin, vs):
    # the same bricks in the same order on both paths: where brick positions
    # alias (hashBrickPos), the stored position is the first one given
    sp, sm = scenes.grid_to_solid_bricks(solid, origin)
    a = s.addCollider(voxel_size=vs, brick_pos=sp, voxels=classified_bricks(solid, origin, sp))
    b = s.addColliderSolid(vs, sp, sm)
    return a, b


def solid_bricks(s, c):
    """Stored bricks holding at least one voxel (the byte path also stores the
    all-empty bricks of the grid's bounding box; they read as empty)."""
    p, v = s.colliderBricks(c)
    keep = (v & 7).any(axis=1)
    return p[keep], v[keep]


def assert_same(s, a, b, t, densities=(1.0, 2.5)):
    pa, va = solid_bricks(s, a)
    pb, vb = solid_bricks(s, b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb), f"{t}: bricks differ"
    assert s.colliderInfo(a) == s.colliderInfo(b), t
    for d in densities:
        ma, mb = s.colliderMassProperties(a, d), s.colliderMassProperties(b, d)
        assert ma[0] == mb[0] and np.array_equal(ma[1], mb[1]) and ma[2] == mb[2], t


def test_solid_path_equals_classified_bytes(oracle):
    s = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=4)
    for t, solid, origin in shapes():
        a, b = build_both(s, solid, origin, 0.2)
        assert_same(s, a, b, t)
        assert s.colliderInfo(b)[4] == int(solid.sum())
    s.close()


def test_solid_path_rejects_duplicate_positions(oracle):
    s = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=4)
    bp = np.array([[0, 0, 0], [0, 0, 0]], dtype=np.int32)
    sm = np.ones((2, 8), dtype=np.uint64)
    with pytest.raises(vxp.VxpError):
        s.addColliderSolid(0.2, bp, sm)
    s.close()


def test_solid_bricks_roundtrip():
    # the occupancy packing used by the solid path: bit x + 8y of word z
    rng = np.random.default_rng(1)
    solid = rng.random((11, 9, 10)) < 0.3
    origin = (-5, 2, -9)
    bp, words = scenes.grid_to_solid_bricks(solid, origin)
    back = np.zeros_like(solid)
    for (bx, by, bz), w in zip(bp, words):
        for z in range(8):
            for y in range(8):
                for x in range(8):
                    if (int(w[z]) >> (x + 8 * y)) & 1:
                        back[bx * 8 + x - origin[0], by * 8 + y - origin[1], bz * 8 + z - origin[2]] = True
    assert np.array_equal(back, solid)


@pytest.mark.gpu
def test_device_build_matches_restatement(product, oracle):
    sp_ = vxp.VoxelPhysicsSystem(lib=product, max_bodies=4)
    so = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=4)
    grounds = [("ground256x2", np.ones((256, 2, 256), dtype=bool), (-128, 0, -128))]
    for t, solid, origin in shapes() + grounds:
        print(f"collider build: {t} {solid.shape}", flush=True)
        a, b = build_both(sp_, solid, origin, 0.2)
        assert_same(sp_, a, b, t)
        sp, sm = scenes.grid_to_solid_bricks(solid, origin)
        c = so.addColliderSolid(0.2, sp, sm)
        pb, vb = sp_.colliderBricks(b)
        pc, vc = so.colliderBricks(c)  # both from occupancy: the same bricks
        assert np.array_equal(pb, pc) and np.array_equal(vb, vc), t
        assert sp_.colliderInfo(b) == so.colliderInfo(c), t
        for d in (1.0, 2.5):
            mb, mc = sp_.colliderMassProperties(b, d), so.colliderMassProperties(c, d)
            assert mb[0] == mc[0] and np.array_equal(mb[1], mc[1]) and mb[2] == mc[2], t
    sp_.close()
    so.close()
