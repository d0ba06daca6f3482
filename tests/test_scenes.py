"""Scene generators: determinism, SplitMix64 known answers, brick packing."""
import numpy as np

from vxp import scenes


def test_splitmix64_known_answers():
    # published SplitMix64 output for seed 0 (Vigna's reference implementation)
    r = scenes.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                0x06C45D188009454F]
    r1, r2 = scenes.SplitMix64(7), scenes.SplitMix64(7)
    bulk = r1.bulk_u64(5)
    assert [int(x) for x in bulk] == [r2.next_u64() for _ in range(5)]
    assert r1.next_u64() == r2.next_u64()


def test_quaternions_are_unit():
    r = scenes.SplitMix64(3)
    for _ in range(100):
        assert abs(np.linalg.norm(r.quat()) - 1) < 1e-12


def test_scene_determinism():
    a, b = scenes.c0_small(), scenes.c0_small()
    assert [(x.position, x.rotation) for x in a.worlds[0]] == [(x.position, x.rotation) for x in b.worlds[0]]
    for ca, cb in zip(a.colliders, b.colliders):
        assert (ca.brick_pos == cb.brick_pos).all() and (ca.voxels == cb.voxels).all()


def test_grid_to_bricks_round_trip():
    rng = np.random.default_rng(0)
    vox = (rng.integers(0, 5, size=(13, 7, 9)) * (rng.random((13, 7, 9)) < 0.3)).astype(np.uint8)
    origin = (-5, 3, -9)
    pos, v = scenes.grid_to_bricks(vox, origin)
    back = np.zeros_like(vox)
    for (bx, by, bz), data in zip(pos, v):
        for i in np.nonzero(data)[0]:
            x, y, z = bx * 8 + i % 8 - origin[0], by * 8 + (i // 8) % 8 - origin[1], bz * 8 + i // 64 - origin[2]
            back[x, y, z] = data[i]
    assert (back == vox).all()


def test_classify_grid_counts_exposed_faces():
    solid = np.ones((3, 3, 3), bool)
    v = scenes.classify_grid(solid)
    t = v & 7
    assert t[1, 1, 1] == 4          # inside
    assert (t[0, 0, 0], t[1, 0, 0], t[1, 1, 0]) == (1, 2, 3)  # corner, edge, face
    assert v[1, 1, 0] >> 3 == 5     # -z face normal


def test_blob_is_one_component():
    from scipy import ndimage
    vox, org = scenes.blob_shape(scenes.SplitMix64(1))
    lab, n = ndimage.label((vox & 7) != 0)
    assert n == 1 and ((vox & 7) != 0).sum() > 500


def test_c1_initial_bodies_disjoint_columns():
    sc = scenes.c1_large(n_bodies=64)
    assert sc.n_dynamic == 64
    ys = [b.position[1] for b in sc.worlds[0][1:]]
    assert min(ys) >= 1.0
