"""The allocation-free entry point vxp_collider_mass_properties_noalloc,
called through the C ABI, equals vxp_collider_mass_properties (which stages
through library-allocated buffers) bit for bit, on two calls in a row with the
second collider smaller than the first (the shared device scratch keeps the
first call's bytes beyond the second's), and rejects a collider larger than its
scratch array before launching anything."""
import ctypes as C

import numpy as np
import pytest

import vxp
from vxp import _abi, scenes

pytestmark = pytest.mark.gpu


def _entry(lib):
    fn = lib.dll.vxp_collider_mass_properties_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.POINTER(C.c_float)]
    mx = lib.dll.vxp_noalloc_max_bricks
    mx.restype = C.c_uint32
    mx.argtypes = []
    return fn, mx


def _allocating(s, cid, density):
    inv_mass, inertia, com = s.colliderMassProperties(cid, density)
    return np.concatenate([[inv_mass], np.asarray(inertia, np.float32).ravel(),
                           np.asarray(com, np.float32)]).astype(np.float32)


def test_noalloc_equals_allocating_entry_point(product):
    fn, mx = _entry(product)
    s = vxp.VoxelPhysicsSystem(lib=product)
    rng = scenes.SplitMix64(7)
    big_vox, big_org = scenes.blob_shape(rng, n=24, p=0.45)  # up to 4^3 bricks
    big = s.addCollider(voxel_size=0.1, **dict(zip(("brick_pos", "voxels"),
                                                    scenes.grid_to_bricks(big_vox, big_org))))
    small_vox, small_org = scenes.readme_cube()
    small = s.addCollider(voxel_size=0.2, **dict(zip(("brick_pos", "voxels"),
                                                      scenes.grid_to_bricks(small_vox, small_org))))
    nb_big = len(s.colliderBricks(big)[0])
    nb_small = len(s.colliderBricks(small)[0])
    print(f"noalloc: {nb_big} then {nb_small} bricks, limit {mx()}", flush=True)
    assert nb_small < nb_big <= mx()
    for cid, dens in ((big, 1.0), (small, 2.5)):
        out = (C.c_float * 13)()
        rc = fn(s.ptr, cid, dens, out)
        assert rc == _abi.VXP_OK, product.last_error()
        got = np.frombuffer(out, np.float32)
        want = _allocating(s, cid, dens)
        assert got.tobytes() == want.tobytes(), (cid, got, want)
    s.close()


def test_noalloc_rejects_colliders_beyond_its_scratch(product):
    fn, mx = _entry(product)
    s = vxp.VoxelPhysicsSystem(lib=product)
    n = int(mx()) + 1
    bp = np.zeros((n, 3), np.int32)
    bp[:, 0] = np.arange(n)  # distinct brick hashes: one stored brick each
    vox = np.zeros((n, 512), np.uint8)
    vox[:, 0] = vxp.voxel_byte(vxp.INSIDE, 0)
    cid = s.addCollider(voxel_size=0.1, brick_pos=bp, voxels=vox)
    assert len(s.colliderBricks(cid)[0]) == n
    out = (C.c_float * 13)()
    assert fn(s.ptr, cid, 1.0, out) == _abi.VXP_ERR_CAPACITY
    assert b"allocation-free limit" in product.last_error()
    s.close()
