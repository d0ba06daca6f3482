"""Body mutators and queries between frames (src/system.zig:462-473,886-1023,
src/body.zig:266-435) on the CPU restatement, pinned by closed forms.

The restatement's arithmetic is checked against independent derivations: the
mass of a voxel lattice, the inertia tensor of a box lattice computed here in
float64 and rotated, exact compositions (an impulse at a point equals the
impulse plus the angular impulse of (p - pos) x J), the type guards, wake-up
and sleep, and the world AABB of an unrotated body.  The GPU library runs the
same mutators in test_gpu_parity.py, frame for frame against this one.
"""
import ctypes as C

import numpy as np
import pytest

import vxp
from vxp import scenes

F32 = np.float32


def _world(oracle, shape=(3, 3, 3), vs=0.2, rot=(0.0, 0.0, 0.0, 1.0), pos=(1.0, 2.0, 3.0)):
    s = vxp.VoxelPhysicsSystem(lib=oracle, max_bodies=16)
    vox, org = scenes.box_shape(*shape)
    bp, v = scenes.grid_to_bricks(vox, org)
    cid = s.addCollider(voxel_size=vs, brick_pos=bp, voxels=v)
    dyn = s.createDynamicBody(pos, rot, cid, 1.0)
    return s, cid, dyn, org


def _lattice_inv_inertia(shape, org, vs):
    """Inverse inertia of a box lattice of point masses vs^3 about its centroid."""
    pts = np.array([(x, y, z) for x in range(shape[0]) for y in range(shape[1])
                    for z in range(shape[2])], dtype=np.float64)
    pts = (pts + np.array(org)) * vs + vs
    r = pts - pts.mean(axis=0)
    m = vs ** 3
    I = np.zeros((3, 3))
    for v in r:
        I += m * (np.dot(v, v) * np.eye(3) - np.outer(v, v))
    return np.linalg.inv(I)


def test_impulse_and_force_scale_by_inverse_mass(oracle):
    s, cid, h, _ = _world(oracle)
    vs = F32(0.2)
    inv_mass = F32(1.0) / (vs * vs * vs * F32(1.0) * F32(27))  # computeMassProperties
    J = np.array([0.5, -1.25, 2.0], dtype=F32)
    s.applyImpulse(h, J)
    np.testing.assert_array_equal(s.getLinearVelocity(h), J * inv_mass)
    s.applyForce(h, J)  # same rule, accumulated
    np.testing.assert_array_equal(s.getLinearVelocity(h), J * inv_mass + J * inv_mass)
    assert s.getAngularVelocity(h) == (0.0, 0.0, 0.0)
    s.close()


def test_torque_uses_rotated_inverse_inertia(oracle):
    # anisotropic 3x1x5 box rotated 90 degrees about z: I_world^-1 = R I^-1 R^T
    # (1 m voxels: the inertia's determinant is far above Mat3.inverse's 1e-6
    # cut-off, src/math.zig:482-484)
    shape, vs = (3, 1, 5), 1.0
    c = np.float32(np.sqrt(0.5))
    rot = (0.0, 0.0, float(c), float(c))
    s, cid, h, org = _world(oracle, shape, vs, rot)
    tau = np.array([0.01, 0.02, -0.03])
    s.applyTorque(h, tau)
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    expect = R @ _lattice_inv_inertia(shape, org, vs) @ R.T @ tau
    np.testing.assert_allclose(s.getAngularVelocity(h), expect, rtol=2e-4, atol=1e-6)
    s.close()
    # at 0.2 m the determinant is below the cut-off: the inverse is the identity
    s, cid, h, org = _world(oracle, shape, 0.2, rot)
    s.applyTorque(h, tau)
    np.testing.assert_allclose(s.getAngularVelocity(h), tau, rtol=1e-6)
    s.close()


def test_impulse_at_point_is_impulse_plus_angular_impulse(oracle):
    J = (0.3, -0.7, 1.1)
    p = (1.4, 2.6, 2.2)
    rot = (0.1, 0.2, 0.3, 0.9273618)
    a, _, ha, _ = _world(oracle, (3, 2, 4), 0.1, rot)
    b, _, hb, _ = _world(oracle, (3, 2, 4), 0.1, rot)
    a.applyImpulseAtPoint(ha, J, p)
    b.applyImpulse(hb, J)
    pos = np.array(b.getPosition(hb), dtype=F32)
    r = np.array(p, dtype=F32) - pos
    Jf = np.array(J, dtype=F32)
    ang = np.array([r[1] * Jf[2] - r[2] * Jf[1], r[2] * Jf[0] - r[0] * Jf[2],
                    r[0] * Jf[1] - r[1] * Jf[0]], dtype=F32)  # Vec3.cross
    b.applyAngularImpulse(hb, ang)
    assert a.bodies()[ha.index].tobytes() == b.bodies()[hb.index].tobytes()
    # the force variant follows the same rule
    a.applyForceAtPoint(ha, J, p)
    b.applyForce(hb, J)
    b.applyTorque(hb, ang)
    assert a.bodies()[ha.index].tobytes() == b.bodies()[hb.index].tobytes()
    a.close()
    b.close()


def test_type_guards_wake_and_sleep(oracle):
    s, cid, dyn, _ = _world(oracle)
    ground = s.createStaticBody((0.0, -1.0, 0.0), (0, 0, 0, 1), cid)
    kin = s.createKinematicBody((4.0, 0.0, 0.0), (0, 0, 0, 1), cid)
    before = s.bodies().copy()
    for h in (ground, kin):
        s.applyForce(h, (1, 2, 3))
        s.applyImpulse(h, (1, 2, 3))
        s.applyTorque(h, (1, 2, 3))
        s.applyImpulseAtPoint(h, (1, 2, 3), (0, 0, 0))
        s.setLinearVelocity(h, (1, 2, 3))
        s.setAngularVelocity(h, (1, 2, 3))
    after = s.bodies()
    for h in (ground, kin):  # dynamic-only mutators leave them unchanged
        assert after[h.index].tobytes() == before[h.index].tobytes()
    s.setPosition(kin, (5.0, 0.5, 0.0))
    assert s.getPosition(kin) == (5.0, 0.5, 0.0)
    s.setPosition(ground, (0.0, -2.0, 0.0))  # any type moves; static does not wake
    assert s.getPosition(ground) == (0.0, -2.0, 0.0) and s.isSleeping(ground)
    s.sleep(ground)
    assert s.isSleeping(ground)
    # sleep zeroes both velocities; wakeUp clears the flag and the timer
    s.setLinearVelocity(dyn, (1.0, 0.0, 0.0))
    s.setAngularVelocity(dyn, (0.0, 2.0, 0.0))
    s.sleep(dyn)
    st = s.bodies()[dyn.index]
    assert st["is_sleeping"] and not st["linear_velocity"].any() and not st["angular_velocity"].any()
    s.wakeUp(dyn)
    st = s.bodies()[dyn.index]
    assert not st["is_sleeping"] and st["sleep_timer"] == 0.0
    s.setRotation(dyn, (0.0, 1.0, 0.0, 0.0))
    assert s.getRotation(dyn) == (0.0, 1.0, 0.0, 0.0)
    s.close()


def test_find_body_by_id_and_invalid_handles(oracle):
    s, cid, a, _ = _world(oracle)
    b = s.createDynamicBody((0.0, 5.0, 0.0), (0, 0, 0, 1), cid, 1.0)
    ids = s.bodies()["id"]
    for h in (a, b):
        f = s.findBodyById(int(ids[h.index]))
        assert (f.index, f.generation) == (h.index, h.generation)
    assert s.findBodyById(999) is None
    s.destroyBody(b)
    assert s.findBodyById(int(ids[b.index])) is None
    with pytest.raises(vxp.VxpError):
        s.applyImpulse(b, (1, 0, 0))
    with pytest.raises(vxp.VxpError):
        s.getWorldAABB(b)
    assert s.getBodyCount() == 1
    s.close()


def test_world_aabb_of_unrotated_body(oracle):
    # identity rotation: the collider's local AABB (voxel bounds times the voxel
    # size) shifted by the position
    vs = 0.25
    s, cid, h, org = _world(oracle, (3, 2, 5), vs, pos=(1.0, -2.0, 0.5))
    lo, hi = s.getWorldAABB(h)
    shape = np.array((3, 2, 5))
    lo_l = np.array(org, dtype=F32) * F32(vs)
    hi_l = (np.array(org) + shape).astype(F32) * F32(vs)
    p = np.array((1.0, -2.0, 0.5), dtype=F32)
    np.testing.assert_array_equal(lo, lo_l + p)
    np.testing.assert_array_equal(hi, hi_l + p)
    s.close()
