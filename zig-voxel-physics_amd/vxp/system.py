"""Python mirror of the reference's `VoxelPhysicsSystem` API over the C ABI.

Method names and argument meaning follow src/system.zig (createDynamicBody,
createStaticBody, update, step, getPosition, ...).  The class is the same over
the product library (HIP pipeline) and, for the checker, over the CPU
restatement; `VoxelPhysicsSystem()` binds the product and fails loudly when the
HIP library is not built.
"""
import ctypes as C

import numpy as np

from . import _abi

# vxp_tuning.broad_phase values (VXP_BROAD_AUTO, _SWEEP, _LBVH)
BROAD_NAMES = ["auto", "sweep", "lbvh"]


class VoxelCollider:
    """VoxelCollider built on the host (src/collider.zig:144-409), uploaded on
    `VoxelPhysicsSystem.addCollider`.  Voxels are set per coordinate like
    setCollisionVoxel; bricks are 8^3 and keyed by floor-div brick position."""

    def __init__(self, voxel_size=0.1):
        self.voxel_size = float(voxel_size)
        self.bricks = {}

    def setCollisionVoxel(self, x, y, z, voxel_byte):
        bp = (x // 8, y // 8, z // 8)
        b = self.bricks.get(bp)
        if b is None:
            b = np.zeros(512, dtype=np.uint8)
            self.bricks[bp] = b
        b[(x % 8) + 8 * (y % 8) + 64 * (z % 8)] = voxel_byte

    def arrays(self):
        keys = sorted(self.bricks, key=lambda p: (p[2], p[1], p[0]))
        pos = np.array(keys, dtype=np.int32).reshape(-1, 3)
        vox = np.stack([self.bricks[k] for k in keys]) if keys else np.zeros((0, 512), np.uint8)
        return pos, vox


class VoxelPhysicsSystem:
    def __init__(self, config=None, n_worlds=1, lib=None, **over):
        self.lib = lib if lib is not None else _abi.product_lib()
        self.cfg = config if config is not None else _abi.default_config(**over)
        self.n_worlds = n_worlds
        self.ptr = C.c_void_p()
        self.lib.check(self.lib.system_create(C.byref(self.cfg), n_worlds, C.byref(self.ptr)),
                       "system_create")

    def close(self):
        if self.ptr:
            self.lib.system_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- recorded steps (product only; results never depend on it) ------------
    def setRecorded(self, on=True):
        """Run the default pipeline's steps from recorded HIP graphs (three per
        step, sized for capacities; the host reads the counts twice per step)."""
        self.lib.check(self.lib.graph_enable(self.ptr, 1 if on else 0), "graph_enable")

    def graphInfo(self):
        g = _abi.GraphInfo()
        self.lib.check(self.lib.graph_info_get(self.ptr, C.byref(g)), "graph_info_get")
        return {f: getattr(g, f) for f, _ in g._fields_}

    # ---- launch tuning (product only; results never depend on it) -------------
    def tuning(self):
        t = _abi.Tuning()
        self.lib.check(self.lib.tuning_get(self.ptr, C.byref(t)), "tuning_get")
        return {"pair_color_max": t.pair_color_max, "windows": t.windows, "fixed": bool(t.fixed),
                "unit_threads": t.unit_threads, "broad_phase": BROAD_NAMES[t.broad_phase],
                "probe_threads": t.probe_threads, "probe_xcd_run": t.probe_xcd_run}

    def setTuning(self, pair_color_max=None, unit_threads=0, broad_phase="auto", probe_threads=0,
                  probe_xcd_run=0):
        """Fix the solver's lane-pair colour limit (None restarts the online tuner
        over the knobs left at their default), and force the solver's workgroup
        size (256 or 512), the broad phase ("sweep" or "lbvh"), the probe's
        workgroup size (64, 128 or 256) and its XCD run (a power of two); 0 /
        "auto" leaves a knob to the tuner."""
        t = _abi.Tuning(pair_color_max=_abi.TUNE_AUTO if pair_color_max is None else pair_color_max,
                        unit_threads=unit_threads, broad_phase=BROAD_NAMES.index(broad_phase),
                        probe_threads=probe_threads, probe_xcd_run=probe_xcd_run)
        self.lib.check(self.lib.tuning_set(self.ptr, C.byref(t)), "tuning_set")

    # ---- colliders / bodies (src/system.zig:294-433) -------------------------
    def addCollider(self, collider=None, voxel_size=None, brick_pos=None, voxels=None):
        if collider is not None:
            brick_pos, voxels = collider.arrays()
            voxel_size = collider.voxel_size
        bp = np.ascontiguousarray(brick_pos, dtype=np.int32)
        vx = np.ascontiguousarray(voxels, dtype=np.uint8)
        out = C.c_uint32()
        self.lib.check(self.lib.collider_create(self.ptr, voxel_size, len(bp),
                                                bp.ctypes.data_as(C.POINTER(C.c_int32)),
                                                vx.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                C.byref(out)), "collider_create")
        return out.value

    def addColliderSolid(self, voxel_size, brick_pos, solid):
        """A collider from solid occupancy (8 uint64 words per brick, bit x + 8y of
        word z), classified with classifyFromNeighbors by the library."""
        bp = np.ascontiguousarray(brick_pos, dtype=np.int32)
        sm = np.ascontiguousarray(solid, dtype=np.uint64)
        out = C.c_uint32()
        self.lib.check(self.lib.collider_create_solid(self.ptr, voxel_size, len(bp),
                                                      bp.ctypes.data_as(C.POINTER(C.c_int32)),
                                                      sm.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                      C.byref(out)), "collider_create_solid")
        return out.value

    def colliderInfo(self, collider):
        """(min_bound, max_bound, local AABB min, local AABB max, solid_count)."""
        b, a, n = (C.c_int32 * 6)(), (C.c_float * 6)(), C.c_uint32()
        self.lib.check(self.lib.collider_info(self.ptr, collider, b, a, C.byref(n)), "collider_info")
        return tuple(b[:3]), tuple(b[3:]), tuple(a[:3]), tuple(a[3:]), n.value

    def colliderMassProperties(self, collider, density):
        """(inv_mass, inverse inertia 3x3, centre of mass)."""
        out = (C.c_float * 13)()
        self.lib.check(self.lib.collider_mass_properties(self.ptr, collider, density, out),
                       "collider_mass_properties")
        return out[0], np.array(out[1:10], dtype=np.float32).reshape(3, 3), tuple(out[10:13])

    def colliderBricks(self, collider):
        """Stored bricks: positions (n, 3) int32 and voxel bytes (n, 512) uint8."""
        n = C.c_uint32()
        self.lib.check(self.lib.collider_bricks(self.ptr, collider, None, None, 0, C.byref(n)),
                       "collider_bricks")
        bp = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        vx = np.zeros((max(n.value, 1), 512), dtype=np.uint8)
        self.lib.check(self.lib.collider_bricks(self.ptr, collider,
                                                bp.ctypes.data_as(C.POINTER(C.c_int32)),
                                                vx.ctypes.data_as(C.POINTER(C.c_uint8)), n.value,
                                                C.byref(n)), "collider_bricks")
        return bp[: n.value], vx[: n.value]

    def _create(self, t, position, rotation, collider, density=1.0, world=0):
        h = _abi.Handle()
        pos = (C.c_float * 3)(*position)
        rot = (C.c_float * 4)(*rotation)
        self.lib.check(self.lib.body_create_w(self.ptr, world, t, pos, rot, collider, density,
                                              C.byref(h)), "body_create")
        return h

    def createDynamicBody(self, position, rotation, collider, density, world=0):
        return self._create(_abi.DYNAMIC, position, rotation, collider, density, world)

    def createStaticBody(self, position, rotation, collider, world=0):
        return self._create(_abi.STATIC, position, rotation, collider, 1.0, world)

    def createKinematicBody(self, position, rotation, collider, world=0):
        return self._create(_abi.KINEMATIC, position, rotation, collider, 1.0, world)

    def createDynamicBodyWithMass(self, position, rotation, collider, inv_mass, inv_inertia,
                                  world=0):
        """createDynamicBodyWithMass (src/system.zig:323-351); inv_inertia is the
        local inverse inertia, 9 floats row-major."""
        h = _abi.Handle()
        self.lib.check(self.lib.body_create_with_mass_w(self.ptr, world, (C.c_float * 3)(*position),
                                                        (C.c_float * 4)(*rotation), collider,
                                                        inv_mass, (C.c_float * 9)(*inv_inertia),
                                                        C.byref(h)), "body_create_with_mass_w")
        return h

    # ---- configuration (src/system.zig:1034-1052) ------------------------------
    def setGravity(self, gravity):
        self.cfg.gravity[:] = list(gravity)
        self.lib.check(self.lib.system_set_config(self.ptr, C.byref(self.cfg)), "system_set_config")

    def getGravity(self):
        return tuple(self.cfg.gravity)

    def setSolverConfig(self, **fields):
        """SolverConfig fields by name (substeps, velocity_iterations, hertz, ...)."""
        old = {k: getattr(self.cfg, k) for k in fields}
        for k, v in fields.items():
            setattr(self.cfg, k, v)
        rc = self.lib.system_set_config(self.ptr, C.byref(self.cfg))
        if rc != _abi.VXP_OK:
            for k, v in old.items():
                setattr(self.cfg, k, v)
        self.lib.check(rc, "system_set_config")

    def destroyBody(self, handle, world=0):
        return self.lib.body_destroy_w(self.ptr, world, handle)

    # ---- stepping (src/system.zig:508-555) ------------------------------------
    def update(self, dt):
        self.lib.check(self.lib.system_update(self.ptr, dt), "system_update")

    def step(self, dt):
        self.lib.check(self.lib.system_step(self.ptr, dt), "system_step")

    def update_frames(self, dt, frames):
        self.lib.check(self.lib.system_update_frames(self.ptr, dt, frames), "system_update_frames")

    # ---- read-back ------------------------------------------------------------
    def bodies(self, world=0):
        """Body states of `world` in slot order as a structured numpy array."""
        n = self.cfg.max_bodies
        arr = (_abi.BodyState * n)()
        cnt = C.c_uint32()
        self.lib.check(self.lib.bodies_get_w(self.ptr, world, arr, n, C.byref(cnt)), "bodies_get")
        out = np.frombuffer(arr, dtype=np.dtype([
            ("position", "<f4", 3), ("rotation", "<f4", 4), ("linear_velocity", "<f4", 3),
            ("angular_velocity", "<f4", 3), ("sleep_timer", "<f4"), ("id", "<u4"),
            ("body_type", "u1"), ("is_sleeping", "u1"), ("active", "u1"), ("pad", "u1")]))
        return out[: cnt.value].copy()

    def stats(self, world=0):
        s = _abi.Stats()
        self.lib.check(self.lib.stats_get_w(self.ptr, world, C.byref(s)), "stats_get")
        return s.as_dict()

    def stats_all(self):
        """Stats of every world (one copy)."""
        arr = (_abi.Stats * self.n_worlds)()
        self.lib.check(self.lib.stats_get_all(self.ptr, arr, self.n_worlds), "stats_get_all")
        return arr

    def manifolds(self, world=0):
        cnt = C.c_uint32()
        self.lib.check(self.lib.manifolds_get_w(self.ptr, world, None, 0, C.byref(cnt)), "manifolds")
        n = cnt.value
        arr = (_abi.Manifold * max(n, 1))()
        self.lib.check(self.lib.manifolds_get_w(self.ptr, world, arr, n, C.byref(cnt)), "manifolds")
        return [arr[i] for i in range(n)]

    def manifold_array(self, world=0):
        """Manifolds as a raw structured array (byte-exact view of vxp_manifold)."""
        ms = self.manifolds(world)
        if not ms:
            return np.zeros(0, dtype=np.uint8).reshape(0, C.sizeof(_abi.Manifold))
        buf = b"".join(bytes(m) for m in ms)
        return np.frombuffer(buf, dtype=np.uint8).reshape(len(ms), C.sizeof(_abi.Manifold))

    def pairs(self, world=0):
        cnt = C.c_uint32()
        self.lib.check(self.lib.pairs_get_w(self.ptr, world, None, 0, C.byref(cnt)), "pairs")
        n = cnt.value
        arr = np.zeros(max(n, 1), dtype=np.uint64)
        self.lib.check(self.lib.pairs_get_w(self.ptr, world,
                                            arr.ctypes.data_as(C.POINTER(C.c_uint64)), n,
                                            C.byref(cnt)), "pairs")
        return arr[:n]

    def islands(self, world=0):
        cnt = C.c_uint32()
        self.lib.check(self.lib.islands_get_w(self.ptr, world, None, None, 0, C.byref(cnt)), "islands")
        n = cnt.value
        ids = np.zeros(max(n, 1), dtype=np.uint32)
        lab = np.zeros(max(n, 1), dtype=np.uint32)
        self.lib.check(self.lib.islands_get_w(self.ptr, world,
                                              ids.ctypes.data_as(C.POINTER(C.c_uint32)),
                                              lab.ctypes.data_as(C.POINTER(C.c_uint32)), n,
                                              C.byref(cnt)), "islands")
        return ids[:n], lab[:n]

    def tree(self, which, world=0):
        """The static (which=0) or dynamic (1) broad-phase tree under the tree
        pair order: (nodes as a structured array in index order, root)."""
        cnt, root = C.c_uint32(), C.c_uint32()
        self.lib.check(self.lib.tree_get_w(self.ptr, world, which, None, 0, C.byref(cnt),
                                           C.byref(root)), "tree_get")
        arr = (_abi.TreeNode * max(cnt.value, 1))()
        self.lib.check(self.lib.tree_get_w(self.ptr, world, which, arr, cnt.value, C.byref(cnt),
                                           C.byref(root)), "tree_get")
        out = np.frombuffer(arr, dtype=np.dtype([
            ("min", "<f4", 3), ("max", "<f4", 3), ("body_id", "<u4"), ("parent", "<u4"),
            ("child1", "<u4"), ("child2", "<u4"), ("height", "<i4"), ("pad", "<u4")]))
        return out[: cnt.value].copy(), root.value

    def coloring(self, world=0):
        cnt = C.c_uint32()
        self.lib.check(self.lib.coloring_get_w(self.ptr, world, None, 0, C.byref(cnt)), "coloring")
        n = cnt.value
        arr = np.zeros(max(n, 1), dtype=np.uint8)
        self.lib.check(self.lib.coloring_get_w(self.ptr, world,
                                               arr.ctypes.data_as(C.POINTER(C.c_uint8)), n,
                                               C.byref(cnt)), "coloring")
        return arr[:n]

    def setState(self, handle, position=None, rotation=None, linear_velocity=None,
                 angular_velocity=None, world=0):
        def arr(v, n):
            return None if v is None else (C.c_float * n)(*v)
        self.lib.check(self.lib.body_set_state_w(self.ptr, world, handle, arr(position, 3),
                                                 arr(rotation, 4), arr(linear_velocity, 3),
                                                 arr(angular_velocity, 3)), "body_set_state")

    # ---- reference-named mutators (src/system.zig:886-961) ---------------------
    def _op(self, handle, op, v=None, point=None, world=0):
        va = None if v is None else (C.c_float * len(v))(*v)
        pa = None if point is None else (C.c_float * 3)(*point)
        self.lib.check(self.lib.body_op_w(self.ptr, world, handle, op, va, pa), "body_op")

    def applyForce(self, handle, force, world=0):
        self._op(handle, _abi.OP_APPLY_FORCE, force, world=world)

    def applyForceAtPoint(self, handle, force, point, world=0):
        self._op(handle, _abi.OP_APPLY_FORCE_AT_POINT, force, point, world)

    def applyImpulse(self, handle, impulse, world=0):
        self._op(handle, _abi.OP_APPLY_IMPULSE, impulse, world=world)

    def applyImpulseAtPoint(self, handle, impulse, point, world=0):
        self._op(handle, _abi.OP_APPLY_IMPULSE_AT_POINT, impulse, point, world)

    def applyTorque(self, handle, torque, world=0):
        self._op(handle, _abi.OP_APPLY_TORQUE, torque, world=world)

    def applyAngularImpulse(self, handle, impulse, world=0):
        self._op(handle, _abi.OP_APPLY_ANGULAR_IMPULSE, impulse, world=world)

    def setLinearVelocity(self, handle, velocity, world=0):
        self._op(handle, _abi.OP_SET_LINEAR_VELOCITY, velocity, world=world)

    def setAngularVelocity(self, handle, velocity, world=0):
        self._op(handle, _abi.OP_SET_ANGULAR_VELOCITY, velocity, world=world)

    def setPosition(self, handle, position, world=0):
        self._op(handle, _abi.OP_SET_POSITION, position, world=world)

    def setRotation(self, handle, rotation, world=0):
        self._op(handle, _abi.OP_SET_ROTATION, rotation, world=world)

    def wakeUp(self, handle, world=0):
        self._op(handle, _abi.OP_WAKE_UP, world=world)

    def sleep(self, handle, world=0):
        self._op(handle, _abi.OP_SLEEP, world=world)

    # ---- reference-named accessors (src/system.zig:462-473,963-1023) -------------
    def getPosition(self, handle, world=0):
        return tuple(self.bodies(world)[handle.index]["position"])

    def getRotation(self, handle, world=0):
        return tuple(self.bodies(world)[handle.index]["rotation"])

    def getTransform(self, handle, world=0):
        b = self.bodies(world)[handle.index]
        return tuple(b["position"]), tuple(b["rotation"])

    def getLinearVelocity(self, handle, world=0):
        return tuple(self.bodies(world)[handle.index]["linear_velocity"])

    def getAngularVelocity(self, handle, world=0):
        return tuple(self.bodies(world)[handle.index]["angular_velocity"])

    def isSleeping(self, handle, world=0):
        return bool(self.bodies(world)[handle.index]["is_sleeping"])

    def getWorldAABB(self, handle, world=0):
        out = (C.c_float * 6)()
        self.lib.check(self.lib.body_world_aabb_w(self.ptr, world, handle, out), "body_world_aabb")
        return tuple(out[:3]), tuple(out[3:])

    def findBodyById(self, body_id, world=0):
        """The handle of the active body with this id, or None."""
        h = _abi.Handle()
        rc = self.lib.body_find_w(self.ptr, world, body_id, C.byref(h))
        if rc == _abi.VXP_ERR_INVALID_HANDLE:
            return None
        self.lib.check(rc, "body_find")
        return h

    def isBodyInSleepingIsland(self, handle, world=0):
        """isBodyInSleepingIsland (src/system.zig:845-853); False for an invalid handle."""
        out = C.c_int()
        rc = self.lib.body_in_sleeping_island_w(self.ptr, world, handle, C.byref(out))
        if rc == _abi.VXP_ERR_INVALID_HANDLE:
            return False
        self.lib.check(rc, "body_in_sleeping_island")
        return bool(out.value)

    def wakeBodyIsland(self, handle, world=0):
        """wakeBodyIsland (src/system.zig:855-866); an invalid handle is ignored."""
        rc = self.lib.body_wake_island_w(self.ptr, world, handle)
        if rc != _abi.VXP_ERR_INVALID_HANDLE:
            self.lib.check(rc, "body_wake_island")

    def getIslandStats(self, world=0):
        """getIslandStats (src/system.zig:868-873): a dict, or None when islands are off."""
        st = _abi.IslandStats()
        rc = self.lib.island_stats_get_w(self.ptr, world, C.byref(st))
        if rc == _abi.VXP_ERR_ARGUMENT and not self.cfg.enable_islands:
            return None
        self.lib.check(rc, "island_stats_get")
        return st.as_dict()

    def getBodyCount(self, world=0):
        return self.stats(world)["total_bodies"]


def from_scene(scene, lib=None, cfg_over=None, worlds=None):
    """Build a VoxelPhysicsSystem holding a generated scene."""
    from .scenes import build_system
    lib = lib if lib is not None else _abi.product_lib()
    cfg = dict(scene.config)
    if cfg_over:
        cfg.update(cfg_over)
    wl = scene.worlds if worlds is None else [scene.worlds[i] for i in worlds]
    cfg["max_bodies"] = max(cfg.get("max_bodies", 0), max(len(w) for w in wl))
    sysobj = VoxelPhysicsSystem.__new__(VoxelPhysicsSystem)
    sysobj.lib = lib
    sysobj.cfg = _abi.default_config(**cfg)
    sysobj.n_worlds = len(wl)
    ptr, handles = build_system(lib, scene, cfg_over=cfg, worlds=worlds)
    sysobj.ptr = ptr
    sysobj.handles = handles
    return sysobj
