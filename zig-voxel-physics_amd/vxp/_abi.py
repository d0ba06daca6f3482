"""ctypes mirror of include/vxp.h and the library loader.

Two libraries export the same C ABI: the product (`libvxp.so`, prefix
``vxp_``, HIP kernels on the GPU) and the CPU restatement used as the checker
(`liboracle`, prefix ``vxo_``, test infrastructure).  `Lib` binds either one
by prefix; the product path never falls back to the CPU one.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
PRODUCT_SO = os.path.join(PKG_DIR, "libvxp.so")
ORACLE_SO = os.path.join(REPO_DIR, "oracle", "_build", "libvxp_oracle.so")

VXP_OK = 0
VXP_ERR_INVALID_HANDLE, VXP_ERR_CAPACITY, VXP_ERR_OOM, VXP_ERR_DEVICE, VXP_ERR_ARGUMENT = -1, -2, -3, -4, -5
STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
EMPTY, CORNER, EDGE, FACE, INSIDE = 0, 1, 2, 3, 4
SOLVE_ORDER_COLOR, SOLVE_ORDER_CANONICAL = 0, 1
PAIR_ORDER_KEY, PAIR_ORDER_TREE = 0, 1
PROFILE_OFF, PROFILE_ALL, PROFILE_SOLVE_UNITS = 0, 1, 2
NO_COLLIDER = 0xFFFFFFFF


def voxel_byte(vtype, normal=0):
    """CollisionVoxel byte: u3 type | u5 normal index << 3 (src/collision_voxel.zig:133-191)."""
    return (vtype & 7) | ((normal & 31) << 3)


class Config(C.Structure):
    _fields_ = [
        ("gravity", C.c_float * 3),
        ("fixed_timestep", C.c_float),
        ("max_substeps", C.c_uint32),
        ("max_bodies", C.c_uint32),
        ("default_friction", C.c_float),
        ("default_restitution", C.c_float),
        ("default_density", C.c_float),
        ("enable_sleeping", C.c_uint8),
        ("enable_warm_starting", C.c_uint8),
        ("enable_islands", C.c_uint8),
        ("enable_coloring", C.c_uint8),
        ("substeps", C.c_uint32),
        ("velocity_iterations", C.c_uint32),
        ("position_iterations", C.c_uint32),
        ("hertz", C.c_float),
        ("damping_ratio", C.c_float),
        ("position_correction_rate", C.c_float),
        ("max_position_correction", C.c_float),
        ("slop", C.c_float),
        ("warm_start_scale", C.c_float),
        ("disable_rotation", C.c_uint8),
        ("warm_starting", C.c_uint8),
        ("enable_friction", C.c_uint8),
        ("enable_restitution", C.c_uint8),
        ("use_brick_sat", C.c_uint8),
        ("use_brick_pair_cache", C.c_uint8),
        ("enable_threading", C.c_uint8),
        ("pad0", C.c_uint8),
        ("brick_pair_cache_size", C.c_uint32),
        ("pair_cap", C.c_uint32),
        ("solve_order", C.c_uint32),
        ("cpu_threads", C.c_uint32),
        ("pair_order", C.c_uint32),
        ("thread_count", C.c_uint32),
    ]


def default_config(**over):
    """PhysicsConfig{} + SolverConfig{} + NarrowPhaseConfig{} defaults
    (src/system.zig:47-88, src/solver.zig:75-119, src/narrow_phase.zig:1325-1343)."""
    c = Config()
    c.gravity[:] = (0.0, -9.81, 0.0)
    c.fixed_timestep = 1.0 / 60.0
    c.max_substeps = 8
    c.max_bodies = 1024
    c.default_friction = 0.6
    c.default_restitution = 0.2
    c.default_density = 1.0
    c.enable_sleeping = 1
    c.enable_warm_starting = 1
    c.enable_islands = 1
    c.enable_coloring = 0
    c.substeps = 4
    c.velocity_iterations = 1
    c.position_iterations = 2
    c.hertz = 30.0
    c.damping_ratio = 1.0
    c.position_correction_rate = 0.2
    c.max_position_correction = 0.04
    c.slop = 0.005
    c.warm_start_scale = 0.8
    c.disable_rotation = 0
    c.warm_starting = 1
    c.enable_friction = 1
    c.enable_restitution = 1
    c.use_brick_sat = 1
    c.use_brick_pair_cache = 1
    c.brick_pair_cache_size = 1024
    c.pair_cap = 0
    c.solve_order = SOLVE_ORDER_COLOR
    c.cpu_threads = 1
    c.pair_order = PAIR_ORDER_KEY
    for k, v in over.items():
        if k == "gravity":
            c.gravity[:] = v
        else:
            setattr(c, k, v)
    return c


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "total_bodies", "dynamic_bodies", "static_bodies", "sleeping_bodies", "active_bodies",
        "broad_phase_pairs", "narrow_phase_tested", "narrow_phase_colliding",
        "contacts_generated", "voxels_checked", "pairs_skipped_sleeping",
        "cache_hits", "cache_misses", "manifolds_skipped_sleeping",
        "island_count", "largest_island_size", "substeps_this_frame",
        "manifolds", "solver_manifolds", "colors", "position_iterations_run",
        "steps_executed", "awake_bodies_stepped", "solver_points")] + [
        ("max_penetration", C.c_float), ("accumulated_time", C.c_float),
        ("contact_candidates", C.c_uint32), ("active_islands", C.c_uint32),
        ("sleeping_islands", C.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class IslandStats(C.Structure):
    """IslandManager.IslandStats (src/islands.zig:222-243)."""
    _fields_ = [(n, C.c_uint32) for n in (
        "total_islands", "active_islands", "sleeping_islands", "largest_island_size",
        "total_bodies", "total_contacts", "rebuilds_skipped", "rebuilds_performed")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TreeNode(C.Structure):
    """TreeNode (src/broad_phase.zig:40-62) as vxp_tree_get_w exports it."""
    _fields_ = [("min", C.c_float * 3), ("max", C.c_float * 3), ("body_id", C.c_uint32),
                ("parent", C.c_uint32), ("child1", C.c_uint32), ("child2", C.c_uint32),
                ("height", C.c_int32), ("pad", C.c_uint32)]


class Contact(C.Structure):
    _fields_ = [
        ("point", C.c_float * 3),
        ("normal", C.c_float * 3),
        ("penetration", C.c_float),
        ("local_anchor_a", C.c_float * 3),
        ("local_anchor_b", C.c_float * 3),
        ("normal_impulse", C.c_float),
        ("tangent_impulse", C.c_float * 2),
        ("normal_mass", C.c_float),
        ("tangent_mass", C.c_float * 2),
        ("pad", C.c_uint32),
        ("feature_id", C.c_uint64),
    ]


class Manifold(C.Structure):
    _fields_ = [
        ("body_a_id", C.c_uint32),
        ("body_b_id", C.c_uint32),
        ("friction", C.c_float),
        ("restitution", C.c_float),
        ("contact_count", C.c_uint32),
        ("color", C.c_uint32),
        ("contacts", Contact * 4),
    ]


class Handle(C.Structure):
    _fields_ = [("index", C.c_uint32), ("generation", C.c_uint32)]


class BodyState(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("rotation", C.c_float * 4),
        ("linear_velocity", C.c_float * 3),
        ("angular_velocity", C.c_float * 3),
        ("sleep_timer", C.c_float),
        ("id", C.c_uint32),
        ("body_type", C.c_uint8),
        ("is_sleeping", C.c_uint8),
        ("active", C.c_uint8),
        ("pad", C.c_uint8),
    ]


class RequestDone(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("world", C.c_uint32), ("frames", C.c_uint32),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("body_substeps", C.c_uint64)]


class ServerInfo(C.Structure):
    _fields_ = [("worlds", C.c_uint32), ("capacity", C.c_uint32), ("slots", C.c_uint32),
                ("occupied", C.c_uint32), ("queued", C.c_uint32), ("pad", C.c_uint32),
                ("ticks", C.c_uint64), ("completed", C.c_uint64), ("body_substeps", C.c_uint64),
                ("resident_bytes", C.c_uint64), ("budget_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class Tuning(C.Structure):
    _fields_ = [("pair_color_max", C.c_uint32), ("windows", C.c_uint32), ("fixed", C.c_uint32),
                ("unit_threads", C.c_uint32), ("broad_phase", C.c_uint32),
                ("probe_threads", C.c_uint32), ("probe_xcd_run", C.c_uint32)]


TUNE_AUTO = 0xFFFFFFFF

# body mutators of vxp_body_op_w (include/vxp.h)
OP_APPLY_FORCE, OP_APPLY_FORCE_AT_POINT, OP_APPLY_IMPULSE, OP_APPLY_IMPULSE_AT_POINT, \
    OP_APPLY_TORQUE, OP_APPLY_ANGULAR_IMPULSE, OP_SET_LINEAR_VELOCITY, OP_SET_ANGULAR_VELOCITY, \
    OP_SET_POSITION, OP_SET_ROTATION, OP_WAKE_UP, OP_SLEEP = range(12)

assert C.sizeof(Config) == 108, C.sizeof(Config)
assert C.sizeof(Contact) == 88, C.sizeof(Contact)
assert C.sizeof(Manifold) == 376, C.sizeof(Manifold)
assert C.sizeof(BodyState) == 64, C.sizeof(BodyState)

_P = C.c_void_p
_f3 = C.POINTER(C.c_float)

# name -> (restype, argtypes) without prefix; every function exists in both libraries
SIGNATURES = {
    "system_create": (C.c_int, [C.POINTER(Config), C.c_uint32, C.POINTER(_P)]),
    "system_destroy": (C.c_int, [_P]),
    "collider_create": (C.c_int, [_P, C.c_float, C.c_uint32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)]),
    "body_create": (C.c_int, [_P, C.c_uint32, _f3, _f3, C.c_uint32, C.c_float,
                              C.POINTER(Handle)]),
    "body_create_w": (C.c_int, [_P, C.c_uint32, C.c_uint32, _f3, _f3, C.c_uint32, C.c_float,
                                C.POINTER(Handle)]),
    "body_create_with_mass_w": (C.c_int, [_P, C.c_uint32, _f3, _f3, C.c_uint32, C.c_float, _f3,
                                          C.POINTER(Handle)]),
    "system_set_config": (C.c_int, [_P, C.POINTER(Config)]),
    "body_destroy_w": (C.c_int, [_P, C.c_uint32, Handle]),
    "collider_create_solid": (C.c_int, [_P, C.c_float, C.c_uint32, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "collider_info": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                C.POINTER(C.c_uint32)]),
    "collider_mass_properties": (C.c_int, [_P, C.c_uint32, C.c_float, C.POINTER(C.c_float)]),
    "collider_bricks": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_uint8),
                                  C.c_uint32, C.POINTER(C.c_uint32)]),
    "body_op_w": (C.c_int, [_P, C.c_uint32, Handle, C.c_uint32, C.POINTER(C.c_float),
                            C.POINTER(C.c_float)]),
    "body_world_aabb_w": (C.c_int, [_P, C.c_uint32, Handle, C.POINTER(C.c_float)]),
    "body_find_w": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(Handle)]),
    "system_update": (C.c_int, [_P, C.c_float]),
    "system_step": (C.c_int, [_P, C.c_float]),
    "system_update_frames": (C.c_int, [_P, C.c_float, C.c_uint32]),
    "bodies_get_w": (C.c_int, [_P, C.c_uint32, C.POINTER(BodyState), C.c_uint32,
                               C.POINTER(C.c_uint32)]),
    "body_set_state_w": (C.c_int, [_P, C.c_uint32, Handle, _f3, _f3, _f3, _f3]),
    "stats_get_w": (C.c_int, [_P, C.c_uint32, C.POINTER(Stats)]),
    "stats_get_all": (C.c_int, [_P, C.POINTER(Stats), C.c_uint32]),
    "manifolds_get_w": (C.c_int, [_P, C.c_uint32, C.POINTER(Manifold), C.c_uint32,
                                  C.POINTER(C.c_uint32)]),
    "pairs_get_w": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint64), C.c_uint32,
                              C.POINTER(C.c_uint32)]),
    "islands_get_w": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.c_uint32, C.POINTER(C.c_uint32)]),
    "coloring_get_w": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint8), C.c_uint32,
                                 C.POINTER(C.c_uint32)]),
    "island_stats_get_w": (C.c_int, [_P, C.c_uint32, C.POINTER(IslandStats)]),
    "tree_get_w": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(TreeNode), C.c_uint32,
                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "body_in_sleeping_island_w": (C.c_int, [_P, C.c_uint32, Handle, C.POINTER(C.c_int)]),
    "body_wake_island_w": (C.c_int, [_P, C.c_uint32, Handle]),
    "last_error": (C.c_char_p, []),
}

# product-only exports (declared in include/vxp.h)
class GraphInfo(C.Structure):
    _fields_ = [("recorded_steps", C.c_uint64), ("redos", C.c_uint64), ("recordings", C.c_uint64),
                ("recording_ms", C.c_double), ("cap_pairs", C.c_uint32), ("cap_world_pairs", C.c_uint32),
                ("cap_work_items", C.c_uint32), ("cap_candidates", C.c_uint32)]


PRODUCT_ONLY = {
    "config_default": (C.c_int, [C.POINTER(Config)]),
    # recorded step (HIP graphs)
    "graph_enable": (C.c_int, [_P, C.c_int]),
    "graph_info_get": (C.c_int, [_P, C.POINTER(GraphInfo)]),
    "server_graph_enable": (C.c_int, [_P, C.c_int]),
    "server_graph_info_get": (C.c_int, [_P, C.POINTER(GraphInfo)]),
    "profile_enable": (C.c_int, [_P, C.c_int]),
    "profile_get": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "profile_solver_phases": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint64)]),
    "tuning_get": (C.c_int, [_P, C.POINTER(Tuning)]),
    "tuning_set": (C.c_int, [_P, C.POINTER(Tuning)]),
    # serving (continuous batching over resident worlds)
    "server_create": (C.c_int, [C.POINTER(Config), C.c_uint32, C.c_uint32, C.c_uint64,
                                C.POINTER(_P)]),
    "server_destroy": (C.c_int, [_P]),
    "server_collider_create": (C.c_int, [_P, C.c_float, C.c_uint32, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)]),
    "server_world_create": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "server_body_create": (C.c_int, [_P, C.c_uint32, C.c_uint32, _f3, _f3, C.c_uint32, C.c_float,
                                     C.POINTER(Handle)]),
    "server_submit": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint64, C.c_double]),
    "server_tick": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "server_poll": (C.c_int, [_P, C.POINTER(RequestDone), C.c_uint32, C.POINTER(C.c_uint32)]),
    "server_bodies_get": (C.c_int, [_P, C.c_uint32, C.POINTER(BodyState), C.c_uint32,
                                    C.POINTER(C.c_uint32)]),
    "server_stats_get": (C.c_int, [_P, C.c_uint32, C.POINTER(Stats)]),
    "server_manifolds_get": (C.c_int, [_P, C.c_uint32, C.POINTER(Manifold), C.c_uint32,
                                       C.POINTER(C.c_uint32)]),
    "server_info_get": (C.c_int, [_P, C.POINTER(ServerInfo)]),
}
KERNEL_NAMES = ["step_begin", "fat_aabb", "pairs", "pair_process", "cache_replay", "probe",
                "reduce", "compact", "islands", "solve_prep", "solve_units", "solve_tail"]


class VxpError(RuntimeError):
    pass


class Lib:
    """Binds one library's C ABI by prefix ('vxp' product, 'vxo' oracle)."""

    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise VxpError(f"{path} is not built (run __graft_entry__.build())")
        self.path = path
        self.prefix = prefix
        self.dll = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(self.dll, f"{prefix}_{name}")
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)
        if prefix == "vxp":
            for name, (res, args) in PRODUCT_ONLY.items():
                fn = getattr(self.dll, f"{prefix}_{name}")
                fn.restype = res
                fn.argtypes = args
                setattr(self, name, fn)

    def check(self, rc, what=""):
        if rc != VXP_OK:
            msg = self.last_error()
            raise VxpError(f"{self.prefix}_{what} failed rc={rc}: {msg.decode() if msg else ''}")


_cache = {}


def product_lib():
    """The HIP pipeline.  Raises if libvxp.so is missing: there is no CPU fallback."""
    if "vxp" not in _cache:
        _cache["vxp"] = Lib(PRODUCT_SO, "vxp")
    return _cache["vxp"]


def oracle_lib():
    """The CPU restatement (checker / CPU baseline only)."""
    if "vxo" not in _cache:
        _cache["vxo"] = Lib(ORACLE_SO, "vxo")
    return _cache["vxo"]
