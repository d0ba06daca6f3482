"""Python binding of the serving loop (vxp_server_* in include/vxp.h).

A Server keeps many worlds resident in device memory and advances the ones
with pending requests together, one frame per tick (continuous batching):
requests that arrive while others are on the GPU join the batch as slots free
up.  The loop itself (queue, admission, slot moves, HBM budget) is native; this
class only forwards calls and converts results.
"""
import ctypes as C

import numpy as np

from . import _abi

DONE_DTYPE = np.dtype([("tag", "<u8"), ("world", "<u4"), ("frames", "<u4"), ("submit_s", "<f8"),
                       ("admit_s", "<f8"), ("done_s", "<f8"), ("body_substeps", "<u8")])


class Server:
    def __init__(self, config, max_worlds, slots, budget_bytes, lib=None):
        self.lib = lib if lib is not None else _abi.product_lib()
        self.cfg = config if isinstance(config, _abi.Config) else _abi.default_config(**config)
        self.ptr = C.c_void_p()
        self.lib.check(self.lib.server_create(C.byref(self.cfg), max_worlds, slots,
                                              int(budget_bytes), C.byref(self.ptr)), "server_create")

    def close(self):
        if self.ptr:
            self.lib.server_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_collider(self, voxel_size, brick_pos, voxels):
        bp = np.ascontiguousarray(brick_pos, dtype=np.int32)
        vx = np.ascontiguousarray(voxels, dtype=np.uint8)
        out = C.c_uint32()
        self.lib.check(self.lib.server_collider_create(self.ptr, voxel_size, len(bp),
                                                       bp.ctypes.data_as(C.POINTER(C.c_int32)),
                                                       vx.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                       C.byref(out)), "server_collider_create")
        return out.value

    def create_world(self):
        out = C.c_uint32()
        self.lib.check(self.lib.server_world_create(self.ptr, C.byref(out)), "server_world_create")
        return out.value

    def create_body(self, world, body_type, position, rotation, collider, density=1.0):
        h = _abi.Handle()
        self.lib.check(self.lib.server_body_create(self.ptr, world, body_type,
                                                   (C.c_float * 3)(*position),
                                                   (C.c_float * 4)(*rotation), collider, density,
                                                   C.byref(h)), "server_body_create")
        return h

    def submit(self, world, frames, tag, submit_s=-1.0):
        self.lib.check(self.lib.server_submit(self.ptr, world, frames, tag, submit_s), "server_submit")

    def set_recorded(self, on=True):
        """Step the batch from recorded HIP graphs (vxp_server_graph_enable)."""
        self.lib.check(self.lib.server_graph_enable(self.ptr, 1 if on else 0), "server_graph_enable")

    def graph_info(self):
        g = _abi.GraphInfo()
        self.lib.check(self.lib.server_graph_info_get(self.ptr, C.byref(g)), "server_graph_info_get")
        return {f: getattr(g, f) for f, _ in g._fields_}

    def tick(self):
        n = C.c_uint32()
        self.lib.check(self.lib.server_tick(self.ptr, C.byref(n)), "server_tick")
        return n.value

    def poll(self, cap=4096):
        """Retired requests as a structured array (oldest first)."""
        out = []
        buf = (_abi.RequestDone * cap)()
        while True:
            n = C.c_uint32()
            self.lib.check(self.lib.server_poll(self.ptr, buf, cap, C.byref(n)), "server_poll")
            if n.value:
                out.append(np.frombuffer(bytes(buf)[: n.value * C.sizeof(_abi.RequestDone)],
                                         dtype=DONE_DTYPE).copy())
            if n.value < cap:
                break
        return np.concatenate(out) if out else np.zeros(0, dtype=DONE_DTYPE)

    def info(self):
        i = _abi.ServerInfo()
        self.lib.check(self.lib.server_info_get(self.ptr, C.byref(i)), "server_info_get")
        return i.as_dict()

    def bodies(self, world):
        n = self.cfg.max_bodies
        arr = (_abi.BodyState * n)()
        cnt = C.c_uint32()
        self.lib.check(self.lib.server_bodies_get(self.ptr, world, arr, n, C.byref(cnt)),
                       "server_bodies_get")
        return np.frombuffer(bytes(arr), dtype=np.uint8).reshape(n, C.sizeof(_abi.BodyState))[: cnt.value]

    def stats(self, world):
        s = _abi.Stats()
        self.lib.check(self.lib.server_stats_get(self.ptr, world, C.byref(s)), "server_stats_get")
        return s.as_dict()

    def manifold_array(self, world):
        cnt = C.c_uint32()
        self.lib.check(self.lib.server_manifolds_get(self.ptr, world, None, 0, C.byref(cnt)),
                       "server_manifolds_get")
        n = cnt.value
        arr = (_abi.Manifold * max(n, 1))()
        self.lib.check(self.lib.server_manifolds_get(self.ptr, world, arr, n, C.byref(cnt)),
                       "server_manifolds_get")
        return np.frombuffer(bytes(arr), dtype=np.uint8)[: n * C.sizeof(_abi.Manifold)].reshape(
            n, C.sizeof(_abi.Manifold))


def from_scene(scene, slots, budget_bytes, lib=None, max_worlds=None, cfg_over=None):
    """A Server holding every world of `scene` (resident, idle)."""
    cfg = dict(scene.config)
    if cfg_over:
        cfg.update(cfg_over)
    cfg["max_bodies"] = max(cfg.get("max_bodies", 0), max(len(w) for w in scene.worlds))
    srv = Server(cfg, max_worlds or len(scene.worlds), slots, budget_bytes, lib=lib)
    cids = [srv.add_collider(c.voxel_size, c.brick_pos, c.voxels) for c in scene.colliders]
    for bodies in scene.worlds:
        w = srv.create_world()
        for b in bodies:
            srv.create_body(w, b.body_type, b.position, b.rotation, cids[b.collider], b.density)
    return srv
