"""Island sleep bookkeeping and its API (src/islands.zig:196-562,
src/system.zig:705-757,845-873): isBodyInSleepingIsland, wakeBodyIsland,
getIslandStats and the island counts of the step statistics.

CPU: the CPU restatement against rules read off observables only — the
rebuild rule (a build is skipped exactly when the manifold key list equals the
one of the previous executed step, the list the last build hashed), the island
statistics against the exported labels, and the sleep timers against the
pre-step velocities of the active bodies in slot order (the list the reference
hands to updateSleepStates, which an island's body id indexes).  A closed
form pins the slot-order reading: with a ground (id 1) and one cube (id 2)
the cube's island reads past the two-body list, so it counts as still and
falls asleep exactly on the 30th float32 accumulation of dt at or above 0.5 s
after its last build, whatever the cube does; the ground's island reads the
cube.  GPU: the HIP pipeline frame for frame against the restatement,
through wakes and a destroy.
"""
import numpy as np
import pytest

import vxp
from vxp import _abi, scenes

F32 = np.float32
TH = F32(0.1) * F32(0.1)  # ISLAND_SLEEP_*_THRESHOLD squared (src/islands.zig:17-22)


def _lsq(v):
    v = v.astype(F32)
    return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]  # Vec3.dot (src/math.zig:93-95)


class IslandModel:
    """The island manager's observable rules, driven by exported state."""

    def __init__(self, dt):
        self.dt = F32(dt)
        self.last_keys = None
        self.need = True
        self.total = 0
        self.performed = self.skipped = 0
        self.timer, self.asleep = {}, {}
        self.counts = (0, 0)  # (active, sleeping) after the last sleep update
        self.label = {}

    def wake(self, body_id):
        lab = self.label.get(body_id)
        if lab is not None:
            self.asleep[lab] = False
            self.timer[lab] = F32(0)
        self.need = True

    def step(self, pre, keys, ids, labels):
        """One executed step: `pre` = bodies before it, keys/ids/labels after."""
        rebuild = self.need or keys != self.last_keys or self.total == 0
        members = {}
        for i, lab in zip(ids.tolist(), labels.tolist()):
            members.setdefault(lab, []).append(i)
        if rebuild:
            self.performed += 1
            self.timer, self.asleep = {}, {}
            self.total = len(members)
            self.label = dict(zip(ids.tolist(), labels.tolist()))
            self.need = False
            pre_counts = (self.total, 0)
        else:
            self.skipped += 1
            pre_counts = self.counts
        self.last_keys = keys
        act = pre[pre["active"] == 1]
        for lab, mem in members.items():
            still = True
            for i in mem:
                if i >= len(act):
                    continue
                b = act[i]
                if b["body_type"] != _abi.DYNAMIC:
                    continue
                if _lsq(b["linear_velocity"]) > TH or _lsq(b["angular_velocity"]) > TH:
                    still = False
            if still:
                t = F32(self.timer.get(lab, F32(0)) + self.dt)
                self.timer[lab] = t
                if t >= F32(0.5):
                    self.asleep[lab] = True
            else:
                self.timer[lab] = F32(0)
                self.asleep[lab] = False
        n_sleep = sum(1 for lab in members if self.asleep.get(lab, False))
        self.counts = (self.total - n_sleep, n_sleep)
        return pre_counts, members

    def sleeping_of(self, body_id):
        lab = self.label.get(body_id)
        return lab is not None and self.asleep.get(lab, False)


def _keys(s, w=0):
    return tuple((min(m.body_a_id, m.body_b_id), max(m.body_a_id, m.body_b_id))
                 for m in s.manifolds(w))


def _run_against_model(s, dt, frames, actions=None, w=0):
    """Step `s`, checking every observable against IslandModel; returns the
    frames at which some island was asleep."""
    model = IslandModel(dt)
    asleep_frames = []
    for f in range(frames):
        if actions and f in actions:
            actions[f](s, model)
        pre = s.bodies(w)
        s.update(dt)
        st = s.stats(w)
        if st["steps_executed"] == 0:
            continue
        ids, labels = s.islands(w)
        (pa, ps), members = model.step(pre, _keys(s, w), ids, labels)
        assert (st["island_count"], st["active_islands"], st["sleeping_islands"]) == \
            (model.total, pa, ps), f"frame {f}: step stats"
        ist = s.getIslandStats(w)
        sizes = [len(m) for m in members.values()]
        want = dict(total_islands=model.total, active_islands=model.counts[0],
                    sleeping_islands=model.counts[1], rebuilds_skipped=model.skipped,
                    rebuilds_performed=model.performed)
        for k, v in want.items():
            assert ist[k] == v, f"frame {f}: {k} {ist[k]} vs {v}"
        assert ist["largest_island_size"] == (max(sizes) if sizes else 0)
        assert ist["total_bodies"] == len(ids)
        now = s.bodies(w)
        for slot in np.nonzero(now["active"])[0]:
            h = _abi.Handle(int(slot), _gen(s, w, slot))
            assert s.isBodyInSleepingIsland(h, w) == model.sleeping_of(int(now["id"][slot])), \
                f"frame {f}: body slot {slot}"
        if model.counts[1]:
            asleep_frames.append(f)
    return asleep_frames


def _gen(s, w, slot):
    for h in s.handles[w]:
        if h.index == slot:
            return h.generation
    return 0


def _two_body(lib, height=0.6):
    vs = 0.2
    sc = scenes.Scene("ground_cube", scenes.scene_config("g", 1 / 60, 4, 8, 4), 1 / 60, 200)
    g = scenes._add_collider(sc, *scenes.hand_ground(16, 16), vs, "ground16")
    c = scenes._add_collider(sc, *scenes.readme_cube(), vs, "cube3")
    sc.worlds.append([scenes.BodySpec(_abi.STATIC, (0.0, -vs * 0.5, 0.0), (0, 0, 0, 1), g),
                      scenes.BodySpec(_abi.DYNAMIC, (0.0, height, 0.0), (0, 0, 0, 1), c, 1.0)])
    return sc, vxp.from_scene(sc, lib=lib)


def test_rebuild_rule_stats_and_timers(oracle):
    sc = scenes.c0_small(n_side=2, frames=200)
    s = vxp.from_scene(sc, lib=oracle)
    asleep = _run_against_model(s, sc.dt, 200)
    assert asleep, "no island fell asleep in 200 frames"
    s.close()


def test_wake_and_destroy(oracle):
    sc = scenes.c0_small(n_side=3, frames=240)
    s = vxp.from_scene(sc, lib=oracle)

    def wake(slot):
        def act(sysobj, model):
            h = sysobj.handles[0][slot]
            sysobj.wakeBodyIsland(h)
            model.wake(int(sysobj.bodies()["id"][slot]))
            assert not sysobj.isBodyInSleepingIsland(h)
        return act

    def destroy(slot):
        def act(sysobj, model):
            sysobj.destroyBody(sysobj.handles[0][slot])
        return act

    # the destroy shifts every later body down the active list
    acts = {120: wake(3), 150: destroy(2), 180: wake(5), 200: wake(1)}
    _run_against_model(s, sc.dt, 240, acts)
    s.close()


@pytest.mark.parametrize("push", [False, True])
def test_slot_order_quirk_closed_form(oracle, push):
    """push: from frame 30 the cube is kept sliding at 0.5 m/s; its own island
    still falls asleep (it never reads the cube), the ground's never does."""
    sc, s = _two_body(oracle)
    ground, cube = s.handles[0][0], s.handles[0][1]
    dt = F32(sc.dt)
    prev_keys, prev_perf = None, 0
    c_acc = g_acc = F32(0)
    contact_frames = 0
    for f in range(120 if push else 200):
        if push and f >= 30:
            vy = float(s.bodies()[1]["linear_velocity"][1])
            s.setLinearVelocity(cube, (0.5, vy, 0.0))
        v = s.bodies()[1]
        cube_fast = _lsq(v["linear_velocity"]) > TH or _lsq(v["angular_velocity"]) > TH
        s.update(sc.dt)
        if s.stats()["steps_executed"] == 0:
            break
        ist = s.getIslandStats()
        keys = _keys(s)
        rebuilt = ist["rebuilds_performed"] != prev_perf
        assert rebuilt == (not prev_keys or keys != prev_keys), f"frame {f}"
        prev_perf, prev_keys = ist["rebuilds_performed"], keys
        if rebuilt:
            c_acc = g_acc = F32(0)
        if not keys:
            assert ist["total_islands"] == 0
            assert not s.isBodyInSleepingIsland(cube) and not s.isBodyInSleepingIsland(ground)
            continue
        contact_frames += 1
        # two islands: the edge to the static ground unites nothing
        assert ist["total_islands"] == 2 and ist["largest_island_size"] == 1
        # island {2} reads index 2 of a two-body list: passed over, always still
        c_acc = F32(c_acc + dt)
        # island {1} reads index 1, the cube
        g_acc = F32(0) if cube_fast else F32(g_acc + dt)
        assert s.isBodyInSleepingIsland(cube) == bool(c_acc >= F32(0.5)), f"frame {f}"
        assert s.isBodyInSleepingIsland(ground) == bool(g_acc >= F32(0.5)), f"frame {f}"
    assert contact_frames >= 30 and s.isBodyInSleepingIsland(cube)
    if push:
        assert not s.isBodyInSleepingIsland(ground) and not s.bodies()[1]["is_sleeping"]
    s.close()


def test_islands_disabled(oracle):
    sc = scenes.c0_small(n_side=2, frames=10)
    s = vxp.from_scene(sc, lib=oracle, cfg_over={"enable_islands": 0})
    for _ in range(40):
        s.update(sc.dt)
    assert s.getIslandStats() is None
    assert not s.isBodyInSleepingIsland(s.handles[0][1])
    st = s.stats()
    assert st["island_count"] == st["active_islands"] == st["sleeping_islands"] == 0
    s.close()


def test_invalid_handles(oracle):
    sc, s = _two_body(oracle)
    h = s.handles[0][1]
    s.destroyBody(h)
    assert s.isBodyInSleepingIsland(h) is False
    s.wakeBodyIsland(h)  # ignored
    assert s.lib.body_wake_island_w(s.ptr, 0, h) == _abi.VXP_ERR_INVALID_HANDLE
    s.close()


@pytest.mark.gpu
def test_island_sleep_gpu_matches_restatement(product, oracle):
    sc = scenes.c0_small(n_side=3, frames=240)
    sp = vxp.from_scene(sc, lib=product)
    so = vxp.from_scene(sc, lib=oracle)
    wakes = {90: 4, 130: 1, 170: 7}
    for f in range(240):
        if f in wakes:
            for s in (sp, so):
                s.wakeBodyIsland(s.handles[0][wakes[f]])
        if f == 150:
            for s in (sp, so):
                s.destroyBody(s.handles[0][2])
        sp.update(sc.dt)
        so.update(sc.dt)
        a, b = sp.stats(), so.stats()
        for k in ("island_count", "active_islands", "sleeping_islands", "largest_island_size"):
            assert a[k] == b[k], f"frame {f}: {k} {a[k]} vs {b[k]}"
        assert sp.getIslandStats() == so.getIslandStats(), f"frame {f}"
        for h in so.handles[0]:
            assert sp.isBodyInSleepingIsland(h) == so.isBodyInSleepingIsland(h), f"frame {f} slot {h.index}"
    assert so.getIslandStats()["sleeping_islands"] > 0
    sp.close()
    so.close()
