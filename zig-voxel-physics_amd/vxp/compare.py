"""Side-by-side stepping of two systems (e.g. the HIP pipeline and the CPU
restatement) with per-phase comparison: pair set, manifolds, body state."""
import numpy as np

from . import _abi
from .system import from_scene

BODY_FIELDS = ("position", "rotation", "linear_velocity", "angular_velocity", "sleep_timer",
               "is_sleeping", "active", "id")


def manifold_fields(ms):
    """Manifolds -> dict of numpy arrays (contacts padded to 4)."""
    n = len(ms)
    out = {
        "ids": np.zeros((n, 2), np.uint32),
        "count": np.zeros(n, np.uint32),
        "color": np.zeros(n, np.uint32),
        "fr": np.zeros((n, 2), np.float32),
        "fid": np.zeros((n, 4), np.uint64),
        "f": np.zeros((n, 4, 19), np.float32),
    }
    for i, m in enumerate(ms):
        out["ids"][i] = (m.body_a_id, m.body_b_id)
        out["count"][i] = m.contact_count
        out["color"][i] = m.color
        out["fr"][i] = (m.friction, m.restitution)
        for k in range(m.contact_count):
            c = m.contacts[k]
            out["fid"][i, k] = c.feature_id
            out["f"][i, k] = (list(c.point) + list(c.normal) + [c.penetration] +
                              list(c.local_anchor_a) + list(c.local_anchor_b) + [c.normal_impulse] +
                              list(c.tangent_impulse) + [c.normal_mass] + list(c.tangent_mass))
    return out


def _neq_exact(x, y):
    """IEEE inequality with NaN == NaN: exact up to the sign of zero, which the
    reference leaves unspecified (Zig @min/@max lower to minnum/maxnum)."""
    return ~((x == y) | (np.isnan(x) & np.isnan(y)))


def diff_bodies(a, b, rtol=0.0, atol=0.0):
    """First differing (slot, field) or None.  rtol/atol = 0 means exact."""
    for f in BODY_FIELDS:
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            if rtol == 0 and atol == 0:
                bad = np.nonzero(_neq_exact(x, y).reshape(len(x), -1).any(axis=1))[0]
            else:
                bad = np.nonzero((~np.isclose(x, y, rtol=rtol, atol=atol)).reshape(len(x), -1).any(axis=1))[0]
        else:
            bad = np.nonzero((x != y).reshape(len(x), -1).any(axis=1))[0]
        if len(bad):
            i = int(bad[0])
            return f"body slot {i} field {f}: {x[i]} vs {y[i]} ({len(bad)} slots differ)"
    return None


def diff_manifolds(ma, mb, rtol=0.0, atol=0.0):
    if len(ma) != len(mb):
        ka = {(m.body_a_id, m.body_b_id) for m in ma}
        kb = {(m.body_a_id, m.body_b_id) for m in mb}
        return f"manifold count {len(ma)} vs {len(mb)}; only-a {sorted(ka - kb)[:5]} only-b {sorted(kb - ka)[:5]}"
    if not ma:
        return None
    a, b = manifold_fields(ma), manifold_fields(mb)
    for k in ("ids", "count", "color", "fid"):
        bad = np.nonzero((a[k] != b[k]).reshape(len(ma), -1).any(axis=1))[0]
        if len(bad):
            i = int(bad[0])
            return f"manifold {i} {tuple(a['ids'][i])} field {k}: {a[k][i]} vs {b[k][i]}"
    for k in ("fr", "f"):
        x, y = a[k], b[k]
        if rtol == 0 and atol == 0:
            bad = np.nonzero(_neq_exact(x, y).reshape(len(ma), -1).any(axis=1))[0]
        else:
            bad = np.nonzero((~np.isclose(x, y, rtol=rtol, atol=atol)).reshape(len(ma), -1).any(axis=1))[0]
        if len(bad):
            i = int(bad[0])
            j = np.nonzero(_neq_exact(x[i], y[i]).reshape(-1))[0][:4]
            return (f"manifold {i} {tuple(a['ids'][i])} field {k} idx {j.tolist()}: "
                    f"{x[i].reshape(-1)[j]} vs {y[i].reshape(-1)[j]}")
    return None


STAT_KEYS = ("broad_phase_pairs", "narrow_phase_tested", "narrow_phase_colliding",
             "contacts_generated", "voxels_checked", "pairs_skipped_sleeping", "cache_hits",
             "cache_misses", "manifolds_skipped_sleeping", "island_count", "largest_island_size",
             "active_islands", "sleeping_islands",
             "substeps_this_frame", "manifolds", "solver_manifolds", "colors",
             "position_iterations_run", "steps_executed", "awake_bodies_stepped",
             "contact_candidates",
             "solver_points", "sleeping_bodies", "active_bodies",
             "total_bodies", "dynamic_bodies", "static_bodies")


def diff_stats(sa, sb):
    for k in STAT_KEYS:
        if sa[k] != sb[k]:
            return f"stat {k}: {sa[k]} vs {sb[k]}"
    for k in ("max_penetration", "accumulated_time"):
        if sa[k] != sb[k]:
            return f"stat {k}: {sa[k]} vs {sb[k]}"
    return None


def lockstep(scene, lib_a, lib_b, frames, cfg_over=None, worlds=None, check_every=1,
             rtol=0.0, atol=0.0, log=None, dts=None, tune_a=None, prep_a=None):
    """Step both systems `frames` times; return (first_frame_with_diff, message) or (None, None).
    `dts`: the update() argument of frame f is dts[f % len(dts)] (default: the
    scene's dt, one fixed step per frame).  `tune_a`: setTuning arguments of
    the first system (the product); `prep_a`: called with the first system
    before the first frame."""
    sa = from_scene(scene, lib=lib_a, cfg_over=cfg_over, worlds=worlds)
    sb = from_scene(scene, lib=lib_b, cfg_over=cfg_over, worlds=worlds)
    if tune_a:
        sa.setTuning(**tune_a)
    if prep_a:
        prep_a(sa)
    nwl = sa.n_worlds
    for f in range(frames):
        dt = dts[f % len(dts)] if dts else scene.dt
        sa.update(dt)
        sb.update(dt)
        if (f + 1) % check_every and f != frames - 1:
            continue
        for w in range(nwl):
            msg = None
            pa, pb = sa.pairs(w), sb.pairs(w)
            if len(pa) != len(pb) or (pa != pb).any():
                msg = f"pairs differ: {len(pa)} vs {len(pb)}"
            msg = msg or diff_manifolds(sa.manifolds(w), sb.manifolds(w), rtol, atol)
            msg = msg or diff_bodies(sa.bodies(w), sb.bodies(w), rtol, atol)
            msg = msg or diff_stats(sa.stats(w), sb.stats(w))
            if not msg and sa.cfg.enable_islands:
                ia, ib = sa.islands(w), sb.islands(w)
                if len(ia[0]) != len(ib[0]) or (ia[0] != ib[0]).any() or (ia[1] != ib[1]).any():
                    msg = f"island labels differ ({len(ia[0])} vs {len(ib[0])} bodies)"
            if not msg and sa.cfg.enable_coloring:
                ca, cb = sa.coloring(w), sb.coloring(w)
                if len(ca) != len(cb) or (ca != cb).any():
                    msg = "reference colouring differs"
            if msg:
                return f, f"world {w}: {msg}", (sa, sb)
        if log:
            log(f)
    return None, None, (sa, sb)
