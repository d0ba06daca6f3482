"""Seeded synthetic scenes for the five benchmark configurations.

Every scene is built from a SplitMix64 stream (seed 42 + config index by
default), so the same arguments always give the same bytes.  A scene is plain
data: colliders as sparse 8^3 bricks of voxel bytes (type | normal << 3, the
reference's CollisionVoxel packing, src/collision_voxel.zig:133-191) and bodies
as (type, position, quaternion, collider index, density), grouped per world.
The static ground is created first in each world so it gets body id 1 and is
body A of every ground pair (ids start at 1, src/system.zig:245).

Voxel classification follows src/collision_voxel.zig:194-235 (empty-face-count
rule and normal from the empty-neighbour mask); the 1-voxel-thick grounds of the
small scenes use the hand classification of examples/basic.zig:91-122.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream; floats are the top 24 bits (exact in f32)."""

    def __init__(self, seed):
        self.s = seed & _M64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 40) * (1.0 / (1 << 24)))

    def randint(self, lo, hi):  # inclusive
        return lo + self.next_u64() % (hi - lo + 1)

    def bulk_u64(self, n):
        """n draws, vectorised (same sequence as n next_u64 calls)."""
        idx = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            s = np.uint64(self.s) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = s
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.s = (self.s + n * 0x9E3779B97F4A7C15) & _M64
        return z

    def bulk_uniform(self, n):
        return (self.bulk_u64(n) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))

    def quat(self):
        """Uniform random unit quaternion (Shoemake 1992), (x, y, z, w)."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        t2, t3 = 2.0 * math.pi * u2, 2.0 * math.pi * u3
        return (a * math.sin(t2), a * math.cos(t2), b * math.sin(t3), b * math.cos(t3))


# ---- voxel grids -> bricks ---------------------------------------------------

def classify_grid(solid):
    """classifyFromNeighbors on a dense occupancy grid (src/collision_voxel.zig:213-235).
    Returns uint8 voxel bytes of the same shape (indexed [x, y, z])."""
    s = np.pad(solid.astype(bool), 1)
    c = s[1:-1, 1:-1, 1:-1]
    nx, px = s[:-2, 1:-1, 1:-1], s[2:, 1:-1, 1:-1]
    ny, py = s[1:-1, :-2, 1:-1], s[1:-1, 2:, 1:-1]
    nz, pz = s[1:-1, 1:-1, :-2], s[1:-1, 1:-1, 2:]
    mask = ((~nx) * 1 + (~px) * 2 + (~ny) * 4 + (~py) * 8 + (~nz) * 16 + (~pz) * 32).astype(np.int64)
    return (_CLASS_TABLE[mask] * c).astype(np.uint8)


def _normal_index(mask):
    # src/collision_voxel.zig:103-130
    lut = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0),
           (1, -1, 0), (-1, 1, 0), (-1, -1, 0), (1, 0, 1), (1, 0, -1), (-1, 0, 1), (-1, 0, -1),
           (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1), (1, 1, 1), (1, 1, -1), (1, -1, 1),
           (1, -1, -1), (-1, 1, 1), (-1, 1, -1), (-1, -1, 1), (-1, -1, -1)]
    if mask == 0:
        return 0
    d = [0, 0, 0]
    for bit, axis, sgn in ((1, 0, -1), (2, 0, 1), (4, 1, -1), (8, 1, 1), (16, 2, -1), (32, 2, 1)):
        if mask & bit:
            d[axis] += sgn
    t = tuple(d)
    if t in lut:
        return lut.index(t)
    # findClosestNormalIndex((0,0,0)): every dot is 0 -> first index
    best, best_dot = 0, -math.inf
    for i, n in enumerate(lut):
        ln = math.sqrt(sum(v * v for v in n))
        dd = sum(a * b / ln for a, b in zip(t, n))
        if dd > best_dot:
            best, best_dot = i, dd
    return best


def _class_byte(mask):
    cnt = bin(mask).count("1")
    t = _abi.INSIDE if cnt == 0 else _abi.FACE if cnt == 1 else _abi.EDGE if cnt == 2 else _abi.CORNER
    return t | (_normal_index(mask) << 3)


_CLASS_TABLE = np.array([_class_byte(m) for m in range(64)], dtype=np.uint8)


def grid_to_bricks(vox, origin):
    """Dense voxel bytes vox[x, y, z] whose [0,0,0] sits at voxel coordinate
    `origin` -> (brick_pos int32[n,3], voxels uint8[n,512]) with voxels indexed
    x + 8y + 64z (src/collider.zig:66-68).  Only bricks holding a solid voxel."""
    ox, oy, oz = origin
    sx, sy, sz = vox.shape
    b0 = [math.floor(o / 8) for o in origin]
    b1 = [math.floor((o + s - 1) / 8) for o, s in zip(origin, vox.shape)]
    nb = [hi - lo + 1 for lo, hi in zip(b0, b1)]
    big = np.zeros((nb[0] * 8, nb[1] * 8, nb[2] * 8), dtype=np.uint8)
    off = [o - lo * 8 for o, lo in zip(origin, b0)]
    big[off[0]:off[0] + sx, off[1]:off[1] + sy, off[2]:off[2] + sz] = vox
    # [bx,8,by,8,bz,8] -> [bz,by,bx, z,y,x]
    t = big.reshape(nb[0], 8, nb[1], 8, nb[2], 8).transpose(4, 2, 0, 5, 3, 1)
    t = t.reshape(nb[2], nb[1], nb[0], 512)
    keep = (t & 7).any(axis=-1)
    zz, yy, xx = np.nonzero(keep)
    pos = np.stack([xx + b0[0], yy + b0[1], zz + b0[2]], axis=1).astype(np.int32)
    return np.ascontiguousarray(pos), np.ascontiguousarray(t[zz, yy, xx])


@dataclass
class ColliderSpec:
    voxel_size: float
    brick_pos: np.ndarray   # int32 [n, 3]
    voxels: np.ndarray      # uint8 [n, 512]
    name: str = ""


@dataclass
class BodySpec:
    body_type: int
    position: tuple
    rotation: tuple
    collider: int
    density: float = 1.0


@dataclass
class Scene:
    name: str
    config: dict
    dt: float
    frames: int
    colliders: list = field(default_factory=list)
    worlds: list = field(default_factory=list)     # list of lists of BodySpec
    notes: str = ""

    @property
    def n_dynamic(self):
        return sum(1 for w in self.worlds for b in w if b.body_type == _abi.DYNAMIC)


# ---- shapes ------------------------------------------------------------------

def hand_ground(size_x, size_z):
    """1-thick ground, classified as examples/basic.zig:91-122 (generalised size):
    interior .face normal 2, rim edges and corners with the example's normals."""
    vox = np.zeros((size_x, 1, size_z), dtype=np.uint8)
    hx, hz = size_x // 2, size_z // 2
    for ix in range(size_x):
        x = ix - hx
        for iz in range(size_z):
            z = iz - hz
            on_x = x == -hx or x == hx - 1
            on_z = z == -hz or z == hz - 1
            t, n = _abi.FACE, 2
            if on_x and on_z:
                t = _abi.CORNER
                n = 23 if (x < 0 and z < 0) else 19 if (x > 0 and z < 0) else 22 if (x < 0 and z > 0) else 18
            elif on_x:
                t, n = _abi.EDGE, (8 if x < 0 else 6)
            elif on_z:
                t, n = _abi.EDGE, (15 if z < 0 else 14)
            vox[ix, 0, iz] = t | (n << 3)
    return vox, (-hx, 0, -hz)


def readme_cube():
    """3x3x3 cube classified as in README.md:85-116 (Quick Start)."""
    vox = np.zeros((3, 3, 3), dtype=np.uint8)
    for x in (-1, 0, 1):
        for y in (-1, 0, 1):
            for z in (-1, 0, 1):
                on = (x != 0, y != 0, z != 0)
                e = sum(on)
                t = {3: _abi.CORNER, 2: _abi.EDGE, 1: _abi.FACE, 0: _abi.INSIDE}[e]
                n = 2
                if on[0] and x > 0: n = 0
                if on[0] and x < 0: n = 1
                if on[1] and y > 0: n = 2
                if on[1] and y < 0: n = 3
                if on[2] and z > 0: n = 4
                if on[2] and z < 0: n = 5
                vox[x + 1, y + 1, z + 1] = t | (n << 3)
    return vox, (-1, -1, -1)


def box_shape(sx, sy, sz):
    solid = np.ones((sx, sy, sz), dtype=bool)
    return classify_grid(solid), (-(sx // 2), -(sy // 2), -(sz // 2))


def grid_to_solid_bricks(solid, origin):
    """Solid occupancy of a dense grid ([x, y, z], voxel (0,0,0) at `origin`) as
    bricks: positions (n, 3) and 8 uint64 words per brick (bit x + 8y of word
    z), for bricks holding at least one solid voxel."""
    solid = np.asarray(solid, dtype=bool)
    org = np.asarray(origin, dtype=np.int64)
    idx = np.argwhere(solid) + org
    if len(idx) == 0:
        return np.zeros((0, 3), np.int32), np.zeros((0, 8), np.uint64)
    bpos = np.floor_divide(idx, 8)
    keys, inv = np.unique(bpos, axis=0, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    local = idx - keys[inv] * 8
    words = np.zeros((len(keys), 8), dtype=np.uint64)
    bits = (np.uint64(1) << (local[:, 0] + 8 * local[:, 1]).astype(np.uint64))
    np.bitwise_or.at(words, (inv, local[:, 2]), bits)
    return keys.astype(np.int32), words


def _largest_component(solid):
    from scipy import ndimage
    lab, n = ndimage.label(solid)  # default structure: 6-connectivity
    if n <= 1:
        return solid
    counts = np.bincount(lab.ravel())
    counts[0] = 0
    return lab == int(np.argmax(counts))


def blob_shape(rng, n=16, p=0.4):
    """n^3 Bernoulli(p) grid, largest 6-connected component, reclassified."""
    solid = (rng.bulk_uniform(n * n * n) < p).reshape(n, n, n)
    solid = _largest_component(solid)
    return classify_grid(solid), (-(n // 2), -(n // 2), -(n // 2))


def rock_shape(rng, n=32):
    """3-octave lattice value noise (frequencies 1/16, 1/8, 1/4; amplitudes
    1, 1/2, 1/4, normalised) thresholded at 0.5, largest component, with a
    radial falloff so the rock is a closed lump."""
    coords = np.arange(n) + 0.5
    total = np.zeros((n, n, n))
    amp_sum = 0.0
    for freq, amp in ((1 / 16, 1.0), (1 / 8, 0.5), (1 / 4, 0.25)):
        cells = int(math.ceil(n * freq)) + 2
        lat = rng.bulk_uniform(cells ** 3).reshape(cells, cells, cells)
        u = coords * freq
        i0 = np.floor(u).astype(int)
        f = u - i0
        f = f * f * (3 - 2 * f)
        gx, gy, gz = np.meshgrid(i0, i0, i0, indexing="ij")
        fx, fy, fz = np.meshgrid(f, f, f, indexing="ij")
        acc = 0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                    acc = acc + w * lat[gx + dx, gy + dy, gz + dz]
        total += amp * acc
        amp_sum += amp
    total /= amp_sum
    c = (n - 1) / 2.0
    x, y, z = np.meshgrid(np.arange(n) - c, np.arange(n) - c, np.arange(n) - c, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z) / (n / 2.0)
    density = total * 0.6 + (1.0 - r) * 0.6
    solid = _largest_component(density > 0.5)
    return classify_grid(solid), (-(n // 2), -(n // 2), -(n // 2))


def _add_collider(scene, vox, origin, voxel_size, name):
    pos, v = grid_to_bricks(vox, origin)
    scene.colliders.append(ColliderSpec(voxel_size, pos, v, name))
    return len(scene.colliders) - 1


def _rot_aabb_half(half, q):
    x, y, z, w = q
    m = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return np.abs(m) @ np.asarray(half)


class _HeightMap:
    """Drop placement at voxel resolution over a square footprint.  A body keeps
    its sampled x, z and orientation; its rotated voxels are binned into
    vertical columns of `cell` metres, and it rests at max(sampled y, the
    lowest height where each of its columns clears the current column tops).
    Initial bodies never share a column interval, so they start disjoint and
    nest into one another as a pile would."""

    def __init__(self, half, cell=0.2, margin=12.0):
        self.cell = cell
        self.off = half + margin
        self.n = int(math.ceil(2 * self.off / cell)) + 1
        self.top = np.full((self.n, self.n), -1e9)

    def place(self, x, z, y, pts):
        """pts: rotated voxel centres relative to the body origin (N, 3)."""
        c = self.cell
        ix = np.clip(np.floor((x + pts[:, 0] + self.off) / c).astype(np.int64), 0, self.n - 1)
        iz = np.clip(np.floor((z + pts[:, 2] + self.off) / c).astype(np.int64), 0, self.n - 1)
        key = ix * self.n + iz
        order = np.argsort(key, kind="stable")
        ks, ys = key[order], pts[order, 1]
        uniq, start = np.unique(ks, return_index=True)
        lo = np.minimum.reduceat(ys, start) - c * 0.5
        hi = np.maximum.reduceat(ys, start) + c * 0.5
        ui, uk = uniq // self.n, uniq % self.n
        need = float(np.max(self.top[ui, uk] - lo)) + 0.02
        oy = max(y, need)
        self.top[ui, uk] = np.maximum(self.top[ui, uk], oy + hi)
        return oy


def _rot_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _drop(rng, hmap, local_pts, xz, y_range):
    q = rng.quat()
    x, y, z = rng.uniform(-xz, xz), rng.uniform(*y_range), rng.uniform(-xz, xz)
    pts = local_pts @ _rot_matrix(q).T
    oy = hmap.place(x, z, y, pts)
    return (float(x), float(oy), float(z)), q


def _voxel_centres(vox, origin, vs):
    """Centres of the surface voxels (every voxel whose type is not .inside): the
    lowest and highest voxel of any vertical column of a rotated body is a
    surface voxel, so these give the same column profile as all voxels."""
    t = vox & 7
    ii = np.argwhere((t != 0) & (t != _abi.INSIDE))
    return (ii + np.asarray(origin)[None, :] + 0.5) * vs


def _shape_half(vox, origin, vs):
    sx, sy, sz = vox.shape
    lo = np.array(origin, dtype=float) * vs
    hi = (np.array(origin, dtype=float) + np.array([sx, sy, sz])) * vs
    return (hi - lo) / 2.0, (hi + lo) / 2.0


def scene_config(name, dt, substeps, iterations, n_bodies, **over):
    cfg = dict(fixed_timestep=dt, substeps=substeps, velocity_iterations=iterations,
               max_bodies=n_bodies, default_friction=0.6, default_restitution=0.2)
    cfg.update(over)
    return cfg


def c0_small(seed=42, frames=300, n_side=8):
    """64 README cubes (3^3, 0.2 m voxels) on an 8x8 grid, 1 m spacing, heights
    U[1,6], random orientations, over a 64x1x64 hand-classified ground."""
    rng = SplitMix64(seed)
    vs = 0.2
    n = n_side * n_side
    sc = Scene("c0_small", scene_config("c0", 1 / 60, 4, 8, n + 1), 1 / 60, frames)
    g = _add_collider(sc, *hand_ground(64, 64), vs, "ground64")
    cube = _add_collider(sc, *readme_cube(), vs, "cube3")
    bodies = [BodySpec(_abi.STATIC, (0.0, -vs * 0.5, 0.0), (0.0, 0.0, 0.0, 1.0), g)]
    for i in range(n_side):
        for j in range(n_side):
            y = rng.uniform(1.0, 6.0)
            q = rng.quat()
            bodies.append(BodySpec(_abi.DYNAMIC, ((i - (n_side - 1) / 2) * 1.0, y,
                                                  (j - (n_side - 1) / 2) * 1.0), q, cube, 1.0))
    sc.worlds.append(bodies)
    return sc


def _mixed_world(rng, sc, n_bodies, ground, vs, xz, y_range, shape_pool=None, box_cache=None,
                 cell=0.2):
    placer = _HeightMap(xz, cell=cell)
    gspec = BodySpec(_abi.STATIC, (0.0, -2 * vs, 0.0), (0.0, 0.0, 0.0, 1.0), ground)
    bodies = [gspec]
    for k in range(n_bodies):
        if k % 2 == 0:
            sx, sy, sz = rng.randint(3, 16), rng.randint(3, 16), rng.randint(3, 16)
            key = (sx, sy, sz)
            if box_cache is not None and key in box_cache:
                cid, pts = box_cache[key]
            else:
                vox, org = box_shape(sx, sy, sz)
                cid = _add_collider(sc, vox, org, vs, f"box{sx}x{sy}x{sz}")
                pts = _voxel_centres(vox, org, vs)
                if box_cache is not None:
                    box_cache[key] = (cid, pts)
        else:
            if shape_pool is not None:
                cid, pts = shape_pool[rng.randint(0, len(shape_pool) - 1)]
            else:
                vox, org = blob_shape(rng)
                cid = _add_collider(sc, vox, org, vs, "blob16")
                pts = _voxel_centres(vox, org, vs)
        pos, q = _drop(rng, placer, pts, xz, y_range)
        bodies.append(BodySpec(_abi.DYNAMIC, pos, q, cid, 1.0))
    return bodies


def c1_large(seed=43, frames=600, n_bodies=4096, ground_vox=256):
    """4096 bodies (half boxes with sides in {3..16} voxels, half 16^3 40% blobs),
    0.2 m voxels, dropped from 1-40 m into a 32x32 m footprint over a
    256x2x256 ground.  Bodies that would overlap earlier ones are lifted onto
    them (voxel-column heightmap drop), so the scene starts as a deep column."""
    rng = SplitMix64(seed)
    vs = 0.2
    sc = Scene("c1_large", scene_config("c1", 1 / 60, 4, 8, n_bodies + 1), 1 / 60, frames)
    gvox = classify_grid(np.ones((ground_vox, 2, ground_vox), dtype=bool))
    g = _add_collider(sc, gvox, (-ground_vox // 2, 0, -ground_vox // 2), vs, "ground256x2")
    sc.worlds.append(_mixed_world(rng, sc, n_bodies, g, vs, 16.0, (1.0, 40.0)))
    return sc


def c2_rocks(seed=44, frames=400, nx=8, ny=4, nz=8):
    """256 value-noise rocks (32^3 at 0.1 m) in an 8x4x8 lattice with 10%
    jitter over a 128x1x128 ground of 0.2 m voxels (25.6 m, wide enough for the
    lattice footprint)."""
    rng = SplitMix64(seed)
    n = nx * ny * nz
    sc = Scene("c2_rocks", scene_config("c2", 1 / 120, 8, 12, n + 1), 1 / 120, frames)
    g = _add_collider(sc, *hand_ground(128, 128), 0.2, "ground128")
    bodies = [BodySpec(_abi.STATIC, (0.0, -0.1, 0.0), (0.0, 0.0, 0.0, 1.0), g)]
    spacing = 3.3
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                vox, org = rock_shape(rng)
                cid = _add_collider(sc, vox, org, 0.1, "rock32")
                j = [rng.uniform(-0.1, 0.1) * spacing * 0.5 for _ in range(3)]
                p = ((ix - (nx - 1) / 2) * spacing + j[0], 2.0 + iy * spacing + j[1],
                     (iz - (nz - 1) / 2) * spacing + j[2])
                bodies.append(BodySpec(_abi.DYNAMIC, p, rng.quat(), cid, 1.0))
    sc.worlds.append(bodies)
    return sc


def log_uniform_counts(rng, n_worlds, lo, hi):
    return [int(math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(n_worlds)]


def batched_worlds(seed=45, n_worlds=512, lo=8, hi=1024, frames=300, blob_pool=256, name="c3_batched"):
    """Independent worlds with log-uniform body counts, the c1 shape mix, and a
    64x1x64 ground each.  Colliders are shared through a pool of blob shapes and
    one collider per box size (colliders are shareable, src/system.zig:294)."""
    rng = SplitMix64(seed)
    vs = 0.2
    counts = log_uniform_counts(rng, n_worlds, lo, hi)
    sc = Scene(name, scene_config("c3", 1 / 60, 4, 8, max(counts) + 1), 1 / 60, frames)
    g = _add_collider(sc, *hand_ground(64, 64), vs, "ground64")
    pool = []
    for _ in range(blob_pool):
        vox, org = blob_shape(rng)
        cid = _add_collider(sc, vox, org, vs, "blob16")
        pool.append((cid, _voxel_centres(vox, org, vs)))
    box_cache = {}
    for n in counts:
        w = _mixed_world(rng, sc, n, g, vs, 6.0, (1.0, 20.0), shape_pool=pool, box_cache=box_cache,
                         cell=0.4)
        # the 64x1x64 ground is 1 voxel thick at y=0: centre it like c0
        w[0] = BodySpec(_abi.STATIC, (0.0, -vs * 0.5, 0.0), (0.0, 0.0, 0.0, 1.0), g)
        sc.worlds.append(w)
    return sc


_POOL_SHAPES = None  # (blob pool, box table) inherited by serving-pool workers


def _pool_world(args):
    seed, n, ground, vs = args
    blobs, boxes = _POOL_SHAPES
    w = _mixed_world(SplitMix64(seed), None, n, ground, vs, 6.0, (1.0, 20.0), shape_pool=blobs,
                     box_cache=boxes, cell=0.4)
    w[0] = BodySpec(_abi.STATIC, (0.0, -vs * 0.5, 0.0), (0.0, 0.0, 0.0, 1.0), ground)
    return w


def serving_pool(seed=46, n_worlds=2048, lo=8, hi=2048, blob_pool=256, name="c4_serving",
                 workers=16):
    """Resident worlds for the serving loop: body counts log-uniform in [lo, hi],
    the c1 shape mix over a 64x1x64 ground (as batched_worlds).  Every box size
    has its collider up front and each world draws from its own SplitMix64
    stream (seeded from the scene stream), so worlds generate in parallel."""
    global _POOL_SHAPES
    rng = SplitMix64(seed)
    vs = 0.2
    counts = log_uniform_counts(rng, n_worlds, lo, hi)
    sc = Scene(name, scene_config("c4", 1 / 60, 4, 8, max(counts) + 1), 1 / 60, 0)
    g = _add_collider(sc, *hand_ground(64, 64), vs, "ground64")
    blobs = []
    for _ in range(blob_pool):
        vox, org = blob_shape(rng)
        blobs.append((_add_collider(sc, vox, org, vs, "blob16"), _voxel_centres(vox, org, vs)))
    boxes = {}
    for sx in range(3, 17):
        for sy in range(3, 17):
            for sz in range(3, 17):
                vox, org = box_shape(sx, sy, sz)
                boxes[(sx, sy, sz)] = (_add_collider(sc, vox, org, vs, f"box{sx}x{sy}x{sz}"),
                                       _voxel_centres(vox, org, vs))
    seeds = [rng.next_u64() for _ in range(n_worlds)]
    jobs = [(seeds[i], counts[i], g, vs) for i in range(n_worlds)]
    _POOL_SHAPES = (blobs, boxes)
    try:
        if workers and workers > 1 and n_worlds > 32:
            import multiprocessing as mp
            with mp.get_context("fork").Pool(workers) as pool:
                sc.worlds = pool.map(_pool_world, jobs, chunksize=8)
        else:
            sc.worlds = [_pool_world(j) for j in jobs]
    finally:
        _POOL_SHAPES = None
    return sc


def build_system(lib, scene, cfg_over=None, worlds=None):
    """Create a system of `lib` (product or oracle) holding `scene`.
    Returns (system pointer, per-world handle lists)."""
    import ctypes as C
    cfg = dict(scene.config)
    if cfg_over:
        cfg.update(cfg_over)
    wl = scene.worlds if worlds is None else [scene.worlds[i] for i in worlds]
    cfg["max_bodies"] = max(cfg.get("max_bodies", 0), max(len(w) for w in wl))
    c = _abi.default_config(**cfg)
    sysp = C.c_void_p()
    lib.check(lib.system_create(C.byref(c), len(wl), C.byref(sysp)), "system_create")
    cids = []
    for col in scene.colliders:
        out = C.c_uint32()
        bp = np.ascontiguousarray(col.brick_pos, dtype=np.int32)
        vx = np.ascontiguousarray(col.voxels, dtype=np.uint8)
        lib.check(lib.collider_create(sysp, col.voxel_size, len(bp),
                                      bp.ctypes.data_as(C.POINTER(C.c_int32)),
                                      vx.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(out)),
                  "collider_create")
        cids.append(out.value)
    handles = []
    for wi, bodies in enumerate(wl):
        hs = []
        for b in bodies:
            h = _abi.Handle()
            pos = (C.c_float * 3)(*b.position)
            rot = (C.c_float * 4)(*b.rotation)
            lib.check(lib.body_create_w(sysp, wi, b.body_type, pos, rot, cids[b.collider],
                                        b.density, C.byref(h)), "body_create_w")
            hs.append(h)
        handles.append(hs)
    return sysp, handles


# Cubes that rest on the ground from the first frame, so that every
# cube-ground pair reaches the narrow phase in every frame (the thread-pool
# narrow phase of src/system.zig:618-638 takes its branch on the pair count).
def _stored_bricks(brick_pos):
    """Brick positions a collider keeps: its brick map is keyed by
    hashBrickPos (src/collider.zig:187-192), so of the positions that share a
    hash only the first one stores a brick."""
    seen, kept = set(), set()
    m = (1 << 64) - 1
    for x, y, z in brick_pos.tolist():
        h = (x & m) ^ (((y & m) * 0x517cc1b727220a95) & m) ^ (((z & m) * 0x9e3779b97f4a7c15) & m)
        if h not in seen:
            seen.add(h)
            kept.add((x, y, z))
    return kept


def resting_cubes(n_cubes, spacing=1.0, y=0.33, frames=60, name="resting"):
    """`n_cubes` README cubes on a square grid (row-major from the -x/-z
    corner), each over ground that the collider stores, with their lower faces
    0.03 m above it; the ground is body 1 (slot 0).  A cube's fat box overlaps
    the ground's from the start."""
    vs = 0.2
    cells = 64
    while True:
        vox, org = hand_ground(cells, cells)
        bp, _ = grid_to_bricks(vox, org)
        kept = _stored_bricks(bp)
        spots = []
        half = cells * vs / 2 - 1.0
        k = int(half / spacing)
        for i in range(-k, k + 1):
            for j in range(-k, k + 1):
                x, z = i * spacing, j * spacing
                # the cube's footprint (0.6 m) in voxel columns, all in kept bricks
                ok = all((int(math.floor((x + dx) / vs)) // 8, 0, int(math.floor((z + dz) / vs)) // 8)
                         in kept for dx in (-0.3, 0.29) for dz in (-0.3, 0.29))
                if ok:
                    spots.append((x, z))
        if len(spots) >= n_cubes:
            break
        cells *= 2
    sc = Scene(name, scene_config("c0", 1 / 60, 4, 8, n_cubes + 1), 1 / 60, frames)
    g = _add_collider(sc, vox, org, vs, f"ground{cells}")
    cube = _add_collider(sc, *readme_cube(), vs, "cube3")
    bodies = [BodySpec(_abi.STATIC, (0.0, -vs * 0.5, 0.0), (0.0, 0.0, 0.0, 1.0), g)]
    for x, z in spots[:n_cubes]:
        bodies.append(BodySpec(_abi.DYNAMIC, (x, y, z), (0.0, 0.0, 0.0, 1.0), cube, 1.0))
    sc.worlds.append(bodies)
    return sc
