// vxp_oracle.cpp — ORACLE: test infrastructure, never shipped or linked by the
// product.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load it (as liboracle), as the checker and the timed CPU baseline.
//
// A C++ restatement, written from the source text, of zig-voxel-physics'
// per-frame step: src/collider.zig, src/body.zig, src/broad_phase.zig,
// src/narrow_phase.zig, src/islands.zig, src/solver.zig, src/system.zig.
// The reference is pure Zig and no Zig toolchain exists in this environment,
// so the reference itself cannot be built; this file is the CPU path.
//
// Deviations from the reference, each forced by Zig std.AutoHashMap
// iteration order that cannot be reproduced outside Zig's stdlib:
//  * broad-phase pairs are processed in ascending pair key (a<<32|b)
//    instead of AABB-tree order (src/broad_phase.zig:488-526);
//  * manifolds are solved per colour of the solver colouring (ascending
//    colour, then pair key) or in pair-key order (solve_order knob) instead of
//    hash order (src/system.zig:687-702);
//  * mass properties sum voxels over bricks in ascending (z,y,x) brick order
//    instead of hash order (src/body.zig:481-506, src/collider.zig:382-400);
//  * the 512-pair filter cap (src/system.zig:594) is a knob (0 = unlimited).
//
// Parity status of each part and what pins it: see tests/test_oracle_*.py.
#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../include/vxp.h"
#include "ora_math.hpp"

namespace ora {

// ---------------------------------------------------------------------------
// src/collider.zig
// ---------------------------------------------------------------------------
struct Brick {
  IVec3 pos;
  uint8_t v[512];
  AABB local_aabb;
  bool is_empty;
  uint16_t solid_count;
};

inline int32_t floorDiv8(int32_t a) { return (a >= 0) ? a / 8 : -((-a + 7) / 8); }
inline int32_t mod8(int32_t a) { return a - floorDiv8(a) * 8; }

// src/collider.zig:187-192
inline uint64_t hashBrickPos(IVec3 p) {
  uint64_t ux = (uint64_t)(int64_t)p.x, uy = (uint64_t)(int64_t)p.y, uz = (uint64_t)(int64_t)p.z;
  return ux ^ (uy * 0x517cc1b727220a95ull) ^ (uz * 0x9e3779b97f4a7c15ull);
}

struct Collider {
  float voxel_size = 0.1f;
  std::vector<Brick> bricks;  // ascending (z, y, x) brick position
  std::unordered_map<uint64_t, int> map;
  IVec3 min_bound{0, 0, 0}, max_bound{0, 0, 0};
  AABB local_aabb = AABB::empty();
  uint32_t solid_count = 0;

  // src/collider.zig:257-264
  uint8_t getVoxel(IVec3 v) const {
    IVec3 bp{floorDiv8(v.x), floorDiv8(v.y), floorDiv8(v.z)};
    auto it = map.find(hashBrickPos(bp));
    if (it == map.end()) return 0;
    const Brick& b = bricks[it->second];
    return b.v[mod8(v.x) + 8 * mod8(v.y) + 64 * mod8(v.z)];
  }
  // src/collider.zig:276-280 (half-open)
  bool isValidCoord(IVec3 v) const {
    return v.x >= min_bound.x && v.x < max_bound.x && v.y >= min_bound.y && v.y < max_bound.y &&
           v.z >= min_bound.z && v.z < max_bound.z;
  }
  // src/collider.zig:213-220
  Vec3 voxelToLocal(IVec3 v) const {
    float half = voxel_size * 0.5f;
    return {(float)v.x * voxel_size + half, (float)v.y * voxel_size + half,
            (float)v.z * voxel_size + half};
  }
  // src/collider.zig:223-229
  IVec3 localToVoxel(Vec3 p) const {
    return {(int32_t)std::floor(p.x / voxel_size), (int32_t)std::floor(p.y / voxel_size),
            (int32_t)std::floor(p.z / voxel_size)};
  }
  // src/collider.zig:90-133 per brick, 283-333 per collider
  void recomputeMetadata() {
    min_bound = {INT32_MAX, INT32_MAX, INT32_MAX};
    max_bound = {INT32_MIN, INT32_MIN, INT32_MIN};
    solid_count = 0;
    local_aabb = AABB::empty();
    for (Brick& b : bricks) {
      b.solid_count = 0;
      Vec3 mn = Vec3::splat(INFINITY), mx = Vec3::splat(-INFINITY);
      for (int i = 0; i < 512; i++) {
        if (voxelType(b.v[i]) == kEmpty) continue;
        b.solid_count++;
        int cx = i % 8, cy = (i / 8) % 8, cz = i / 64;
        Vec3 lp = Vec3{(float)b.pos.x * 8.0f + (float)cx, (float)b.pos.y * 8.0f + (float)cy,
                       (float)b.pos.z * 8.0f + (float)cz}
                      .scale(voxel_size);
        mn = mn.minComponents(lp);
        mx = mx.maxComponents(lp.add(Vec3::splat(voxel_size)));
      }
      b.is_empty = b.solid_count == 0;
      b.local_aabb = b.is_empty ? AABB::empty() : AABB{mn, mx};
      if (b.is_empty) continue;
      solid_count += b.solid_count;
      for (int i = 0; i < 512; i++) {
        if (voxelType(b.v[i]) == kEmpty) continue;
        int32_t vx = b.pos.x * 8 + i % 8, vy = b.pos.y * 8 + (i / 8) % 8, vz = b.pos.z * 8 + i / 64;
        min_bound = {std::min(min_bound.x, vx), std::min(min_bound.y, vy), std::min(min_bound.z, vz)};
        max_bound = {std::max(max_bound.x, vx + 1), std::max(max_bound.y, vy + 1),
                     std::max(max_bound.z, vz + 1)};
      }
      local_aabb = local_aabb.merge(b.local_aabb);
    }
    if (solid_count == 0) {
      min_bound = {0, 0, 0};
      max_bound = {0, 0, 0};
    }
  }
  // src/collider.zig:376-403 (extra half voxel on top of voxelToLocal's)
  Vec3 getCenterOfMass() const {
    if (solid_count == 0) return {0, 0, 0};
    Vec3 sum{0, 0, 0};
    uint32_t count = 0;
    for (const Brick& b : bricks) {
      if (b.is_empty) continue;
      for (int i = 0; i < 512; i++) {
        if (voxelType(b.v[i]) == kEmpty) continue;
        IVec3 vp{b.pos.x * 8 + i % 8, b.pos.y * 8 + (i / 8) % 8, b.pos.z * 8 + i / 64};
        Vec3 p = voxelToLocal(vp).add(Vec3::splat(voxel_size * 0.5f));
        sum = sum.add(p);
        count++;
      }
    }
    return sum.scale(1.0f / (float)count);
  }
};

struct MassProps {
  float inv_mass;
  Mat3 inv_inertia;
};

// src/body.zig:462-519
MassProps computeMassProperties(const Collider& c, float density) {
  if (c.solid_count == 0) return {0.0f, Mat3::zero()};
  float vv = c.voxel_size * c.voxel_size * c.voxel_size;
  float vm = vv * density;
  float total = vm * (float)c.solid_count;
  Vec3 com = c.getCenterOfMass();
  Mat3 I = Mat3::zero();
  for (const Brick& b : c.bricks) {
    if (b.is_empty) continue;
    for (int i = 0; i < 512; i++) {
      if (voxelType(b.v[i]) == kEmpty) continue;
      IVec3 vp{b.pos.x * 8 + i % 8, b.pos.y * 8 + (i / 8) % 8, b.pos.z * 8 + i / 64};
      Vec3 p = c.voxelToLocal(vp).add(Vec3::splat(c.voxel_size * 0.5f));
      Vec3 r = p.sub(com);
      float r2 = r.dot(r);
      I.m[0][0] += vm * (r2 - r.x * r.x);
      I.m[1][1] += vm * (r2 - r.y * r.y);
      I.m[2][2] += vm * (r2 - r.z * r.z);
      I.m[0][1] -= vm * r.x * r.y;
      I.m[0][2] -= vm * r.x * r.z;
      I.m[1][2] -= vm * r.y * r.z;
    }
  }
  I.m[1][0] = I.m[0][1];
  I.m[2][0] = I.m[0][2];
  I.m[2][1] = I.m[1][2];
  return {1.0f / total, I.inverse()};
}

// ---------------------------------------------------------------------------
// src/body.zig
// ---------------------------------------------------------------------------
struct Body {
  uint32_t id = 0;
  uint8_t type = VXP_BODY_DYNAMIC;
  Vec3 pos{0, 0, 0};
  Quat rot = Quat::identity();
  Vec3 lv{0, 0, 0}, av{0, 0, 0};
  float inv_mass = 1.0f;
  Mat3 inv_inertia_local = Mat3::identity();
  float friction = 0.5f, restitution = 0.0f, lin_damp = 0.3f, ang_damp = 0.5f;
  int collider = -1;
  bool sleeping = false;
  float sleep_timer = 0.0f;
  AABB fat;  // node AABB in the broad phase tree
  Transform xf() const { return {pos, rot}; }
  // src/body.zig:386-392
  Mat3 worldInvInertia() const {
    if (inv_mass == 0.0f) return Mat3::zero();
    Mat3 R = rot.toMat3();
    return R.mul(inv_inertia_local).mul(R.transpose());
  }
};

struct Slot {
  Body body;
  uint32_t generation = 0;
  bool active = false;
};

// ---------------------------------------------------------------------------
// src/narrow_phase.zig — contacts and manifolds
// ---------------------------------------------------------------------------
struct Contact {  // src/narrow_phase.zig:336-393
  Vec3 point, normal;
  float pen;
  Vec3 anchor_a, anchor_b;
  float normal_impulse = 0, tangent_impulse[2] = {0, 0};
  float normal_mass = 0, tangent_mass[2] = {0, 0};
  uint64_t feature_id = 0;
};

// src/narrow_phase.zig:384-392
inline uint64_t computeFeatureId(IVec3 a, IVec3 b) {
  uint64_t ax = (uint64_t)((int64_t)a.x & 0xFFFF);
  uint64_t ay = (uint64_t)(((int64_t)a.y & 0xFFFF) << 16);
  uint64_t az = (uint64_t)(((int64_t)a.z & 0xFFFF) << 32);
  uint64_t bx = (uint64_t)(((int64_t)b.x & 0xF) << 48);
  uint64_t by = (uint64_t)(((int64_t)b.y & 0xF) << 52);
  uint64_t bz = (uint64_t)(((int64_t)b.z & 0xF) << 56);
  return ax | ay | az | bx | by | bz;
}

struct Manifold {  // src/narrow_phase.zig:396-502
  uint32_t a = 0, b = 0;
  Contact c[4];
  int n = 0;
  float friction = 0.5f, restitution = 0.0f;
  uint32_t color = 0xFFFFFFFFu;

  // src/narrow_phase.zig:451-481: streaming top-4
  void addContact(const Contact& ct) {
    for (int i = 0; i < n; i++) {
      if (c[i].feature_id == ct.feature_id) {
        float ni = c[i].normal_impulse, t0 = c[i].tangent_impulse[0], t1 = c[i].tangent_impulse[1];
        c[i] = ct;
        c[i].normal_impulse = ni;
        c[i].tangent_impulse[0] = t0;
        c[i].tangent_impulse[1] = t1;
        return;
      }
    }
    if (n < 4) {
      c[n++] = ct;
      return;
    }
    float minp = ct.pen;
    int mi = -1;
    for (int i = 0; i < n; i++) {
      if (c[i].pen < minp) {
        minp = c[i].pen;
        mi = i;
      }
    }
    if (mi >= 0) c[mi] = ct;
  }
};

// ContactManager.addManifold warm start (src/narrow_phase.zig:1049-1059): each
// new contact takes the impulses of the previous frame's contact with the same
// feature id (the first one).
inline void warmStartFrom(Manifold& m, const Manifold& old) {
  for (int i = 0; i < m.n; i++)
    for (int j = 0; j < old.n; j++)
      if (m.c[i].feature_id == old.c[j].feature_id) {
        m.c[i].normal_impulse = old.c[j].normal_impulse;
        m.c[i].tangent_impulse[0] = old.c[j].tangent_impulse[0];
        m.c[i].tangent_impulse[1] = old.c[j].tangent_impulse[1];
        break;
      }
}

inline uint64_t pairKey(uint32_t a, uint32_t b) {
  uint32_t lo = std::min(a, b), hi = std::max(a, b);
  return ((uint64_t)lo << 32) | hi;
}

// ---- SAT: src/narrow_phase.zig:504-569 -------------------------------------
static void projectOBB(const OBB& o, Vec3 axis, float& mn, float& mx) {
  Vec3 v[8];
  o.getVertices(v);
  mn = INFINITY;
  mx = -INFINITY;
  for (int i = 0; i < 8; i++) {
    float p = v[i].dot(axis);
    mn = zmin(mn, p);
    mx = zmax(mx, p);
  }
}
static bool overlap1(float amin, float amax, float bmin, float bmax) {
  return !(amax < bmin || bmax < amin);
}
bool satEarlyOut(const OBB& a, const OBB& b) {
  Vec3 aa[3], bb[3];
  a.getAxes(aa);
  b.getAxes(bb);
  float p0, p1, q0, q1;
  for (int i = 0; i < 3; i++) {
    projectOBB(a, aa[i], p0, p1);
    projectOBB(b, aa[i], q0, q1);
    if (!overlap1(p0, p1, q0, q1)) return true;
  }
  for (int i = 0; i < 3; i++) {
    projectOBB(a, bb[i], p0, p1);
    projectOBB(b, bb[i], q0, q1);
    if (!overlap1(p0, p1, q0, q1)) return true;
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Vec3 cr = aa[i].cross(bb[j]);
      float len = cr.length();
      if (len < 0.001f) continue;
      Vec3 ax = cr.scale(1.0f / len);
      projectOBB(a, ax, p0, p1);
      projectOBB(b, ax, q0, q1);
      if (!overlap1(p0, p1, q0, q1)) return true;
    }
  return false;
}

// ---- src/narrow_phase.zig:644-698 -------------------------------------------
struct Region {
  IVec3 mn, mx;
  bool valid;
};
Region getOverlapRegion(const Transform& ta, const Collider& ca, const Transform& tb,
                        const Collider& cb) {
  OBB ob = OBB::fromAABB(cb.local_aabb, tb);
  Vec3 vb[8];
  ob.getVertices(vb);
  Quat inv = ta.rotation.conjugate();
  Vec3 tmin = Vec3::splat(INFINITY), tmax = Vec3::splat(-INFINITY);
  for (int i = 0; i < 8; i++) {
    Vec3 rel = inv.rotate(vb[i].sub(ta.position));
    tmin = tmin.minComponents(rel);
    tmax = tmax.maxComponents(rel);
  }
  float vs = ca.voxel_size;
  tmin = tmin.sub(Vec3::splat(vs));
  tmax = tmax.add(Vec3::splat(vs));
  Vec3 mnc = tmin.clamp(ca.local_aabb.min, ca.local_aabb.max);
  Vec3 mxc = tmax.clamp(ca.local_aabb.min, ca.local_aabb.max);
  if (mnc.x >= mxc.x || mnc.y >= mxc.y || mnc.z >= mxc.z) return {{0, 0, 0}, {0, 0, 0}, false};
  IVec3 a = ca.localToVoxel(mnc), b = ca.localToVoxel(mxc);
  return {{std::max(a.x, ca.min_bound.x), std::max(a.y, ca.min_bound.y), std::max(a.z, ca.min_bound.z)},
          {std::min(b.x, ca.max_bound.x), std::min(b.y, ca.max_bound.y), std::min(b.z, ca.max_bound.z)},
          true};
}

// ---- src/narrow_phase.zig:860-922 -------------------------------------------
static Contact generateContact(uint8_t va, IVec3 ca, Vec3 wa, const Transform& ta, uint8_t vb,
                               IVec3 cb, Vec3 wb, const Transform& tb, Vec3 diff, float dsq,
                               float vs) {
  float dist = std::sqrt(dsq);
  Vec3 cp = wa.add(wb).scale(0.5f);
  float pen = vs - dist;
  Vec3 n;
  Vec3 na = ta.rotation.rotate(getNormalVec3(voxelNormal(va)));
  Vec3 nb = tb.rotation.rotate(getNormalVec3(voxelNormal(vb)));
  uint8_t ta_ = voxelType(va), tb_ = voxelType(vb);
  if (ta_ == kFace) {
    n = na;
  } else if (tb_ == kFace) {
    n = nb.negate();
  } else if (dist > 0.001f) {
    Vec3 dn = diff.scale(1.0f / dist);
    if (dn.dot(na) > 0.3f)
      n = na;
    else if (dn.dot(nb.negate()) > 0.3f)
      n = nb.negate();
    else
      n = dn;
  } else {
    n = na;
  }
  if (ta_ == kInside || tb_ == kInside) pen = zmax(pen, vs);
  Contact c;
  c.point = cp;
  c.normal = n;
  c.pen = pen;
  c.anchor_a = ta.inverseTransformPoint(cp);
  c.anchor_b = tb.inverseTransformPoint(cp);
  c.feature_id = computeFeatureId(ca, cb);
  return c;
}

struct DetectResult {
  Manifold m;
  uint32_t voxels_checked = 0;
  uint32_t candidates = 0;
  bool has_collision = false;
};

// src/narrow_phase.zig:722-857: detectVoxelCollisions + checkNeighborhood
DetectResult detectVoxelCollisions(const Body& A, const Collider& ca, const Body& B,
                                   const Collider& cb) {
  DetectResult r;
  Transform ta = A.xf(), tb = B.xf();
  r.m.a = A.id;
  r.m.b = B.id;
  r.m.friction = std::sqrt(A.friction * B.friction);
  r.m.restitution = zmax(A.restitution, B.restitution);
  Region reg = getOverlapRegion(ta, ca, tb, cb);
  if (!reg.valid) return r;
  float vs = ca.voxel_size;
  float cds = vs * vs * 1.0f * 1.0f;
  float spec = cds + 0.02f * 0.02f + 2.0f * std::sqrt(cds) * 0.02f;
  for (int32_t z = reg.mn.z; z <= reg.mx.z; z++)
    for (int32_t y = reg.mn.y; y <= reg.mx.y; y++)
      for (int32_t x = reg.mn.x; x <= reg.mx.x; x++) {
        IVec3 ia{x, y, z};
        uint8_t va = ca.getVoxel(ia);
        if (voxelType(va) == kEmpty) continue;
        Vec3 wa = ta.transformPoint(ca.voxelToLocal(ia));
        Vec3 lb = tb.inverseTransformPoint(wa);
        IVec3 c = cb.localToVoxel(lb);
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              IVec3 ib{c.x + dx, c.y + dy, c.z + dz};
              if (!cb.isValidCoord(ib)) continue;
              uint8_t vb = cb.getVoxel(ib);
              if (voxelType(vb) == kEmpty) continue;
              r.voxels_checked++;
              if (!shouldCheckCollision(voxelType(va), voxelType(vb))) continue;
              Vec3 wb = tb.transformPoint(cb.voxelToLocal(ib));
              Vec3 diff = wb.sub(wa);
              float dsq = diff.lengthSquared();
              if (dsq < spec) {
                r.candidates++;
                r.m.addContact(generateContact(va, ia, wa, ta, vb, ib, wb, tb, diff, dsq, vs));
              }
            }
      }
  r.has_collision = r.m.n > 0;
  return r;
}

// ---- BrickPairCache: src/narrow_phase.zig:65-333 -----------------------------
struct CacheKey {
  uint32_t a, b;
  uint64_t th;
  bool eql(const CacheKey& o) const { return a == o.a && b == o.b && th == o.th; }
};
struct CacheValue {
  Contact c[4];
  uint8_t count = 0;
  bool has_collision = false;
  uint32_t voxels_checked = 0;
};

// src/narrow_phase.zig:148-169
uint64_t computeTransformHash(Vec3 rp, Quat rq) {
  int32_t qx = (int32_t)(rp.x * 100.0f), qy = (int32_t)(rp.y * 100.0f), qz = (int32_t)(rp.z * 100.0f);
  int16_t rw = (int16_t)zclamp(rq.w * 1000.0f, -32767.0f, 32767.0f);
  int16_t rx = (int16_t)zclamp(rq.x * 1000.0f, -32767.0f, 32767.0f);
  int16_t ry = (int16_t)zclamp(rq.y * 1000.0f, -32767.0f, 32767.0f);
  int16_t rz = (int16_t)zclamp(rq.z * 1000.0f, -32767.0f, 32767.0f);
  uint64_t h = 0;
  h = h * 31 + (uint64_t)(uint32_t)qx;
  h = h * 31 + (uint64_t)(uint32_t)qy;
  h = h * 31 + (uint64_t)(uint32_t)qz;
  h = h * 31 + (uint64_t)(uint16_t)rw;
  h = h * 31 + (uint64_t)(uint16_t)rx;
  h = h * 31 + (uint64_t)(uint16_t)ry;
  h = h * 31 + (uint64_t)(uint16_t)rz;
  return h;
}
// src/narrow_phase.zig:122-134 (brick positions are always zero on the step path)
uint64_t hashKey(const CacheKey& k) {
  uint64_t h = 0;
  h = h * 31 + k.a;
  h = h * 31 + k.b;
  for (int i = 0; i < 6; i++) h = h * 31 + 0;
  h = h * 31 + k.th;
  return h;
}

struct BrickPairCache {
  struct Entry {
    CacheKey key;
    CacheValue value;
    uint64_t last_used = 0;
    bool valid = false;
  };
  std::vector<Entry> e;
  uint64_t frame = 0;
  void init(size_t cap) { e.assign(cap, Entry{}); }
  // src/narrow_phase.zig:228-252; returns entry index or -1
  long get(const CacheKey& k) {
    size_t cap = e.size();
    size_t idx = hashKey(k) % cap;
    size_t maxp = std::min<size_t>(cap, 16);
    for (size_t p = 0; p < maxp; p++) {
      Entry& en = e[idx];
      if (!en.valid) return -1;
      if (en.key.eql(k)) {
        en.last_used = frame;
        return (long)idx;
      }
      idx = (idx + 1) % cap;
    }
    return -1;
  }
  // src/narrow_phase.zig:255-296; returns the entry index written
  size_t put(const CacheKey& k, const CacheValue& v) {
    size_t cap = e.size();
    size_t start = hashKey(k) % cap, idx = start;
    size_t maxp = std::min<size_t>(cap, 16);
    size_t oldest = start;
    uint64_t oldest_frame = UINT64_MAX;
    for (size_t p = 0; p < maxp; p++) {
      Entry& en = e[idx];
      if (!en.valid) {
        en.key = k;
        en.value = v;
        en.last_used = frame;
        en.valid = true;
        return idx;
      }
      if (en.key.eql(k)) {
        en.value = v;
        en.last_used = frame;
        return idx;
      }
      if (en.last_used < oldest_frame) {
        oldest_frame = en.last_used;
        oldest = idx;
      }
      idx = (idx + 1) % cap;
    }
    e[oldest].key = k;
    e[oldest].value = v;
    e[oldest].last_used = frame;
    e[oldest].valid = true;
    return oldest;
  }
};

// ---------------------------------------------------------------------------
// Solver colouring shared with the GPU pipeline: greedy colouring of the
// manifolds that the solver visits, in ascending (prio(key), key) order, where
// two manifolds conflict iff they share an awake dynamic body (the only bodies
// the solver writes: src/solver.zig:316,322,328,336,893-918).  Manifolds of one
// colour write disjoint bodies, so per-colour Gauss-Seidel in any in-colour
// order equals sequential Gauss-Seidel in (colour, key) order.
// ---------------------------------------------------------------------------
inline uint32_t colorPrio(uint64_t key) {
  uint64_t h = key;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return (uint32_t)(h >> 32);
}

// ---------------------------------------------------------------------------
// src/solver.zig
// ---------------------------------------------------------------------------
struct SoftParams {
  float bias_rate, mass_coeff, impulse_coeff;
};
// src/solver.zig:57-72
SoftParams computeSoftParams(float hertz, float zeta, float dt) {
  if (hertz <= 0.0f) return {0.0f, 1.0f, 1.0f};
  float omega = 2.0f * kPi * hertz;
  float a1 = 2.0f * zeta + omega * dt;
  float a2 = dt * omega * a1;
  float a3 = 1.0f / (1.0f + a2);
  return {omega / a1, a2 * a3, a3};
}

// src/solver.zig:967-981
Vec3 computeOrthogonal(Vec3 v) {
  if (std::fabs(v.x) < std::fabs(v.y)) {
    if (std::fabs(v.x) < std::fabs(v.z)) return Vec3{0, -v.z, v.y}.normalizeOrZero();
    return Vec3{-v.y, v.x, 0}.normalizeOrZero();
  }
  if (std::fabs(v.y) < std::fabs(v.z)) return Vec3{-v.z, 0, v.x}.normalizeOrZero();
  return Vec3{-v.y, v.x, 0}.normalizeOrZero();
}

// src/solver.zig:984-1005
Quat integrateRotationSmall(Quat r, Vec3 d) {
  Quat hw{d.x * 0.5f, d.y * 0.5f, d.z * 0.5f, 0.0f};
  Quat dq{hw.w * r.x + hw.x * r.w + hw.y * r.z - hw.z * r.y,
          hw.w * r.y - hw.x * r.z + hw.y * r.w + hw.z * r.x,
          hw.w * r.z + hw.x * r.y - hw.y * r.x + hw.z * r.w,
          hw.w * r.w - hw.x * r.x - hw.y * r.y - hw.z * r.z};
  return Quat{r.x + dq.x, r.y + dq.y, r.z + dq.z, r.w + dq.w}.normalize();
}

struct BodyState {  // src/solver.zig:198-339
  Vec3 pos;
  Quat rot;
  Vec3 lv, av;
  float inv_mass;
  Mat3 inv_i;
  uint8_t type;
  bool sleeping;
  Body* body;

  void applyImpulseAtPoint(Vec3 imp, Vec3 r) {
    if (type != VXP_BODY_DYNAMIC || sleeping) return;
    lv = lv.add(imp.scale(inv_mass));
    av = av.add(inv_i.mulVec(r.cross(imp)));
  }
  Vec3 pointVel(Vec3 p) const { return lv.add(av.cross(p.sub(pos))); }
  // src/solver.zig:246-296
  void writeBack() {
    if (type == VXP_BODY_STATIC) return;
    bool fin_q = std::isfinite(rot.x) && std::isfinite(rot.y) && std::isfinite(rot.z) &&
                 std::isfinite(rot.w);
    if (!pos.isFinite() || !fin_q) {
      pos = body->pos.isFinite() ? body->pos : Vec3{0, 2, 0};
      rot = Quat::identity();
      lv = {0, 0, 0};
      av = {0, 0, 0};
    }
    if (!lv.isFinite()) lv = {0, 0, 0};
    if (!av.isFinite()) av = {0, 0, 0};
    float ls = lv.lengthSquared(), as = av.lengthSquared();
    if (ls > 50.0f * 50.0f) lv = lv.scale(50.0f / std::sqrt(ls));
    if (as > 30.0f * 30.0f) av = av.scale(30.0f / std::sqrt(as));
    if (ls < 0.002f) lv = {0, 0, 0};
    if (as < 0.002f) av = {0, 0, 0};
    rot = rot.normalize();
    body->pos = pos;
    body->rot = rot;
    body->lv = lv;
    body->av = av;
  }
};

struct CPoint {  // src/solver.zig:122-174
  Vec3 point, normal, la, lb, ra, rb;
  float base_sep, normal_mass, tangent_mass[2];
  Vec3 t1, t2;
  float bias, ni, ti[2], vn0, pen;
};
struct CManifold {  // src/solver.zig:177-196
  int ia, ib;
  float friction, restitution;
  int n;
  CPoint p[4];
  uint64_t key;
};

struct Solver {
  const vxp_config* cfg;
  Vec3 gravity;
  std::vector<BodyState> st;
  std::unordered_map<uint32_t, int> idx;
  std::vector<CManifold> cms;
  float max_pen = 0;
  uint32_t pos_iters = 0;

  void clear() {
    st.clear();
    idx.clear();
    cms.clear();
    max_pen = 0;
    pos_iters = 0;
  }
  // src/solver.zig:429-435, 231-243
  void addBody(Body* b) {
    if (idx.count(b->id)) return;
    idx[b->id] = (int)st.size();
    st.push_back({b->pos, b->rot, b->lv, b->av, b->inv_mass, b->worldInvInertia(), b->type,
                  b->sleeping, b});
  }
  // src/solver.zig:449-491
  void addManifold(const Manifold& m) {
    auto fa = idx.find(m.a), fb = idx.find(m.b);
    if (fa == idx.end() || fb == idx.end()) return;
    CManifold cm;
    cm.ia = fa->second;
    cm.ib = fb->second;
    cm.friction = m.friction;
    cm.restitution = m.restitution;
    cm.n = m.n;
    cm.key = pairKey(m.a, m.b);
    const BodyState& A = st[cm.ia];
    const BodyState& B = st[cm.ib];
    for (int i = 0; i < m.n; i++) {
      const Contact& c = m.c[i];
      Vec3 wa = A.rot.rotate(c.anchor_a).add(A.pos);
      Vec3 wb = B.rot.rotate(c.anchor_b).add(B.pos);
      Vec3 ad = wb.sub(wa);
      CPoint& p = cm.p[i];
      p.point = c.point;
      p.normal = c.normal;
      p.la = c.anchor_a;
      p.lb = c.anchor_b;
      p.ra = wa.sub(A.pos);
      p.rb = wb.sub(B.pos);
      p.base_sep = -c.pen - ad.dot(c.normal);
      p.normal_mass = 0;
      p.tangent_mass[0] = p.tangent_mass[1] = 0;
      p.t1 = p.t2 = {0, 0, 0};
      p.bias = 0;
      p.ni = c.normal_impulse;
      p.ti[0] = c.tangent_impulse[0];
      p.ti[1] = c.tangent_impulse[1];
      p.vn0 = 0;
      p.pen = c.pen;
    }
    cms.push_back(cm);
  }

  // src/solver.zig:564-605
  void applyGravity(float h) {
    for (BodyState& s : st) {
      if (s.type != VXP_BODY_DYNAMIC || s.sleeping) continue;
      s.lv = s.lv.add(gravity.scale(h));
      float ld = s.body->lin_damp, ad = s.body->ang_damp;
      float ls = s.lv.length(), as = s.av.length();
      if (ls < 0.5f) ld += 5.0f * (1.0f - (ls / 0.5f));
      if (as < 0.5f) ad += 5.0f * (1.0f - (as / 0.5f));
      float ldec = zmax(0.0f, 1.0f - ld * h), adec = zmax(0.0f, 1.0f - ad * h);
      s.lv = s.lv.scale(ldec);
      s.av = s.av.scale(adec);
      if (s.lv.lengthSquared() < 0.001f * 0.001f) s.lv = {0, 0, 0};
      if (s.av.lengthSquared() < 0.001f * 0.001f) s.av = {0, 0, 0};
    }
  }
  // src/solver.zig:608-699
  void prepare(const SoftParams& sp, float h) {
    float dt_inv = h > kEps ? 1.0f / h : 0.0f;
    for (CManifold& cm : cms) {
      BodyState& A = st[cm.ia];
      BodyState& B = st[cm.ib];
      for (int i = 0; i < cm.n; i++) {
        CPoint& p = cm.p[i];
        Vec3 wa = A.rot.rotate(p.la).add(A.pos);
        Vec3 wb = B.rot.rotate(p.lb).add(B.pos);
        p.ra = wa.sub(A.pos);
        p.rb = wb.sub(B.pos);
        p.point = wa.add(wb).scale(0.5f);
        Vec3 ad = wb.sub(wa);
        float sep = p.base_sep + ad.dot(p.normal);
        Vec3 rna = p.ra.cross(p.normal), rnb = p.rb.cross(p.normal);
        float k = A.inv_mass + B.inv_mass;
        k += rna.dot(A.inv_i.mulVec(rna));
        k += rnb.dot(B.inv_i.mulVec(rnb));
        p.normal_mass = k > kEps ? 1.0f / k : 0.0f;
        Vec3 rel = B.pointVel(p.point).sub(A.pointVel(p.point));
        float vn = rel.dot(p.normal);
        p.vn0 = vn;
        if (sep > 0.0f)
          p.bias = sep * dt_inv;
        else
          p.bias = zmax(sp.bias_rate * sep, -0.2f);
        if (cfg->enable_restitution && vn < -1.0f) p.bias = zmax(p.bias, -cm.restitution * vn);
        if (cfg->enable_friction) {
          p.t1 = computeOrthogonal(p.normal);
          p.t2 = p.normal.cross(p.t1);
          Vec3 a1 = p.ra.cross(p.t1), b1 = p.rb.cross(p.t1);
          float k1 = A.inv_mass + B.inv_mass;
          k1 += a1.dot(A.inv_i.mulVec(a1));
          k1 += b1.dot(B.inv_i.mulVec(b1));
          p.tangent_mass[0] = k1 > kEps ? 1.0f / k1 : 0.0f;
          Vec3 a2 = p.ra.cross(p.t2), b2 = p.rb.cross(p.t2);
          float k2 = A.inv_mass + B.inv_mass;
          k2 += a2.dot(A.inv_i.mulVec(a2));
          k2 += b2.dot(B.inv_i.mulVec(b2));
          p.tangent_mass[1] = k2 > kEps ? 1.0f / k2 : 0.0f;
        }
      }
    }
  }
  // src/solver.zig:702-729
  void warmStart() {
    float s = cfg->warm_start_scale;
    for (CManifold& cm : cms) {
      BodyState& A = st[cm.ia];
      BodyState& B = st[cm.ib];
      for (int i = 0; i < cm.n; i++) {
        CPoint& p = cm.p[i];
        float sn = p.ni * s, s1 = p.ti[0] * s, s2 = p.ti[1] * s;
        Vec3 imp = p.normal.scale(sn).add(p.t1.scale(s1)).add(p.t2.scale(s2));
        A.applyImpulseAtPoint(imp.negate(), p.ra);
        B.applyImpulseAtPoint(imp, p.rb);
        p.ni = sn;
        p.ti[0] = s1;
        p.ti[1] = s2;
      }
    }
  }
  // src/solver.zig:790-838
  void friction(const CManifold& cm, CPoint& p, BodyState& A, BodyState& B) {
    Vec3 rel = B.pointVel(p.point).sub(A.pointVel(p.point));
    float maxf = cm.friction * p.ni;
    {
      float vt = rel.dot(p.t1);
      float fi = -p.tangent_mass[0] * vt;
      float old = p.ti[0];
      p.ti[0] = zclamp(old + fi, -maxf, maxf);
      fi = p.ti[0] - old;
      Vec3 imp = p.t1.scale(fi);
      A.applyImpulseAtPoint(imp.negate(), p.ra);
      B.applyImpulseAtPoint(imp, p.rb);
    }
    {
      Vec3 rel2 = B.pointVel(p.point).sub(A.pointVel(p.point));
      float vt = rel2.dot(p.t2);
      float fi = -p.tangent_mass[1] * vt;
      float old = p.ti[1];
      p.ti[1] = zclamp(old + fi, -maxf, maxf);
      fi = p.ti[1] - old;
      Vec3 imp = p.t2.scale(fi);
      A.applyImpulseAtPoint(imp.negate(), p.ra);
      B.applyImpulseAtPoint(imp, p.rb);
    }
  }
  // src/solver.zig:732-787
  void solveVelocity(const SoftParams& sp, bool use_bias) {
    for (CManifold& cm : cms) {
      BodyState& A = st[cm.ia];
      BodyState& B = st[cm.ib];
      for (int i = 0; i < cm.n; i++) {
        CPoint& p = cm.p[i];
        Vec3 rel = B.pointVel(p.point).sub(A.pointVel(p.point));
        float vn = rel.dot(p.normal);
        float imp;
        if (use_bias)
          imp = -sp.mass_coeff * p.normal_mass * (vn + p.bias) - sp.impulse_coeff * p.ni;
        else
          imp = -p.normal_mass * vn;
        float nw = zmax(p.ni + imp, 0.0f);
        imp = nw - p.ni;
        p.ni = nw;
        Vec3 ni = p.normal.scale(imp);
        A.applyImpulseAtPoint(ni.negate(), p.ra);
        B.applyImpulseAtPoint(ni, p.rb);
        if (cfg->enable_friction && cm.friction > 0.0f) friction(cm, p, A, B);
      }
    }
  }
  // src/solver.zig:341-356
  void integrate(float h) {
    for (BodyState& s : st) {
      if (s.type == VXP_BODY_STATIC || s.sleeping) continue;
      s.pos = s.pos.add(s.lv.scale(h));
      if (!cfg->disable_rotation) s.rot = integrateRotation(s.rot, s.av, h);
      if (s.type == VXP_BODY_DYNAMIC) {
        Mat3 R = s.rot.toMat3();
        s.inv_i = R.mul(s.body->inv_inertia_local).mul(R.transpose());
      }
    }
  }
  // src/solver.zig:860-923
  bool solvePosition() {
    float min_sep = 0.0f;
    for (CManifold& cm : cms) {
      BodyState& A = st[cm.ia];
      BodyState& B = st[cm.ib];
      if (A.type == VXP_BODY_STATIC && B.type == VXP_BODY_STATIC) continue;
      for (int i = 0; i < cm.n; i++) {
        CPoint& p = cm.p[i];
        Vec3 wa = A.rot.rotate(p.la).add(A.pos);
        Vec3 wb = B.rot.rotate(p.lb).add(B.pos);
        Vec3 ad = wb.sub(wa);
        float sep = p.base_sep + ad.dot(p.normal);
        min_sep = zmin(min_sep, sep);
        max_pen = zmax(max_pen, -sep);
        if (sep >= -cfg->slop) continue;
        float corr = zmax(cfg->position_correction_rate * (sep + cfg->slop), -cfg->max_position_correction);
        Vec3 ra = wa.sub(A.pos), rb = wb.sub(B.pos);
        Vec3 rna = ra.cross(p.normal), rnb = rb.cross(p.normal);
        float k = A.inv_mass + B.inv_mass;
        k += rna.dot(A.inv_i.mulVec(rna));
        k += rnb.dot(B.inv_i.mulVec(rnb));
        if (k < kEps) continue;
        Vec3 P = p.normal.scale(-corr / k);
        if (A.type == VXP_BODY_DYNAMIC && !A.sleeping) {
          A.pos = A.pos.sub(P.scale(A.inv_mass));
          A.rot = integrateRotationSmall(A.rot, A.inv_i.mulVec(ra.cross(P.negate())));
        }
        if (B.type == VXP_BODY_DYNAMIC && !B.sleeping) {
          B.pos = B.pos.add(P.scale(B.inv_mass));
          B.rot = integrateRotationSmall(B.rot, B.inv_i.mulVec(rb.cross(P)));
        }
      }
    }
    return min_sep >= -3.0f * cfg->slop;
  }
  // src/solver.zig:494-561
  void solve(float dt) {
    if (dt <= 0.0f || st.empty()) return;
    bool has_c = !cms.empty();
    float h = dt / (float)cfg->substeps;
    SoftParams sp = computeSoftParams(cfg->hertz, cfg->damping_ratio, h);
    SoftParams hard{0.0f, 1.0f, 1.0f};
    for (uint32_t s = 0; s < cfg->substeps; s++) {
      applyGravity(h);
      if (has_c) {
        prepare(sp, h);
        if (cfg->warm_starting && s == 0) warmStart();
        for (uint32_t it = 0; it < cfg->velocity_iterations; it++) solveVelocity(sp, true);
      }
      integrate(h);
      if (has_c) solveVelocity(hard, false);
    }
    if (has_c) {
      for (uint32_t it = 0; it < cfg->position_iterations; it++) {
        pos_iters++;
        if (solvePosition()) break;
      }
    }
    for (BodyState& s : st) s.writeBack();
  }
};

// ---------------------------------------------------------------------------
// src/islands.zig — labels and the reference colouring as free functions
// ---------------------------------------------------------------------------
// UnionFind over manifold endpoints, uniting across edges where neither body is
// static (src/islands.zig:25-138,369-380); the label of a body is the minimum
// body id of its component.  Returns (body id, label) ascending by id.
std::vector<std::pair<uint32_t, uint32_t>> islandLabels(
    const std::vector<std::pair<uint32_t, uint32_t>>& edges,
    const std::function<bool(uint32_t)>& isStatic) {
  std::map<uint32_t, uint32_t> parent;
  std::function<uint32_t(uint32_t)> findp = [&](uint32_t x) {
    while (parent[x] != x) x = parent[x];
    return x;
  };
  for (auto& e : edges) {
    parent.emplace(e.first, e.first);
    parent.emplace(e.second, e.second);
  }
  for (auto& e : edges) {
    if (isStatic(e.first) || isStatic(e.second)) continue;
    uint32_t ra = findp(e.first), rb = findp(e.second);
    if (ra == rb) continue;
    if (ra < rb) parent[rb] = ra; else parent[ra] = rb;  // root = min id
  }
  std::vector<std::pair<uint32_t, uint32_t>> out;
  for (auto& kv : parent) out.push_back({kv.first, findp(kv.first)});
  return out;
}

// GraphColoring.build (src/islands.zig:668-736) over manifolds in the given
// order: static bodies count, colours saturate at 255.
std::vector<uint8_t> refColoring(const std::vector<std::pair<uint32_t, uint32_t>>& ms) {
  std::vector<uint8_t> colors;
  std::unordered_map<uint32_t, std::vector<uint32_t>> body_m;
  for (uint32_t i = 0; i < ms.size(); i++) {
    body_m[ms[i].first].push_back(i);
    body_m[ms[i].second].push_back(i);
  }
  for (uint32_t i = 0; i < ms.size(); i++) {
    bool used[256] = {false};
    for (uint32_t id : {ms[i].first, ms[i].second})
      for (uint32_t j : body_m[id])
        if (j < colors.size()) used[colors[j]] = true;
    uint32_t c = 0;
    while (c < 255 && used[c]) c++;
    colors.push_back((uint8_t)c);
  }
  return colors;
}

// ---------------------------------------------------------------------------
// src/broad_phase.zig:40-529 — AABBTree, restated node for node so that the
// pair order of findOverlappingPairs (and with it the 512-pair filter and the
// brick-pair cache order) can be reproduced
// ---------------------------------------------------------------------------
struct AabbTree {
  static constexpr uint32_t kNull = 0xFFFFFFFFu;  // NULL_NODE
  struct Node {
    AABB box;
    uint32_t user = 0, parent = kNull, child1 = kNull, child2 = kNull;
    int32_t height = 0;
    bool isLeaf() const { return child1 == kNull; }
  };
  std::vector<Node> nodes;
  uint32_t root = kNull, free_list = kNull, node_count = 0;
  std::map<uint32_t, uint32_t> body_to_node;

  // AABB.surfaceArea (src/math.zig:628-631)
  static float area(const AABB& b) {
    Vec3 s = b.max.sub(b.min);
    return 2.0f * (s.x * s.y + s.y * s.z + s.z * s.x);
  }
  uint32_t allocateNode() {  // src/broad_phase.zig:112-140
    uint32_t id;
    if (free_list != kNull) {
      id = free_list;
      free_list = nodes[id].parent;
    } else {
      id = (uint32_t)nodes.size();
      nodes.emplace_back();
    }
    nodes[id] = Node{};
    nodes[id].box = AABB::empty();
    node_count++;
    return id;
  }
  void freeNode(uint32_t id) {  // src/broad_phase.zig:142-147
    nodes[id].parent = free_list;
    nodes[id].height = -1;
    free_list = id;
    node_count--;
  }
  void insertBody(uint32_t body, const AABB& fat) {  // src/broad_phase.zig:150-163
    uint32_t id = allocateNode();
    nodes[id].box = fat;
    nodes[id].user = body;
    nodes[id].height = 0;
    body_to_node[body] = id;
    insertLeaf(id);
  }
  void removeBody(uint32_t body) {  // src/broad_phase.zig:166-172
    auto it = body_to_node.find(body);
    if (it == body_to_node.end()) return;
    removeLeaf(it->second);
    freeNode(it->second);
    body_to_node.erase(it);
  }
  void moveBody(uint32_t body, const AABB& fat) {  // src/broad_phase.zig:190-212
    uint32_t id = body_to_node.at(body);
    removeLeaf(id);
    nodes[id].box = fat;
    insertLeaf(id);
  }
  int32_t h(uint32_t i) const { return nodes[i].height; }
  void insertLeaf(uint32_t leaf) {  // src/broad_phase.zig:215-314
    if (root == kNull) {
      root = leaf;
      nodes[leaf].parent = kNull;
      return;
    }
    const AABB lb = nodes[leaf].box;
    uint32_t index = root;
    while (!nodes[index].isLeaf()) {
      const Node& n = nodes[index];
      const float a = area(n.box);
      const float ca = area(n.box.merge(lb));
      const float cost = 2.0f * ca;
      const float inh = 2.0f * (ca - a);
      auto childCost = [&](uint32_t c) {
        const float na = area(lb.merge(nodes[c].box));
        return nodes[c].isLeaf() ? na + inh : (na - area(nodes[c].box)) + inh;
      };
      const float c1 = childCost(n.child1), c2 = childCost(n.child2);
      if (cost < c1 && cost < c2) break;
      index = c1 < c2 ? n.child1 : n.child2;
    }
    const uint32_t sibling = index, old_parent = nodes[sibling].parent;
    const uint32_t np = allocateNode();
    nodes[np].parent = old_parent;
    nodes[np].user = 0;
    nodes[np].box = lb.merge(nodes[sibling].box);
    nodes[np].height = nodes[sibling].height + 1;
    if (old_parent != kNull) {
      if (nodes[old_parent].child1 == sibling)
        nodes[old_parent].child1 = np;
      else
        nodes[old_parent].child2 = np;
    } else {
      root = np;
    }
    nodes[np].child1 = sibling;
    nodes[np].child2 = leaf;
    nodes[sibling].parent = np;
    nodes[leaf].parent = np;
    for (uint32_t w = np; w != kNull; w = nodes[w].parent) {
      w = balance(w);
      const uint32_t c1 = nodes[w].child1, c2 = nodes[w].child2;
      nodes[w].height = 1 + std::max(h(c1), h(c2));
      nodes[w].box = nodes[c1].box.merge(nodes[c2].box);
    }
  }
  void removeLeaf(uint32_t leaf) {  // src/broad_phase.zig:317-362
    if (leaf == root) {
      root = kNull;
      return;
    }
    const uint32_t parent = nodes[leaf].parent, grand = nodes[parent].parent;
    const uint32_t sibling = nodes[parent].child1 == leaf ? nodes[parent].child2 : nodes[parent].child1;
    if (grand == kNull) {
      root = sibling;
      nodes[sibling].parent = kNull;
      freeNode(parent);
      return;
    }
    if (nodes[grand].child1 == parent)
      nodes[grand].child1 = sibling;
    else
      nodes[grand].child2 = sibling;
    nodes[sibling].parent = grand;
    freeNode(parent);
    for (uint32_t i = grand; i != kNull; i = nodes[i].parent) {
      i = balance(i);
      const uint32_t c1 = nodes[i].child1, c2 = nodes[i].child2;
      nodes[i].box = nodes[c1].box.merge(nodes[c2].box);
      nodes[i].height = 1 + std::max(h(c1), h(c2));
    }
  }
  uint32_t balance(uint32_t A) {  // src/broad_phase.zig:365-458
    if (nodes[A].isLeaf() || nodes[A].height < 2) return A;
    const uint32_t B = nodes[A].child1, C = nodes[A].child2;
    const int32_t bal = h(C) - h(B);
    if (bal > 1) {  // C moves up
      const uint32_t F = nodes[C].child1, G = nodes[C].child2;
      nodes[C].child1 = A;
      nodes[C].parent = nodes[A].parent;
      nodes[A].parent = C;
      relink(C, A);
      if (h(F) > h(G)) {
        nodes[C].child2 = F;
        nodes[A].child2 = G;
        nodes[G].parent = A;
        nodes[A].box = nodes[B].box.merge(nodes[G].box);
        nodes[C].box = nodes[A].box.merge(nodes[F].box);
        nodes[A].height = 1 + std::max(h(B), h(G));
        nodes[C].height = 1 + std::max(h(A), h(F));
      } else {
        nodes[C].child2 = G;
        nodes[A].child2 = F;
        nodes[F].parent = A;
        nodes[A].box = nodes[B].box.merge(nodes[F].box);
        nodes[C].box = nodes[A].box.merge(nodes[G].box);
        nodes[A].height = 1 + std::max(h(B), h(F));
        nodes[C].height = 1 + std::max(h(A), h(G));
      }
      return C;
    }
    if (bal < -1) {  // B moves up
      const uint32_t D = nodes[B].child1, E = nodes[B].child2;
      nodes[B].child1 = A;
      nodes[B].parent = nodes[A].parent;
      nodes[A].parent = B;
      relink(B, A);
      if (h(D) > h(E)) {
        nodes[B].child2 = D;
        nodes[A].child1 = E;
        nodes[E].parent = A;
        nodes[A].box = nodes[C].box.merge(nodes[E].box);
        nodes[B].box = nodes[A].box.merge(nodes[D].box);
        nodes[A].height = 1 + std::max(h(C), h(E));
        nodes[B].height = 1 + std::max(h(A), h(D));
      } else {
        nodes[B].child2 = E;
        nodes[A].child1 = D;
        nodes[D].parent = A;
        nodes[A].box = nodes[C].box.merge(nodes[D].box);
        nodes[B].box = nodes[A].box.merge(nodes[E].box);
        nodes[A].height = 1 + std::max(h(C), h(D));
        nodes[B].height = 1 + std::max(h(A), h(E));
      }
      return B;
    }
    return A;
  }
  // the parent of `up` (which now stands where `old` stood) points to `up`
  void relink(uint32_t up, uint32_t old) {
    const uint32_t p = nodes[up].parent;
    if (p == kNull) {
      root = up;
    } else if (nodes[p].child1 == old) {
      nodes[p].child1 = up;
    } else {
      nodes[p].child2 = up;
    }
  }
  // query (src/broad_phase.zig:461-485): depth first, child2 popped first
  void query(const AABB& q, std::vector<uint32_t>& out) const {
    if (root == kNull) return;
    std::vector<uint32_t> stack{root};
    while (!stack.empty()) {
      const uint32_t id = stack.back();
      stack.pop_back();
      if (id == kNull) continue;
      const Node& n = nodes[id];
      if (!n.box.overlaps(q)) continue;
      if (n.isLeaf()) {
        out.push_back(n.user);
      } else {
        stack.push_back(n.child1);
        stack.push_back(n.child2);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// src/system.zig — VoxelPhysicsSystem
// ---------------------------------------------------------------------------
struct Shared {
  std::vector<std::unique_ptr<Collider>> colliders;
};

struct World {
  vxp_config cfg;
  Shared* shared;
  std::vector<Slot> slots;
  uint32_t body_count = 0, next_id = 1;
  std::vector<uint32_t> free_list;
  std::vector<int32_t> id_to_slot{-1};
  float acc = 0.0f;
  vxp_stats stats{};
  // broad phase
  bool bp_dirty = true;
  std::vector<uint64_t> pairs;
  AabbTree static_tree, dynamic_tree;  // maintained when pair_order is the tree order
  // narrow phase
  std::map<uint64_t, Manifold> cur, prev;  // ContactManager maps; std::map iterates by key
  BrickPairCache cache;
  // IslandManager (src/islands.zig:196-270): islands.items, body_to_island,
  // stats and the rebuild bookkeeping persist across steps
  struct Island {
    std::vector<uint32_t> bodies;  // body ids, ascending
    bool sleeping = false;
    float sleep_timer = 0.0f;
  };
  uint64_t prev_graph_hash = 0;
  bool needs_rebuild = true;
  std::vector<Island> islands;
  std::map<uint32_t, uint32_t> body_island;  // body id -> island index
  vxp_island_stats mgr{};
  std::vector<std::pair<uint32_t, uint32_t>> island_labels;  // (body id, min id of island)
  // reference colouring artifact, pair-key order
  std::vector<uint8_t> ref_colors;
  Solver solver;

  const Collider* coll(const Body& b) const {
    return b.collider < 0 ? nullptr : shared->colliders[b.collider].get();
  }
  Body* find(uint32_t id) {
    if (id >= id_to_slot.size() || id_to_slot[id] < 0) return nullptr;
    return &slots[id_to_slot[id]].body;
  }
  // src/body.zig:430-435
  AABB worldAABB(const Body& b) const {
    const Collider* c = coll(b);
    if (c) return c->local_aabb.transform(b.xf());
    return {b.pos, b.pos};
  }

  // src/system.zig:475-499
  int allocate(vxp_handle* h) {
    uint32_t index;
    if (!free_list.empty()) {
      index = free_list.back();
      free_list.pop_back();
    } else {
      size_t i = 0;
      for (; i < slots.size(); i++)
        if (!slots[i].active) break;
      if (i == slots.size()) return VXP_ERR_CAPACITY;
      index = (uint32_t)i;
    }
    slots[index].active = true;
    body_count++;
    h->index = index;
    h->generation = slots[index].generation;
    return VXP_OK;
  }

  // src/system.zig:294-402
  // createDynamicBodyWithMass (src/system.zig:323-351) passes `given` instead
  // of computing the mass properties from the collider.
  int createBody(uint32_t type, Vec3 pos, Quat rot, int collider, float density, vxp_handle* h,
                 const MassProps* given = nullptr) {
    int rc = allocate(h);
    if (rc) return rc;
    Slot& s = slots[h->index];
    Body b;
    b.id = next_id;
    b.type = (uint8_t)type;
    b.pos = pos;
    b.rot = rot;
    b.collider = collider;
    if (type == VXP_BODY_DYNAMIC) {
      MassProps mp = given ? *given
                     : collider >= 0 ? computeMassProperties(*shared->colliders[collider], density)
                                     : MassProps{0.0f, Mat3::zero()};
      b.inv_mass = mp.inv_mass;
      b.inv_inertia_local = mp.inv_inertia;
      b.sleeping = false;
    } else {
      b.inv_mass = 0.0f;  // BodyProperties.static_props
      b.inv_inertia_local = Mat3::zero();
      b.lin_damp = 0.0f;
      b.ang_damp = 0.0f;
      b.sleeping = type == VXP_BODY_STATIC;
    }
    b.friction = cfg.default_friction;
    b.restitution = cfg.default_restitution;
    next_id++;
    b.fat = worldAABB(b).inflate(0.1f);  // insertBody (src/broad_phase.zig:150-163)
    if (cfg.pair_order == VXP_PAIR_ORDER_TREE)  // addStaticBody / addDynamicBody
      (type == VXP_BODY_STATIC ? static_tree : dynamic_tree).insertBody(b.id, b.fat);
    s.body = b;
    if (id_to_slot.size() <= b.id) id_to_slot.resize(b.id + 1, -1);
    id_to_slot[b.id] = (int32_t)h->index;
    bp_dirty = true;
    stats.total_bodies++;
    if (type == VXP_BODY_DYNAMIC) stats.dynamic_bodies++;
    if (type == VXP_BODY_STATIC) stats.static_bodies++;
    return VXP_OK;
  }

  bool valid(vxp_handle h) const {
    return h.index < slots.size() && slots[h.index].active && slots[h.index].generation == h.generation;
  }
  // src/system.zig:405-433
  int destroyBody(vxp_handle h) {
    if (!valid(h)) return VXP_ERR_INVALID_HANDLE;
    Slot& s = slots[h.index];
    stats.total_bodies--;
    if (s.body.type == VXP_BODY_DYNAMIC) stats.dynamic_bodies--;
    if (s.body.type == VXP_BODY_STATIC) stats.static_bodies--;
    id_to_slot[s.body.id] = -1;
    bp_dirty = true;
    if (cfg.pair_order == VXP_PAIR_ORDER_TREE) {  // BroadPhase.removeBody
      static_tree.removeBody(s.body.id);
      dynamic_tree.removeBody(s.body.id);
    }
    s.active = false;
    s.generation++;
    body_count--;
    free_list.push_back(h.index);
    return VXP_OK;
  }

  // src/system.zig:570-583 with src/broad_phase.zig:175-212 (fat-AABB rule)
  void broadPhase() {
    for (Slot& s : slots) {
      if (!s.active || s.body.type != VXP_BODY_DYNAMIC || s.body.sleeping) continue;
      Body& b = s.body;
      AABB aabb = worldAABB(b);
      AABB fat = aabb.inflate(0.1f);
      if (b.fat.containsAABB(aabb)) {
        AABB large = fat.inflate(0.1f * 4.0f);
        if (large.containsAABB(b.fat)) continue;
      }
      AABB ext = fat;
      Vec3 d = b.lv.scale(2.0f);
      if (d.x < 0) ext.min.x += d.x; else ext.max.x += d.x;
      if (d.y < 0) ext.min.y += d.y; else ext.max.y += d.y;
      if (d.z < 0) ext.min.z += d.z; else ext.max.z += d.z;
      b.fat = ext;
      if (cfg.pair_order == VXP_PAIR_ORDER_TREE) dynamic_tree.moveBody(b.id, ext);
      bp_dirty = true;
    }
  }

  // findOverlappingPairs (src/broad_phase.zig:488-526,606-619) in the trees'
  // order: every dynamic-tree leaf in node-index order queries the dynamic
  // tree (keeping partners of higher id), then every dynamic-tree leaf queries
  // the static tree.
  void findPairsTree() {
    pairs.clear();
    std::vector<uint32_t> hits;
    for (int pass = 0; pass < 2; pass++) {
      const AabbTree& other = pass == 0 ? dynamic_tree : static_tree;
      for (const AabbTree::Node& n : dynamic_tree.nodes) {
        if (n.height < 0 || !n.isLeaf()) continue;
        hits.clear();
        other.query(n.box, hits);
        for (uint32_t o : hits)
          if (pass == 1 || o > n.user) pairs.push_back(pairKey(n.user, o));
      }
    }
  }

  // Pair set of src/broad_phase.zig:488-526 (dynamic tree self-pairs, dynamic
  // leaves against the static tree), found by sort-and-sweep on fat min.x and
  // emitted in ascending pair key.
  void findPairs() {
    if (!bp_dirty) return;
    if (cfg.pair_order == VXP_PAIR_ORDER_TREE) {
      findPairsTree();
      bp_dirty = false;
      return;
    }
    struct It {
      float x0;
      uint32_t slot;
    };
    std::vector<It> items;
    for (uint32_t i = 0; i < slots.size(); i++)
      if (slots[i].active) items.push_back({slots[i].body.fat.min.x, i});
    // NaN boxes (a body whose collider is empty has an infinite local AABB, whose
    // transform is NaN) overlap nothing; they sort last so the order stays a
    // strict weak ordering and finite boxes keep their early exit.
    std::sort(items.begin(), items.end(), [](const It& a, const It& b) {
      bool na = std::isnan(a.x0), nb = std::isnan(b.x0);
      if (na != nb) return nb;
      if (na) return a.slot < b.slot;
      return a.x0 < b.x0 || (a.x0 == b.x0 && a.slot < b.slot);
    });
    pairs.clear();
    for (size_t i = 0; i < items.size(); i++) {
      const Body& bi = slots[items[i].slot].body;
      for (size_t j = i + 1; j < items.size(); j++) {
        const Body& bj = slots[items[j].slot].body;
        if (bj.fat.min.x > bi.fat.max.x) break;
        if (!bi.fat.overlaps(bj.fat)) continue;
        bool si = bi.type == VXP_BODY_STATIC, sj = bj.type == VXP_BODY_STATIC;
        if (si && sj) continue;
        pairs.push_back(pairKey(bi.id, bj.id));
      }
    }
    std::sort(pairs.begin(), pairs.end());
    bp_dirty = false;
  }

  // processPair minus the both-sleeping branch, which the system filter
  // already removed (src/narrow_phase.zig:1414-1525)
  struct PairWork {
    Body *a, *b;
    bool sat_pass = false, use_cache = false, hit = false, need_detect = false;
    CacheKey key{};
    CacheValue hit_value;
    DetectResult det;
  };

  void prepPair(PairWork& w, bool count_tested = true) {
    if (count_tested) stats.narrow_phase_tested++;
    const Collider* ca = coll(*w.a);
    const Collider* cb = coll(*w.b);
    if (!ca || !cb) return;
    Transform ta = w.a->xf(), tb = w.b->xf();
    if (cfg.use_brick_sat) {
      if (satEarlyOut(OBB::fromAABB(ca->local_aabb, ta), OBB::fromAABB(cb->local_aabb, tb))) return;
    }
    w.sat_pass = true;
    w.use_cache = cfg.use_brick_pair_cache &&
                  (w.a->type == VXP_BODY_STATIC || w.b->type == VXP_BODY_STATIC || w.a->sleeping ||
                   w.b->sleeping);
    if (w.use_cache) {
      Vec3 rp = tb.position.sub(ta.position);
      Quat rq = ta.rotation.conjugate().mul(tb.rotation);
      w.key = {std::min(w.a->id, w.b->id), std::max(w.a->id, w.b->id), computeTransformHash(rp, rq)};
    }
  }

  // ContactManager.addManifold (src/narrow_phase.zig:1011-1021,1049-1059)
  void addManifold(Manifold m) {
    uint64_t key = pairKey(m.a, m.b);
    auto it = prev.find(key);
    if (it != prev.end()) warmStartFrom(m, it->second);
    cur[key] = m;
  }

  void finishHit(PairWork& w) {
    const CacheValue& v = w.hit_value;
    stats.cache_hits++;
    if (!v.has_collision) return;
    Manifold m;
    m.a = w.a->id;
    m.b = w.b->id;
    m.friction = std::sqrt(w.a->friction * w.b->friction);
    m.restitution = zmax(w.a->restitution, w.b->restitution);
    for (int i = 0; i < v.count; i++) m.addContact(v.c[i]);
    stats.narrow_phase_colliding++;
    stats.contacts_generated += v.count;
    stats.voxels_checked += v.voxels_checked;
    addManifold(m);
  }
  void finishDetect(PairWork& w) {
    stats.voxels_checked += w.det.voxels_checked;
    stats.contact_candidates += w.det.candidates;
    if (w.det.has_collision) {
      stats.narrow_phase_colliding++;
      stats.contacts_generated += (uint32_t)w.det.m.n;
      addManifold(w.det.m);
    }
  }
  static CacheValue toCacheValue(const DetectResult& d) {
    CacheValue v;
    v.has_collision = d.has_collision;
    v.voxels_checked = d.voxels_checked;
    if (d.has_collision) {
      v.count = (uint8_t)std::min(d.m.n, 4);
      for (int i = 0; i < v.count; i++) v.c[i] = d.m.c[i];
    }
    return v;
  }

  // src/system.zig:586-667
  void narrowPhase() {
    std::swap(prev, cur);  // ContactManager.beginFrame
    cur.clear();
    cache.frame++;
    stats.narrow_phase_tested = stats.narrow_phase_colliding = stats.contacts_generated = 0;
    stats.voxels_checked = stats.cache_hits = stats.cache_misses = 0;
    stats.contact_candidates = 0;
    findPairs();
    stats.broad_phase_pairs = (uint32_t)pairs.size();
    std::vector<PairWork> work;
    uint32_t skipped = 0;
    for (uint64_t k : pairs) {
      Body* a = find((uint32_t)(k >> 32));
      Body* b = find((uint32_t)k);
      bool as = a ? a->sleeping : false, bs = b ? b->sleeping : false;
      if (as && bs) {
        skipped++;
        auto it = prev.find(k);
        if (it != prev.end()) cur[k] = it->second;
      } else if (cfg.pair_cap == 0 || work.size() < cfg.pair_cap) {
        if (a && b) work.push_back(PairWork{a, b});
      }
    }
    stats.pairs_skipped_sleeping = skipped;
    if (cfg.enable_threading) {
      threadPoolNarrowPhase(work);
      return;
    }
    if (cfg.cpu_threads <= 1) {
      // literal sequential processPair order
      for (PairWork& w : work) {
        prepPair(w);
        if (!w.sat_pass) continue;
        if (w.use_cache) {
          long e = cache.get(w.key);
          if (e >= 0) {
            w.hit_value = cache.e[e].value;
            finishHit(w);
            continue;
          }
          stats.cache_misses++;
          w.det = detectVoxelCollisions(*w.a, *coll(*w.a), *w.b, *coll(*w.b));
          cache.put(w.key, toCacheValue(w.det));
          finishDetect(w);
        } else {
          w.det = detectVoxelCollisions(*w.a, *coll(*w.a), *w.b, *coll(*w.b));
          finishDetect(w);
        }
      }
      return;
    }
    // Thread-pooled CPU baseline (src/narrow_phase.zig:1538-1632 semantics):
    // the cache gets/puts replay sequentially in pair order (puts reserve
    // their slot at miss time, values are filled after detection), detection
    // runs on worker threads, results merge in pair order.  Same results as
    // the sequential branch.
    std::vector<size_t> miss_slot(work.size(), SIZE_MAX);
    for (size_t i = 0; i < work.size(); i++) {
      PairWork& w = work[i];
      prepPair(w);
      if (!w.sat_pass) continue;
      if (w.use_cache) {
        long e = cache.get(w.key);
        if (e >= 0) {
          w.hit = true;
          w.hit_value = cache.e[e].value;
          continue;
        }
        miss_slot[i] = cache.put(w.key, CacheValue{});
      }
      w.need_detect = true;
    }
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        size_t i = next.fetch_add(1);
        if (i >= work.size()) break;
        PairWork& w = work[i];
        if (w.need_detect) w.det = detectVoxelCollisions(*w.a, *coll(*w.a), *w.b, *coll(*w.b));
      }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 1; t < cfg.cpu_threads; t++) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    for (size_t i = 0; i < work.size(); i++) {
      PairWork& w = work[i];
      if (!w.sat_pass) continue;
      if (w.hit) {
        finishHit(w);
        continue;
      }
      if (w.use_cache) {
        stats.cache_misses++;
        size_t e = miss_slot[i];
        if (cache.e[e].valid && cache.e[e].key.eql(w.key)) cache.e[e].value = toCacheValue(w.det);
      }
      finishDetect(w);
    }
  }

  // The narrow phase of a system with a thread pool (src/system.zig:618-638):
  // it sees only the first 256 active bodies in slot order; with fewer than
  // MIN_PARALLEL_PAIRS (8) pairs or one thread, processPair runs on the pairs
  // whose bodies are both listed (src/narrow_phase.zig:1546-1561); otherwise
  // each pair goes through processPairIntoBuffer (src/narrow_phase.zig:1187-1290):
  // pairs_tested counts every pair before the body lookup, the brick-pair
  // cache is read but never written, and the per-thread buffers (256
  // manifolds, src/narrow_phase.zig:35,1105-1110) merge in pair order.  No
  // buffer can fill: at most 512 pairs reach the narrow phase and two or more
  // threads split them into ranges of ceil(n / threads) <= 256 pairs.  Only
  // threads <= 1 versus >= 2 changes results; thread_count 0 (the host's CPU
  // count) counts as >= 2.
  void threadPoolNarrowPhase(std::vector<PairWork>& work) {
    std::unordered_set<const Body*> listed;
    for (Slot& sl : slots)
      if (sl.active && listed.size() < 256) listed.insert(&sl.body);
    const uint32_t threads = cfg.thread_count ? cfg.thread_count : 2u;
    const bool parallel = work.size() >= 8 && threads > 1;
    for (PairWork& w : work) {
      if (parallel) stats.narrow_phase_tested++;
      if (!listed.count(w.a) || !listed.count(w.b)) continue;
      prepPair(w, !parallel);
      if (!w.sat_pass) continue;
      if (w.use_cache) {
        long e = cache.get(w.key);
        if (e >= 0) {
          w.hit_value = cache.e[e].value;
          finishHit(w);
          continue;
        }
        stats.cache_misses++;
        w.det = detectVoxelCollisions(*w.a, *coll(*w.a), *w.b, *coll(*w.b));
        if (!parallel) cache.put(w.key, toCacheValue(w.det));
        finishDetect(w);
      } else {
        w.det = detectVoxelCollisions(*w.a, *coll(*w.a), *w.b, *coll(*w.b));
        finishDetect(w);
      }
    }
  }

  bool awakeDynamic(uint32_t id) {
    Body* b = find(id);
    return b && b->type == VXP_BODY_DYNAMIC && !b->sleeping;
  }

  // Solver colouring (see colorPrio above); returns colour count.
  uint32_t solverColoring(std::vector<const Manifold*>& ms, std::vector<uint32_t>& colors) {
    std::vector<size_t> order(ms.size());
    for (size_t i = 0; i < ms.size(); i++) order[i] = i;
    std::vector<std::pair<uint64_t, size_t>> pk(ms.size());
    for (size_t i = 0; i < ms.size(); i++)
      pk[i] = {((uint64_t)colorPrio(pairKey(ms[i]->a, ms[i]->b)) << 32) | (uint32_t)i, i};
    std::sort(pk.begin(), pk.end());
    std::unordered_map<uint32_t, std::vector<uint32_t>> used;  // body -> colours taken
    colors.assign(ms.size(), 0);
    uint32_t ncol = 0;
    for (auto& e : pk) {
      const Manifold* m = ms[e.second];
      std::vector<char> taken;
      for (uint32_t id : {m->a, m->b}) {
        if (!awakeDynamic(id)) continue;
        for (uint32_t c : used[id]) {
          if (taken.size() <= c) taken.resize(c + 1, 0);
          taken[c] = 1;
        }
      }
      uint32_t c = 0;
      while (c < taken.size() && taken[c]) c++;
      colors[e.second] = c;
      ncol = std::max(ncol, c + 1);
      for (uint32_t id : {m->a, m->b})
        if (awakeDynamic(id)) used[id].push_back(c);
    }
    return ncol;
  }

  // IslandManager.buildIslandsWithBodyTypes (src/islands.zig:318-441). The
  // graph hash runs over the manifolds in pair-key order (the order this
  // restatement keeps them in); islands are numbered by ascending label (the
  // minimum body id of the component), which no query observes.
  void buildIslands() {
    uint64_t h = 0;
    const uint64_t prime = 0x100000001b3ull;
    for (auto& kv : cur) {
      h ^= kv.first;
      h *= prime;
    }
    h ^= (uint64_t)cur.size();
    h *= prime;
    if (!needs_rebuild && h == prev_graph_hash && mgr.total_islands > 0) {
      mgr.rebuilds_skipped++;
      return;
    }
    prev_graph_hash = h;
    needs_rebuild = false;
    const uint32_t skipped = mgr.rebuilds_skipped, performed = mgr.rebuilds_performed + 1;
    for (Island& is : islands) is = Island{};  // Island.clear (src/islands.zig:176-184)
    body_island.clear();
    mgr = vxp_island_stats{};
    mgr.rebuilds_skipped = skipped;
    mgr.rebuilds_performed = performed;
    island_labels.clear();
    if (cur.empty()) return;
    uint32_t max_id = 0;
    for (Slot& s : slots)
      if (s.active) max_id = std::max(max_id, s.body.id);
    auto isStatic = [&](uint32_t id) {
      if (id > max_id) return false;
      Body* b = find(id);
      return b == nullptr || b->type == VXP_BODY_STATIC;
    };
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    for (auto& kv : cur) edges.push_back({kv.second.a, kv.second.b});
    island_labels = islandLabels(edges, isStatic);
    mgr.total_contacts = (uint32_t)cur.size();
    std::map<uint32_t, uint32_t> index;  // label -> island index
    for (auto& kv : island_labels) index.emplace(kv.second, (uint32_t)index.size());
    if (islands.size() < index.size()) islands.resize(index.size());
    for (auto& kv : island_labels) {
      const uint32_t i = index[kv.second];
      body_island[kv.first] = i;
      islands[i].bodies.push_back(kv.first);
    }
    mgr.total_islands = (uint32_t)index.size();
    for (uint32_t i = 0; i < mgr.total_islands; i++) {
      const uint32_t n = (uint32_t)islands[i].bodies.size();
      mgr.total_bodies += n;
      mgr.largest_island_size = std::max(mgr.largest_island_size, n);
    }
    mgr.active_islands = mgr.total_islands;  // every island was just cleared
  }

  // IslandManager.updateSleepStates (src/islands.zig:444-495), called with the
  // active bodies copied in slot order (src/system.zig:747-756): an island's
  // body id indexes that list, so id k reads the k-th active slot's body and
  // ids at or past the number of active bodies are passed over.
  void updateIslandSleep(float dt) {
    std::vector<const Body*> act;
    for (Slot& s : slots)
      if (s.active) act.push_back(&s.body);
    const float lt = 0.1f * 0.1f, at = 0.1f * 0.1f;  // ISLAND_SLEEP_*_THRESHOLD squared
    for (Island& is : islands) {
      if (is.bodies.empty()) continue;
      bool still = true;
      for (uint32_t id : is.bodies) {
        if (id >= act.size()) continue;
        const Body& b = *act[id];
        if (b.type != VXP_BODY_DYNAMIC) continue;
        if (b.lv.lengthSquared() > lt || b.av.lengthSquared() > at) {
          still = false;
          break;
        }
      }
      if (still) {
        is.sleep_timer += dt;
        if (is.sleep_timer >= 0.5f) is.sleeping = true;  // ISLAND_SLEEP_TIME_THRESHOLD
      } else {
        is.sleep_timer = 0.0f;
        is.sleeping = false;
      }
    }
    mgr.sleeping_islands = mgr.active_islands = 0;
    for (uint32_t i = 0; i < mgr.total_islands; i++) {
      if (islands[i].sleeping)
        mgr.sleeping_islands++;
      else
        mgr.active_islands++;
    }
  }

  // IslandManager.isBodyInSleepingIsland (src/islands.zig:554-562)
  bool inSleepingIsland(uint32_t id) const {
    auto it = body_island.find(id);
    return it != body_island.end() && it->second < islands.size() && islands[it->second].sleeping;
  }

  // IslandManager.wakeBodyIsland / wakeIsland (src/islands.zig:497-516)
  void wakeIsland(uint32_t id) {
    auto it = body_island.find(id);
    if (it != body_island.end() && it->second < islands.size()) {
      islands[it->second].sleeping = false;
      islands[it->second].sleep_timer = 0.0f;
    }
    needs_rebuild = true;
  }

  // GraphColoring.build (src/islands.zig:668-736) over the current manifolds
  // in pair-key order: static bodies count, colours saturate at 255.
  void buildRefColoring() {
    std::vector<std::pair<uint32_t, uint32_t>> ms;
    for (auto& kv : cur) ms.push_back({kv.second.a, kv.second.b});
    ref_colors = refColoring(ms);
  }

  // src/system.zig:670-783
  void solve(float dt) {
    solver.clear();
    solver.cfg = &cfg;
    solver.gravity = {cfg.gravity[0], cfg.gravity[1], cfg.gravity[2]};
    for (Slot& s : slots)
      if (s.active) solver.addBody(&s.body);
    std::vector<const Manifold*> ms;
    uint32_t skipped = 0;
    for (auto& kv : cur) {
      Body* a = find(kv.second.a);
      Body* b = find(kv.second.b);
      bool as = a ? a->sleeping : true, bs = b ? b->sleeping : true;
      kv.second.color = 0xFFFFFFFFu;
      if (as && bs) {
        skipped++;
        continue;
      }
      ms.push_back(&kv.second);
    }
    stats.manifolds_skipped_sleeping = skipped;
    std::vector<uint32_t> colors;
    uint32_t ncol = solverColoring(ms, colors);
    for (size_t i = 0; i < ms.size(); i++) const_cast<Manifold*>(ms[i])->color = colors[i];
    std::vector<size_t> order(ms.size());
    for (size_t i = 0; i < ms.size(); i++) order[i] = i;
    if (cfg.solve_order == VXP_SOLVE_ORDER_COLOR)
      std::stable_sort(order.begin(), order.end(),
                       [&](size_t x, size_t y) { return colors[x] < colors[y]; });
    for (size_t i : order) solver.addManifold(*ms[i]);
    stats.manifolds = (uint32_t)cur.size();
    stats.solver_manifolds = (uint32_t)solver.cms.size();
    stats.solver_points = 0;
    for (const CManifold& cm : solver.cms) stats.solver_points += (uint32_t)cm.n;
    stats.colors = ncol;
    if (cfg.enable_islands) {  // src/system.zig:705-757
      buildIslands();
      stats.island_count = mgr.total_islands;
      stats.active_islands = mgr.active_islands;
      stats.sleeping_islands = mgr.sleeping_islands;
      stats.largest_island_size = mgr.largest_island_size;
      updateIslandSleep(dt);
    }
    if (cfg.enable_coloring) buildRefColoring();
    solver.solve(dt);
    stats.max_penetration = solver.max_pen;
    stats.position_iterations_run = solver.pos_iters;
    // storeImpulses (src/solver.zig:933-958)
    if (cfg.enable_warm_starting && !solver.cms.empty()) {
      for (CManifold& cm : solver.cms) {
        auto it = cur.find(cm.key);
        if (it == cur.end()) continue;
        Manifold& m = it->second;
        int cnt = std::min(cm.n, m.n);
        for (int i = 0; i < cnt; i++) {
          m.c[i].normal_impulse = cm.p[i].ni;
          m.c[i].tangent_impulse[0] = cm.p[i].ti[0];
          m.c[i].tangent_impulse[1] = cm.p[i].ti[1];
          m.c[i].normal_mass = cm.p[i].normal_mass;
          m.c[i].tangent_mass[0] = cm.p[i].tangent_mass[0];
          m.c[i].tangent_mass[1] = cm.p[i].tangent_mass[1];
        }
      }
    }
  }

  // src/system.zig:786-792, src/body.zig:415-427
  void updateSleepStates(float dt) {
    for (Slot& s : slots) {
      Body& b = s.body;
      if (!s.active || b.type != VXP_BODY_DYNAMIC || b.sleeping) continue;
      if (b.lv.length() < 0.1f && b.av.length() < 0.1f) {
        b.sleep_timer += dt;
        if (b.sleep_timer >= 0.5f) {
          b.sleeping = true;
          b.lv = {0, 0, 0};
          b.av = {0, 0, 0};
        }
      } else {
        b.sleep_timer = 0.0f;
      }
    }
  }

  bool allDynamicSleeping() {
    for (Slot& s : slots)
      if (s.active && s.body.type == VXP_BODY_DYNAMIC && !s.body.sleeping) return false;
    return true;
  }

  // src/system.zig:537-555
  void step(float dt) {
    if (cfg.enable_sleeping && allDynamicSleeping()) return;
    stats.steps_executed++;
    for (Slot& s : slots)
      if (s.active && s.body.type == VXP_BODY_DYNAMIC && !s.body.sleeping) stats.awake_bodies_stepped++;
    broadPhase();
    narrowPhase();
    solve(dt);
    if (cfg.enable_sleeping) updateSleepStates(dt);
  }

  // src/system.zig:508-534
  void update(float dt) {
    uint32_t keep_total = stats.total_bodies, keep_dyn = stats.dynamic_bodies,
             keep_st = stats.static_bodies;
    std::memset(&stats, 0, sizeof(stats));  // PhysicsStats.reset (src/system.zig:129-154)
    stats.total_bodies = keep_total;
    stats.dynamic_bodies = keep_dyn;
    stats.static_bodies = keep_st;
    acc += dt;
    stats.accumulated_time = acc;
    uint32_t n = 0;
    while (acc >= cfg.fixed_timestep && n < cfg.max_substeps) {
      step(cfg.fixed_timestep);
      acc -= cfg.fixed_timestep;
      n++;
    }
    if (acc > cfg.fixed_timestep * (float)cfg.max_substeps) acc = cfg.fixed_timestep;
    stats.substeps_this_frame = n;
    uint32_t sl = 0, ac = 0;
    for (Slot& s : slots)
      if (s.active && s.body.type == VXP_BODY_DYNAMIC) (s.body.sleeping ? sl : ac)++;
    stats.sleeping_bodies = sl;
    stats.active_bodies = ac;
  }
};

}  // namespace ora

// ---------------------------------------------------------------------------
// C API (mirrors include/vxp.h with the vxo_ prefix)
// ---------------------------------------------------------------------------
using namespace ora;

struct vxo_system {
  Shared shared;
  std::vector<std::unique_ptr<World>> worlds;
};

static thread_local std::string g_err;

// Fields that setGravity / setSolverConfig leave alone must not change.
static bool configChangeAllowed(const vxp_config& a, const vxp_config& b) {
  vxp_config x = a, y = b;
  auto clear = [](vxp_config& c) {
    for (float& g : c.gravity) g = 0.0f;
    c.substeps = c.velocity_iterations = c.position_iterations = 0;
    c.hertz = c.damping_ratio = c.position_correction_rate = c.max_position_correction = 0.0f;
    c.slop = c.warm_start_scale = 0.0f;
    c.disable_rotation = c.warm_starting = c.enable_friction = c.enable_restitution = 0;
    c.pad0 = 0;
    c.cpu_threads = 0;
  };
  clear(x);
  clear(y);
  return std::memcmp(&x, &y, sizeof(vxp_config)) == 0 && b.substeps > 0;
}

extern "C" {

const char* vxo_last_error(void) { return g_err.c_str(); }

int vxo_system_create(const vxp_config* cfg, uint32_t n_worlds, vxo_system** out) {
  if (!cfg || !out || n_worlds == 0 || cfg->pair_order > VXP_PAIR_ORDER_TREE) return VXP_ERR_ARGUMENT;
  auto* s = new vxo_system();
  for (uint32_t i = 0; i < n_worlds; i++) {
    auto w = std::make_unique<World>();
    w->cfg = *cfg;
    w->shared = &s->shared;
    w->slots.resize(cfg->max_bodies);
    w->cache.init(cfg->brick_pair_cache_size ? cfg->brick_pair_cache_size : 1024);
    s->worlds.push_back(std::move(w));
  }
  *out = s;
  return VXP_OK;
}

int vxo_system_destroy(vxo_system* s) {
  delete s;
  return VXP_OK;
}

int vxo_collider_create(vxo_system* s, float voxel_size, uint32_t n, const int32_t* bp,
                        const uint8_t* vox, uint32_t* out) {
  auto c = std::make_unique<Collider>();
  c->voxel_size = voxel_size;
  // setCollisionVoxel for every non-empty voxel, brick by brick in the given
  // order (src/collider.zig:232-239,267-273): the brick map is keyed by
  // hashBrickPos, so positions whose hashes collide share one brick, which
  // keeps the first position and takes the later writes.
  std::unordered_map<uint64_t, size_t> tmp;
  std::vector<Brick> merged;
  for (uint32_t i = 0; i < n; i++) {
    IVec3 p{bp[3 * i], bp[3 * i + 1], bp[3 * i + 2]};
    uint64_t h = hashBrickPos(p);
    auto it = tmp.find(h);
    if (it == tmp.end()) {
      Brick b{};
      b.pos = p;
      it = tmp.emplace(h, merged.size()).first;
      merged.push_back(b);
    }
    Brick& b = merged[it->second];
    const uint8_t* src = vox + 512ull * i;
    for (int j = 0; j < 512; j++)
      if (voxelType(src[j]) != kEmpty) b.v[j] = src[j];
  }
  std::sort(merged.begin(), merged.end(), [](const Brick& a, const Brick& b) {
    if (a.pos.z != b.pos.z) return a.pos.z < b.pos.z;
    if (a.pos.y != b.pos.y) return a.pos.y < b.pos.y;
    return a.pos.x < b.pos.x;
  });
  for (size_t i = 0; i < merged.size(); i++) c->map[hashBrickPos(merged[i].pos)] = (int)i;
  c->bricks = std::move(merged);
  c->recomputeMetadata();
  *out = (uint32_t)s->shared.colliders.size();
  s->shared.colliders.push_back(std::move(c));
  return VXP_OK;
}

// classifyFromNeighbors (src/collision_voxel.zig:213-235) for every voxel of
// the given bricks (occupancy: 8 words per brick, bit x + 8y of word z; a
// neighbour in a brick that is not given is empty), then the same collider
// build as vxo_collider_create.
int vxo_collider_create_solid(vxo_system* s, float voxel_size, uint32_t n, const int32_t* bp,
                              const uint64_t* solid, uint32_t* out) {
  if (!(voxel_size > 0.0f) || (n && (!bp || !solid))) return VXP_ERR_ARGUMENT;
  std::map<std::array<int32_t, 3>, uint32_t> at;
  for (uint32_t i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++)
      if (bp[3 * i + k] < -(1 << 20) || bp[3 * i + k] >= (1 << 20)) return VXP_ERR_ARGUMENT;
    if (!at.emplace(std::array<int32_t, 3>{bp[3 * i], bp[3 * i + 1], bp[3 * i + 2]}, i).second)
      return VXP_ERR_ARGUMENT;
  }
  // solid voxel at global voxel coordinates
  auto isSolid = [&](int32_t x, int32_t y, int32_t z) {
    auto it = at.find({floorDiv8(x), floorDiv8(y), floorDiv8(z)});
    if (it == at.end()) return false;
    const int lx = x - floorDiv8(x) * 8, ly = y - floorDiv8(y) * 8, lz = z - floorDiv8(z) * 8;
    return ((solid[(size_t)it->second * 8 + lz] >> (lx + 8 * ly)) & 1ull) != 0;
  };
  std::vector<uint8_t> vox((size_t)n * 512, 0);
  for (uint32_t i = 0; i < n; i++)
    for (int v = 0; v < 512; v++) {
      const int32_t x = bp[3 * i] * 8 + (v & 7), y = bp[3 * i + 1] * 8 + ((v >> 3) & 7),
                    z = bp[3 * i + 2] * 8 + (v >> 6);
      if (!isSolid(x, y, z)) continue;
      const bool nb[6] = {isSolid(x - 1, y, z), isSolid(x + 1, y, z), isSolid(x, y - 1, z),
                          isSolid(x, y + 1, z), isSolid(x, y, z - 1), isSolid(x, y, z + 1)};
      uint32_t empty = 0;
      for (int k = 0; k < 6; k++)
        if (!nb[k]) empty |= 1u << k;
      vox[(size_t)i * 512 + v] = classifyVoxel(true, empty);
    }
  return vxo_collider_create(s, voxel_size, n, bp, vox.data(), out);
}

int vxo_body_create_w(vxo_system* s, uint32_t world, uint32_t type, const float pos[3],
                      const float rot[4], uint32_t collider, float density, vxp_handle* out) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  if (collider != 0xFFFFFFFFu && collider >= s->shared.colliders.size()) return VXP_ERR_ARGUMENT;
  return s->worlds[world]->createBody(type, Vec3{pos[0], pos[1], pos[2]},
                                      Quat{rot[0], rot[1], rot[2], rot[3]},
                                      collider == 0xFFFFFFFFu ? -1 : (int)collider, density, out);
}
int vxo_body_create(vxo_system* s, uint32_t type, const float pos[3], const float rot[4],
                    uint32_t collider, float density, vxp_handle* out) {
  return vxo_body_create_w(s, 0, type, pos, rot, collider, density, out);
}
int vxo_body_create_with_mass_w(vxo_system* s, uint32_t world, const float pos[3],
                                const float rot[4], uint32_t collider, float inv_mass,
                                const float inv_inertia[9], vxp_handle* out) {
  if (!s || !out || !pos || !rot || !inv_inertia || world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  if (collider != 0xFFFFFFFFu && collider >= s->shared.colliders.size()) return VXP_ERR_ARGUMENT;
  MassProps mp;
  mp.inv_mass = inv_mass;
  for (int i = 0; i < 9; i++) mp.inv_inertia.m[i / 3][i % 3] = inv_inertia[i];
  return s->worlds[world]->createBody(VXP_BODY_DYNAMIC, Vec3{pos[0], pos[1], pos[2]},
                                      Quat{rot[0], rot[1], rot[2], rot[3]},
                                      collider == 0xFFFFFFFFu ? -1 : (int)collider, 1.0f, out, &mp);
}
// setGravity / setSolverConfig (src/system.zig:1034-1052)
int vxo_system_set_config(vxo_system* s, const vxp_config* cfg) {
  if (!s || !cfg || s->worlds.empty()) return VXP_ERR_ARGUMENT;
  if (!configChangeAllowed(s->worlds[0]->cfg, *cfg)) {
    g_err = "only gravity and the SolverConfig fields can change after creation";
    return VXP_ERR_ARGUMENT;
  }
  for (auto& w : s->worlds) w->cfg = *cfg;
  return VXP_OK;
}
int vxo_body_destroy_w(vxo_system* s, uint32_t world, vxp_handle h) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  return s->worlds[world]->destroyBody(h);
}

int vxo_system_update(vxo_system* s, float dt) {
  for (auto& w : s->worlds) w->update(dt);
  return VXP_OK;
}
int vxo_system_step(vxo_system* s, float dt) {
  for (auto& w : s->worlds) w->step(dt);
  return VXP_OK;
}
int vxo_system_update_frames(vxo_system* s, float dt, uint32_t frames) {
  for (uint32_t f = 0; f < frames; f++)
    for (auto& w : s->worlds) w->update(dt);
  return VXP_OK;
}

int vxo_bodies_get_w(vxo_system* s, uint32_t world, vxp_body_state* out, uint32_t cap,
                     uint32_t* count) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  uint32_t n = (uint32_t)std::min<size_t>(cap, w.slots.size());
  for (uint32_t i = 0; i < n; i++) {
    const Slot& sl = w.slots[i];
    const Body& b = sl.body;
    vxp_body_state& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.active = sl.active;
    if (!sl.active) continue;
    o.position[0] = b.pos.x; o.position[1] = b.pos.y; o.position[2] = b.pos.z;
    o.rotation[0] = b.rot.x; o.rotation[1] = b.rot.y; o.rotation[2] = b.rot.z; o.rotation[3] = b.rot.w;
    o.linear_velocity[0] = b.lv.x; o.linear_velocity[1] = b.lv.y; o.linear_velocity[2] = b.lv.z;
    o.angular_velocity[0] = b.av.x; o.angular_velocity[1] = b.av.y; o.angular_velocity[2] = b.av.z;
    o.sleep_timer = b.sleep_timer;
    o.id = b.id;
    o.body_type = b.type;
    o.is_sleeping = b.sleeping;
  }
  *count = n;
  return VXP_OK;
}

int vxo_body_set_state_w(vxo_system* s, uint32_t world, vxp_handle h, const float pos[3],
                         const float rot[4], const float lv[3], const float av[3]) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  if (!w.valid(h)) return VXP_ERR_INVALID_HANDLE;
  Body& b = w.slots[h.index].body;
  if (pos) b.pos = {pos[0], pos[1], pos[2]};
  if (rot) b.rot = {rot[0], rot[1], rot[2], rot[3]};
  if (b.type == VXP_BODY_DYNAMIC) {
    if (lv) b.lv = {lv[0], lv[1], lv[2]};
    if (av) b.av = {av[0], av[1], av[2]};
  }
  if (b.type != VXP_BODY_STATIC) {  // wakeUp (src/body.zig:400-404)
    b.sleeping = false;
    b.sleep_timer = 0.0f;
  }
  return VXP_OK;
}

// VoxelPhysicsSystem.applyForce ... sleep (src/system.zig:886-961) forwarding
// to the VoxelBody methods of src/body.zig:266-414.
int vxo_body_op_w(vxo_system* s, uint32_t world, vxp_handle h, uint32_t op, const float* v,
                  const float point[3]) {
  if (world >= s->worlds.size() || op > VXP_OP_SLEEP) return VXP_ERR_ARGUMENT;
  if (op != VXP_OP_WAKE_UP && op != VXP_OP_SLEEP && !v) return VXP_ERR_ARGUMENT;
  if ((op == VXP_OP_APPLY_FORCE_AT_POINT || op == VXP_OP_APPLY_IMPULSE_AT_POINT) && !point)
    return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  if (!w.valid(h)) return VXP_ERR_INVALID_HANDLE;
  Body& b = w.slots[h.index].body;
  const bool dyn = b.type == VXP_BODY_DYNAMIC;
  // wakeUp (src/body.zig:400-404)
  auto wakeUp = [&] {
    if (b.type == VXP_BODY_STATIC) return;
    b.sleeping = false;
    b.sleep_timer = 0.0f;
  };
  // applyForce / applyImpulse (src/body.zig:327-333, 355-361)
  auto applyLinear = [&](Vec3 f) {
    if (!dyn || b.inv_mass == 0.0f) return;
    b.lv = b.lv.add(f.scale(b.inv_mass));
    wakeUp();
  };
  // applyTorque / applyAngularImpulse (src/body.zig:346-353, 374-381)
  auto applyAngular = [&](Vec3 t) {
    if (!dyn) return;
    b.av = b.av.add(b.worldInvInertia().mulVec(t));
    wakeUp();
  };
  Vec3 a{0, 0, 0};
  if (v) a = {v[0], v[1], v[2]};
  switch (op) {
    case VXP_OP_APPLY_FORCE:
    case VXP_OP_APPLY_IMPULSE:
      applyLinear(a);
      break;
    case VXP_OP_APPLY_FORCE_AT_POINT:  // src/body.zig:336-344
    case VXP_OP_APPLY_IMPULSE_AT_POINT:  // src/body.zig:364-372
      if (!dyn) break;
      applyLinear(a);
      applyAngular(Vec3{point[0], point[1], point[2]}.sub(b.pos).cross(a));
      break;
    case VXP_OP_APPLY_TORQUE:
    case VXP_OP_APPLY_ANGULAR_IMPULSE:
      applyAngular(a);
      break;
    case VXP_OP_SET_LINEAR_VELOCITY:  // src/body.zig:304-308
      if (!dyn) break;
      b.lv = a;
      wakeUp();
      break;
    case VXP_OP_SET_ANGULAR_VELOCITY:  // src/body.zig:316-320
      if (!dyn) break;
      b.av = a;
      wakeUp();
      break;
    case VXP_OP_SET_POSITION:  // src/body.zig:281-284
      b.pos = a;
      wakeUp();
      break;
    case VXP_OP_SET_ROTATION:  // src/body.zig:292-295
      b.rot = {v[0], v[1], v[2], v[3]};
      wakeUp();
      break;
    case VXP_OP_WAKE_UP:
      wakeUp();
      break;
    case VXP_OP_SLEEP:  // src/body.zig:407-412
      if (b.type == VXP_BODY_STATIC) break;
      b.sleeping = true;
      b.lv = {0, 0, 0};
      b.av = {0, 0, 0};
      break;
  }
  return VXP_OK;
}

// VoxelPhysicsSystem.getWorldAABB (src/system.zig:1003-1008)
int vxo_body_world_aabb_w(vxo_system* s, uint32_t world, vxp_handle h, float out[6]) {
  if (world >= s->worlds.size() || !out) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  if (!w.valid(h)) return VXP_ERR_INVALID_HANDLE;
  AABB a = w.worldAABB(w.slots[h.index].body);
  const float r[6] = {a.min.x, a.min.y, a.min.z, a.max.x, a.max.y, a.max.z};
  std::memcpy(out, r, sizeof(r));
  return VXP_OK;
}

// VoxelPhysicsSystem.findBodyById (src/system.zig:462-473): a linear scan
int vxo_body_find_w(vxo_system* s, uint32_t world, uint32_t id, vxp_handle* out) {
  if (world >= s->worlds.size() || !out) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  for (size_t i = 0; i < w.slots.size(); i++)
    if (w.slots[i].active && w.slots[i].body.id == id) {
      out->index = (uint32_t)i;
      out->generation = w.slots[i].generation;
      return VXP_OK;
    }
  return VXP_ERR_INVALID_HANDLE;
}

int vxo_tree_get_w(vxo_system* s, uint32_t world, uint32_t which, vxp_tree_node* out, uint32_t cap,
                   uint32_t* count, uint32_t* root) {
  if (world >= s->worlds.size() || which > 1 || !count || !root) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  if (w.cfg.pair_order != VXP_PAIR_ORDER_TREE) return VXP_ERR_ARGUMENT;
  const AabbTree& t = which ? w.dynamic_tree : w.static_tree;
  *count = (uint32_t)t.nodes.size();
  *root = t.root;
  for (uint32_t i = 0; out && i < *count && i < cap; i++) {
    const AabbTree::Node& n = t.nodes[i];
    vxp_tree_node& o = out[i];
    o.min[0] = n.box.min.x;
    o.min[1] = n.box.min.y;
    o.min[2] = n.box.min.z;
    o.max[0] = n.box.max.x;
    o.max[1] = n.box.max.y;
    o.max[2] = n.box.max.z;
    o.body_id = n.user;
    o.parent = n.parent;
    o.child1 = n.child1;
    o.child2 = n.child2;
    o.height = n.height;
    o.pad = 0;
  }
  return VXP_OK;
}

int vxo_island_stats_get_w(vxo_system* s, uint32_t world, vxp_island_stats* out) {
  if (world >= s->worlds.size() || !out || !s->worlds[world]->cfg.enable_islands) return VXP_ERR_ARGUMENT;
  *out = s->worlds[world]->mgr;
  return VXP_OK;
}

int vxo_body_in_sleeping_island_w(vxo_system* s, uint32_t world, vxp_handle h, int* out) {
  if (world >= s->worlds.size() || !out) return VXP_ERR_ARGUMENT;
  *out = 0;
  World& w = *s->worlds[world];
  if (!w.valid(h)) return VXP_ERR_INVALID_HANDLE;
  if (w.cfg.enable_islands) *out = w.inSleepingIsland(w.slots[h.index].body.id) ? 1 : 0;
  return VXP_OK;
}

int vxo_body_wake_island_w(vxo_system* s, uint32_t world, vxp_handle h) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  if (!w.valid(h)) return VXP_ERR_INVALID_HANDLE;
  if (w.cfg.enable_islands) w.wakeIsland(w.slots[h.index].body.id);
  return vxo_body_op_w(s, world, h, VXP_OP_WAKE_UP, nullptr, nullptr);
}

int vxo_stats_get_w(vxo_system* s, uint32_t world, vxp_stats* out) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  *out = s->worlds[world]->stats;
  return VXP_OK;
}

int vxo_stats_get_all(vxo_system* s, vxp_stats* out, uint32_t cap) {
  if (cap < s->worlds.size()) return VXP_ERR_ARGUMENT;
  for (size_t w = 0; w < s->worlds.size(); w++) out[w] = s->worlds[w]->stats;
  return VXP_OK;
}

static void exportContact(const Contact& c, vxp_contact& o) {
  std::memset(&o, 0, sizeof(o));
  o.point[0] = c.point.x; o.point[1] = c.point.y; o.point[2] = c.point.z;
  o.normal[0] = c.normal.x; o.normal[1] = c.normal.y; o.normal[2] = c.normal.z;
  o.penetration = c.pen;
  o.local_anchor_a[0] = c.anchor_a.x; o.local_anchor_a[1] = c.anchor_a.y; o.local_anchor_a[2] = c.anchor_a.z;
  o.local_anchor_b[0] = c.anchor_b.x; o.local_anchor_b[1] = c.anchor_b.y; o.local_anchor_b[2] = c.anchor_b.z;
  o.normal_impulse = c.normal_impulse;
  o.tangent_impulse[0] = c.tangent_impulse[0];
  o.tangent_impulse[1] = c.tangent_impulse[1];
  o.normal_mass = c.normal_mass;
  o.tangent_mass[0] = c.tangent_mass[0];
  o.tangent_mass[1] = c.tangent_mass[1];
  o.feature_id = c.feature_id;
}

int vxo_manifolds_get_w(vxo_system* s, uint32_t world, vxp_manifold* out, uint32_t cap,
                        uint32_t* count) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  uint32_t n = 0;
  for (auto& kv : w.cur) {
    if (n < cap && out) {
      const Manifold& m = kv.second;
      vxp_manifold& o = out[n];
      std::memset(&o, 0, sizeof(o));
      o.body_a_id = m.a;
      o.body_b_id = m.b;
      o.friction = m.friction;
      o.restitution = m.restitution;
      o.contact_count = (uint32_t)m.n;
      o.color = m.color;
      for (int i = 0; i < m.n; i++) exportContact(m.c[i], o.contacts[i]);
    }
    n++;
  }
  *count = n;
  return VXP_OK;
}

int vxo_pairs_get_w(vxo_system* s, uint32_t world, uint64_t* out, uint32_t cap, uint32_t* count) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  uint32_t n = (uint32_t)w.pairs.size();
  for (uint32_t i = 0; i < n && i < cap; i++) out[i] = w.pairs[i];
  *count = n;
  return VXP_OK;
}

int vxo_islands_get_w(vxo_system* s, uint32_t world, uint32_t* ids, uint32_t* labels, uint32_t cap,
                      uint32_t* count) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  uint32_t n = (uint32_t)w.island_labels.size();
  for (uint32_t i = 0; i < n && i < cap; i++) {
    ids[i] = w.island_labels[i].first;
    labels[i] = w.island_labels[i].second;
  }
  *count = n;
  return VXP_OK;
}

int vxo_coloring_get_w(vxo_system* s, uint32_t world, uint8_t* colors, uint32_t cap,
                       uint32_t* count) {
  if (world >= s->worlds.size()) return VXP_ERR_ARGUMENT;
  World& w = *s->worlds[world];
  uint32_t n = (uint32_t)w.ref_colors.size();
  for (uint32_t i = 0; i < n && i < cap; i++) colors[i] = w.ref_colors[i];
  *count = n;
  return VXP_OK;
}

// ---- unit-level entry points for the known-answer tests ----------------------
void vxo_quat_rotate(const float q[4], const float v[3], float out[3]) {
  Vec3 r = Quat{q[0], q[1], q[2], q[3]}.rotate({v[0], v[1], v[2]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void vxo_quat_from_axis_angle(const float axis[3], float angle, float out[4]) {
  Quat q = Quat::fromAxisAngle({axis[0], axis[1], axis[2]}, angle);
  out[0] = q.x; out[1] = q.y; out[2] = q.z; out[3] = q.w;
}
void vxo_integrate_rotation(const float q[4], const float w[3], float dt, float out[4]) {
  Quat r = integrateRotation({q[0], q[1], q[2], q[3]}, {w[0], w[1], w[2]}, dt);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
void vxo_mat3_inverse(const float m[9], float out[9]) {
  Mat3 a;
  std::memcpy(a.m, m, sizeof(a.m));
  Mat3 r = a.inverse();
  std::memcpy(out, r.m, sizeof(r.m));
}
uint32_t vxo_classify_voxel(uint32_t solid, uint32_t empty_mask) {
  return classifyVoxel(solid != 0, empty_mask);
}
uint32_t vxo_compute_normal_index(uint32_t empty_mask) { return computeNormalIndex(empty_mask); }
uint32_t vxo_should_check(uint32_t ta, uint32_t tb) { return shouldCheckCollision(ta, tb); }
uint64_t vxo_feature_id(const int32_t a[3], const int32_t b[3]) {
  return computeFeatureId({a[0], a[1], a[2]}, {b[0], b[1], b[2]});
}
void vxo_soft_params(float hertz, float zeta, float dt, float out[3]) {
  SoftParams p = computeSoftParams(hertz, zeta, dt);
  out[0] = p.bias_rate; out[1] = p.mass_coeff; out[2] = p.impulse_coeff;
}
void vxo_compute_orthogonal(const float v[3], float out[3]) {
  Vec3 r = computeOrthogonal({v[0], v[1], v[2]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
// SAT on two OBBs given as center[3], half[3], quat[4] each
int vxo_sat_early_out(const float a[10], const float b[10]) {
  OBB oa{{a[0], a[1], a[2]}, {a[3], a[4], a[5]}, {a[6], a[7], a[8], a[9]}};
  OBB ob{{b[0], b[1], b[2]}, {b[3], b[4], b[5]}, {b[6], b[7], b[8], b[9]}};
  return satEarlyOut(oa, ob) ? 1 : 0;
}
// Streaming top-4 manifold: feed n contacts (pen[i], feature[i]); outputs
// slot -> input index.
int vxo_manifold_replay(uint32_t n, const float* pen, const uint64_t* fid, int32_t out_idx[4]) {
  Manifold m;
  for (uint32_t i = 0; i < n; i++) {
    Contact c{};
    c.pen = pen[i];
    c.feature_id = fid[i];
    c.point = {(float)i, 0, 0};
    m.addContact(c);
  }
  for (int i = 0; i < 4; i++) out_idx[i] = i < m.n ? (int32_t)m.c[i].point.x : -1;
  return m.n;
}
// BrickPairCache scripted operations for the known-answer tests:
// ops[i] = 0 get / 1 put / 2 begin frame; keys (a,b,rel_pos[3],rel_rot[4]).
// result[i] = hit entry or written entry (get miss: -1).
void vxo_cache_script(uint32_t cap, uint32_t n, const uint32_t* ops, const uint32_t* ab,
                      const float* xf, int64_t* result) {
  BrickPairCache c;
  c.init(cap);
  for (uint32_t i = 0; i < n; i++) {
    CacheKey k{std::min(ab[2 * i], ab[2 * i + 1]), std::max(ab[2 * i], ab[2 * i + 1]),
               computeTransformHash({xf[7 * i], xf[7 * i + 1], xf[7 * i + 2]},
                                    {xf[7 * i + 3], xf[7 * i + 4], xf[7 * i + 5], xf[7 * i + 6]})};
    if (ops[i] == 0)
      result[i] = c.get(k);
    else if (ops[i] == 1)
      result[i] = (int64_t)c.put(k, CacheValue{});
    else {
      c.frame++;
      result[i] = 0;
    }
  }
}
uint32_t vxo_color_prio(uint64_t key) { return colorPrio(key); }
// VoxelCollider.voxelToBrickPos / voxelToLocalCoord (src/collider.zig:195-210)
void vxo_voxel_to_brick(const int32_t v[3], int32_t brick[3], int32_t local[3]) {
  for (int k = 0; k < 3; k++) {
    brick[k] = floorDiv8(v[k]);
    local[k] = mod8(v[k]);
  }
}
// warm start of one manifold from the previous frame's (impulses 3 per contact)
void vxo_warm_start(uint32_t n_new, const uint64_t* new_fid, uint32_t n_old, const uint64_t* old_fid,
                    const float* old_imp, float* out_imp) {
  Manifold m, o;
  m.n = (int)std::min(n_new, 4u);
  o.n = (int)std::min(n_old, 4u);
  for (int i = 0; i < m.n; i++) m.c[i].feature_id = new_fid[i];
  for (int j = 0; j < o.n; j++) {
    o.c[j].feature_id = old_fid[j];
    o.c[j].normal_impulse = old_imp[3 * j];
    o.c[j].tangent_impulse[0] = old_imp[3 * j + 1];
    o.c[j].tangent_impulse[1] = old_imp[3 * j + 2];
  }
  warmStartFrom(m, o);
  for (int i = 0; i < m.n; i++) {
    out_imp[3 * i] = m.c[i].normal_impulse;
    out_imp[3 * i + 1] = m.c[i].tangent_impulse[0];
    out_imp[3 * i + 2] = m.c[i].tangent_impulse[1];
  }
}
// Island labels of an edge list (src/islands.zig:318-441); is_static[id] for
// ids < n_ids.  Writes (id, label) ascending; returns the count.
uint32_t vxo_island_labels(uint32_t n, const uint32_t* a, const uint32_t* b, uint32_t n_ids,
                           const uint8_t* is_static, uint32_t* ids, uint32_t* labels) {
  std::vector<std::pair<uint32_t, uint32_t>> e;
  for (uint32_t i = 0; i < n; i++) e.push_back({a[i], b[i]});
  auto st = [&](uint32_t id) { return id < n_ids && is_static[id] != 0; };
  auto out = islandLabels(e, st);
  for (size_t i = 0; i < out.size(); i++) {
    ids[i] = out[i].first;
    labels[i] = out[i].second;
  }
  return (uint32_t)out.size();
}
void vxo_ref_coloring(uint32_t n, const uint32_t* a, const uint32_t* b, uint8_t* out) {
  std::vector<std::pair<uint32_t, uint32_t>> ms;
  for (uint32_t i = 0; i < n; i++) ms.push_back({a[i], b[i]});
  auto c = refColoring(ms);
  for (uint32_t i = 0; i < n; i++) out[i] = c[i];
}
// Mass properties of a collider: inv_mass, inv_inertia[9], com[3]
int vxo_mass_properties(vxo_system* s, uint32_t collider, float density, float out[13]) {
  if (collider >= s->shared.colliders.size()) return VXP_ERR_ARGUMENT;
  const Collider& c = *s->shared.colliders[collider];
  MassProps mp = computeMassProperties(c, density);
  out[0] = mp.inv_mass;
  std::memcpy(out + 1, mp.inv_inertia.m, 36);
  Vec3 com = c.getCenterOfMass();
  out[10] = com.x; out[11] = com.y; out[12] = com.z;
  return VXP_OK;
}
int vxo_collider_bricks(vxo_system* s, uint32_t collider, int32_t* bp, uint8_t* vox, uint32_t cap,
                        uint32_t* count) {
  if (collider >= s->shared.colliders.size() || !count) return VXP_ERR_ARGUMENT;
  const Collider& c = *s->shared.colliders[collider];
  *count = (uint32_t)c.bricks.size();
  if (cap && bp && vox)
    for (uint32_t i = 0; i < std::min<uint32_t>(cap, (uint32_t)c.bricks.size()); i++) {
      bp[3 * i] = c.bricks[i].pos.x;
      bp[3 * i + 1] = c.bricks[i].pos.y;
      bp[3 * i + 2] = c.bricks[i].pos.z;
      std::memcpy(vox + (size_t)i * 512, c.bricks[i].v, 512);
    }
  return VXP_OK;
}
int vxo_collider_mass_properties(vxo_system* s, uint32_t collider, float density, float out[13]) {
  return vxo_mass_properties(s, collider, density, out);
}
// Collider metadata: min_bound[3], max_bound[3], local_aabb[6] (as float bits in 6..11)
int vxo_collider_info(vxo_system* s, uint32_t collider, int32_t bounds[6], float aabb[6],
                      uint32_t* solid) {
  if (collider >= s->shared.colliders.size()) return VXP_ERR_ARGUMENT;
  const Collider& c = *s->shared.colliders[collider];
  bounds[0] = c.min_bound.x; bounds[1] = c.min_bound.y; bounds[2] = c.min_bound.z;
  bounds[3] = c.max_bound.x; bounds[4] = c.max_bound.y; bounds[5] = c.max_bound.z;
  aabb[0] = c.local_aabb.min.x; aabb[1] = c.local_aabb.min.y; aabb[2] = c.local_aabb.min.z;
  aabb[3] = c.local_aabb.max.x; aabb[4] = c.local_aabb.max.y; aabb[5] = c.local_aabb.max.z;
  *solid = c.solid_count;
  return VXP_OK;
}
// One detectVoxelCollisions call between two posed colliders (unit tests of
// the narrow phase): poses are pos[3]+quat[4]; returns contact count.
int vxo_detect_pair(vxo_system* s, uint32_t ca, const float xa[7], uint32_t cb, const float xb[7],
                    vxp_manifold* out, uint32_t* voxels_checked, uint32_t* candidates) {
  Body A, B;
  A.id = 1;
  B.id = 2;
  A.pos = {xa[0], xa[1], xa[2]};
  A.rot = {xa[3], xa[4], xa[5], xa[6]};
  B.pos = {xb[0], xb[1], xb[2]};
  B.rot = {xb[3], xb[4], xb[5], xb[6]};
  A.friction = B.friction = 0.6f;
  A.restitution = B.restitution = 0.2f;
  DetectResult r = detectVoxelCollisions(A, *s->shared.colliders[ca], B, *s->shared.colliders[cb]);
  std::memset(out, 0, sizeof(*out));
  out->body_a_id = 1;
  out->body_b_id = 2;
  out->friction = r.m.friction;
  out->restitution = r.m.restitution;
  out->contact_count = (uint32_t)r.m.n;
  for (int i = 0; i < r.m.n; i++) exportContact(r.m.c[i], out->contacts[i]);
  *voxels_checked = r.voxels_checked;
  *candidates = r.candidates;
  return r.m.n;
}

}  // extern "C"
