// ora_math.hpp — ORACLE (test infrastructure only; never linked into the product).
//
// CPU restatement of zig-voxel-physics src/math.zig and src/collision_voxel.zig.
// Every operation keeps the reference's evaluation order; the file is compiled
// with -ffp-contract=off so no multiply-add is fused (Zig's default strict
// float mode has no contraction).  Zig @min/@max on floats are fminf/fmaxf,
// std.math.clamp(x, lo, hi) is @max(lo, @min(x, hi)).
//
// sin/cos: Zig's @sin/@cos on f32 are evaluated here as (float)sin((double)x)
// (correctly rounded except in double-rounding corner cases, like the musl
// sinf/cosf that Zig's compiler_rt follows).  The GPU pipeline uses the same
// rule, so the two agree bit for bit.
#pragma once
#include <cmath>
#include <cstdint>

namespace ora {

constexpr float kEps = 1e-6f;          // src/math.zig:4
constexpr float kPi = 3.14159265358979323846f;  // src/math.zig:3 (std.math.pi as f32)

inline float zmin(float a, float b) { return std::fmin(a, b); }
inline float zmax(float a, float b) { return std::fmax(a, b); }
inline float zclamp(float x, float lo, float hi) { return zmax(lo, zmin(x, hi)); }
inline float zsin(float x) { return (float)std::sin((double)x); }
inline float zcos(float x) { return (float)std::cos((double)x); }

struct Vec3 {  // src/math.zig:6-176
  float x, y, z;
  static Vec3 make(float a, float b, float c) { return Vec3{a, b, c}; }
  static Vec3 splat(float v) { return Vec3{v, v, v}; }
  Vec3 add(Vec3 o) const { return {x + o.x, y + o.y, z + o.z}; }
  Vec3 sub(Vec3 o) const { return {x - o.x, y - o.y, z - o.z}; }
  Vec3 scale(float s) const { return {x * s, y * s, z * s}; }
  Vec3 negate() const { return {-x, -y, -z}; }
  float dot(Vec3 o) const { return x * o.x + y * o.y + z * o.z; }
  Vec3 cross(Vec3 o) const {
    return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
  }
  float lengthSquared() const { return dot(*this); }
  float length() const { return std::sqrt(lengthSquared()); }
  // src/math.zig:101-105
  Vec3 normalize() const {
    float len = length();
    if (len < kEps) return {0, 0, 0};
    return scale(1.0f / len);
  }
  // src/math.zig:107-111
  Vec3 normalizeOrZero() const {
    float ls = lengthSquared();
    if (ls < kEps * kEps) return {0, 0, 0};
    return scale(1.0f / std::sqrt(ls));
  }
  Vec3 minComponents(Vec3 o) const { return {zmin(x, o.x), zmin(y, o.y), zmin(z, o.z)}; }
  Vec3 maxComponents(Vec3 o) const { return {zmax(x, o.x), zmax(y, o.y), zmax(z, o.z)}; }
  Vec3 clamp(Vec3 lo, Vec3 hi) const {
    return {zclamp(x, lo.x, hi.x), zclamp(y, lo.y, hi.y), zclamp(z, lo.z, hi.z)};
  }
  bool isFinite() const { return std::isfinite(x) && std::isfinite(y) && std::isfinite(z); }
};

struct IVec3 {  // src/math.zig:178-216
  int32_t x, y, z;
  Vec3 toVec3() const { return {(float)x, (float)y, (float)z}; }
  bool eql(IVec3 o) const { return x == o.x && y == o.y && z == o.z; }
};

struct Mat3;

struct Quat {  // src/math.zig:218-385
  float x, y, z, w;
  static Quat identity() { return {0, 0, 0, 1}; }
  // src/math.zig:230-240
  static Quat fromAxisAngle(Vec3 axis, float angle) {
    float half = angle * 0.5f;
    float s = zsin(half);
    Vec3 na = axis.normalize();
    return {na.x * s, na.y * s, na.z * s, zcos(half)};
  }
  // src/math.zig:258-265
  Quat mul(Quat o) const {
    return {w * o.x + x * o.w + y * o.z - z * o.y, w * o.y - x * o.z + y * o.w + z * o.x,
            w * o.z + x * o.y - y * o.x + z * o.w, w * o.w - x * o.x - y * o.y - z * o.z};
  }
  // src/math.zig:298-308
  Quat normalize() const {
    float len = std::sqrt(x * x + y * y + z * z + w * w);
    if (len < kEps) return identity();
    float inv = 1.0f / len;
    return {x * inv, y * inv, z * inv, w * inv};
  }
  // src/math.zig:267-275: re-normalizes only outside [0.99, 1.01]
  Vec3 rotate(Vec3 v) const {
    float ls = x * x + y * y + z * z + w * w;
    Quat q = (ls < 0.99f || ls > 1.01f) ? normalize() : *this;
    Vec3 qv{q.x, q.y, q.z};
    Vec3 uv = qv.cross(v);
    Vec3 uuv = qv.cross(uv);
    return v.add(uv.scale(2.0f * q.w)).add(uuv.scale(2.0f));
  }
  Quat conjugate() const { return {-x, -y, -z, w}; }
  Mat3 toMat3() const;
};

struct Mat3 {  // src/math.zig:387-528
  float m[3][3];
  static Mat3 identity() { return {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}}; }
  static Mat3 zero() { return {{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}}; }
  Vec3 mulVec(Vec3 v) const {
    return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
            m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
            m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
  }
  Mat3 mul(const Mat3& o) const {
    Mat3 r;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        r.m[i][j] = m[i][0] * o.m[0][j] + m[i][1] * o.m[1][j] + m[i][2] * o.m[2][j];
    return r;
  }
  Mat3 transpose() const {
    return {{{m[0][0], m[1][0], m[2][0]}, {m[0][1], m[1][1], m[2][1]}, {m[0][2], m[1][2], m[2][2]}}};
  }
  float determinant() const {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
           m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  }
  // src/math.zig:482-507: identity when |det| < 1e-6
  Mat3 inverse() const {
    float det = determinant();
    if (std::fabs(det) < kEps) return identity();
    float id = 1.0f / det;
    Mat3 r;
    r.m[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    r.m[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    r.m[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
    r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    return r;
  }
};

// src/math.zig:350-372
inline Mat3 Quat::toMat3() const {
  float x2 = x + x, y2 = y + y, z2 = z + z;
  float xx = x * x2, xy = x * y2, xz = x * z2;
  float yy = y * y2, yz = y * z2, zz = z * z2;
  float wx = w * x2, wy = w * y2, wz = w * z2;
  return {{{1.0f - (yy + zz), xy + wz, xz - wy},
           {xy - wz, 1.0f - (xx + zz), yz + wx},
           {xz + wy, yz - wx, 1.0f - (xx + yy)}}};
}

struct Transform {  // src/math.zig:530-582
  Vec3 position;
  Quat rotation;
  // src/math.zig:543-546: normalize, then rotate
  Vec3 transformPoint(Vec3 p) const { return rotation.normalize().rotate(p).add(position); }
  // src/math.zig:552-554: conjugate without normalizing
  Vec3 inverseTransformPoint(Vec3 p) const { return rotation.conjugate().rotate(p.sub(position)); }
};

struct AABB {  // src/math.zig:583-700
  Vec3 min, max;
  static AABB empty() {
    return {Vec3::splat(INFINITY), Vec3::splat(-INFINITY)};
  }
  Vec3 center() const { return min.add(max).scale(0.5f); }
  Vec3 extents() const { return max.sub(min).scale(0.5f); }
  AABB expand(Vec3 p) const { return {min.minComponents(p), max.maxComponents(p)}; }
  AABB merge(const AABB& o) const { return {min.minComponents(o.min), max.maxComponents(o.max)}; }
  AABB inflate(float m) const {
    Vec3 mm = Vec3::splat(m);
    return {min.sub(mm), max.add(mm)};
  }
  // src/math.zig:655-659 (inclusive)
  bool overlaps(const AABB& o) const {
    return min.x <= o.max.x && max.x >= o.min.x && min.y <= o.max.y && max.y >= o.min.y &&
           min.z <= o.max.z && max.z >= o.min.z;
  }
  bool containsAABB(const AABB& o) const {
    return o.min.x >= min.x && o.max.x <= max.x && o.min.y >= min.y && o.max.y <= max.y &&
           o.min.z >= min.z && o.max.z <= max.z;
  }
  void getVertices(Vec3 v[8]) const {
    v[0] = {min.x, min.y, min.z};
    v[1] = {max.x, min.y, min.z};
    v[2] = {min.x, max.y, min.z};
    v[3] = {max.x, max.y, min.z};
    v[4] = {min.x, min.y, max.z};
    v[5] = {max.x, min.y, max.z};
    v[6] = {min.x, max.y, max.z};
    v[7] = {max.x, max.y, max.z};
  }
  // src/math.zig:686-693
  AABB transform(const Transform& t) const {
    Vec3 c[8];
    getVertices(c);
    AABB r = empty();
    for (int i = 0; i < 8; i++) r = r.expand(t.transformPoint(c[i]));
    return r;
  }
};

struct OBB {  // src/math.zig:702-753
  Vec3 center, half_extents;
  Quat rotation;
  static OBB fromAABB(const AABB& a, const Transform& t) {
    return {t.transformPoint(a.center()), a.extents(), t.rotation};
  }
  void getAxes(Vec3 ax[3]) const {
    ax[0] = rotation.rotate({1, 0, 0});
    ax[1] = rotation.rotate({0, 1, 0});
    ax[2] = rotation.rotate({0, 0, 1});
  }
  void getVertices(Vec3 v[8]) const {
    Vec3 ax[3];
    getAxes(ax);
    Vec3 ex = ax[0].scale(half_extents.x), ey = ax[1].scale(half_extents.y),
         ez = ax[2].scale(half_extents.z);
    v[0] = center.sub(ex).sub(ey).sub(ez);
    v[1] = center.add(ex).sub(ey).sub(ez);
    v[2] = center.sub(ex).add(ey).sub(ez);
    v[3] = center.add(ex).add(ey).sub(ez);
    v[4] = center.sub(ex).sub(ey).add(ez);
    v[5] = center.add(ex).sub(ey).add(ez);
    v[6] = center.sub(ex).add(ey).add(ez);
    v[7] = center.add(ex).add(ey).add(ez);
  }
};

// src/math.zig:755-768
inline Quat integrateRotation(Quat q, Vec3 w, float dt) {
  Vec3 omega = w.scale(dt);
  float len = omega.length();
  if (len < kEps) return q;
  float half = len * 0.5f;
  Vec3 axis = omega.scale(1.0f / len);
  Quat d = Quat::fromAxisAngle(axis, half * 2.0f);
  return d.mul(q).normalize();
}

// ---- src/collision_voxel.zig ------------------------------------------------
enum VoxelType : uint8_t { kEmpty = 0, kCorner = 1, kEdge = 2, kFace = 3, kInside = 4 };

// src/collision_voxel.zig:47-76
constexpr int8_t kNormalLut[26][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1},  {1, 1, 0},
    {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1},  {1, 0, -1},  {-1, 0, 1},  {-1, 0, -1},
    {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {1, 1, 1},  {1, 1, -1},  {1, -1, 1},
    {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

inline uint8_t voxelType(uint8_t v) { return v & 7u; }
inline uint8_t voxelNormal(uint8_t v) { return v >> 3; }

// src/collision_voxel.zig:79-83
inline Vec3 getNormalVec3(uint32_t idx) {
  if (idx >= 26) return {0, 0, 0};
  IVec3 n{kNormalLut[idx][0], kNormalLut[idx][1], kNormalLut[idx][2]};
  return n.toVec3().normalize();
}

// src/collision_voxel.zig:86-100
inline uint32_t findClosestNormalIndex(Vec3 dir) {
  uint32_t best = 0;
  float best_dot = -INFINITY;
  for (uint32_t i = 0; i < 26; i++) {
    Vec3 n = getNormalVec3(i);
    float d = dir.dot(n);
    if (d > best_dot) {
      best_dot = d;
      best = i;
    }
  }
  return best;
}

// src/collision_voxel.zig:103-130
inline uint32_t computeNormalIndex(uint32_t empty) {
  if (__builtin_popcount(empty & 63u) == 0) return 0;
  int dx = 0, dy = 0, dz = 0;
  if (empty & 1u) dx -= 1;
  if (empty & 2u) dx += 1;
  if (empty & 4u) dy -= 1;
  if (empty & 8u) dy += 1;
  if (empty & 16u) dz -= 1;
  if (empty & 32u) dz += 1;
  for (uint32_t i = 0; i < 26; i++)
    if (kNormalLut[i][0] == dx && kNormalLut[i][1] == dy && kNormalLut[i][2] == dz) return i;
  return findClosestNormalIndex(IVec3{dx, dy, dz}.toVec3());
}

// src/collision_voxel.zig:194-210
inline uint8_t classifyVoxel(bool solid, uint32_t empty) {
  if (!solid) return 0;
  int c = __builtin_popcount(empty & 63u);
  uint8_t t = c == 0 ? kInside : c == 1 ? kFace : c == 2 ? kEdge : kCorner;
  return (uint8_t)(t | (computeNormalIndex(empty) << 3));
}

// src/narrow_phase.zig:701-709 (same table as CollisionVoxel.shouldCollide)
inline bool shouldCheckCollision(uint8_t ta, uint8_t tb) {
  if (ta == kEmpty || tb == kEmpty) return false;
  if (ta == kCorner || tb == kCorner) return true;
  if (ta == kEdge && tb == kEdge) return true;
  return false;
}

}  // namespace ora
