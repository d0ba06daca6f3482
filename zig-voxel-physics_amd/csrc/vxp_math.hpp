// vxp_math.hpp — FP32 vector/quaternion math of the step pipeline, usable on
// the host and on gfx950.  Each function keeps the evaluation order of
// zig-voxel-physics src/math.zig (cited per function); the library is built
// with -ffp-contract=off so no multiply-add is fused, as in Zig's strict float
// mode.  sin/cos of an f32 argument are evaluated in double precision and
// rounded to f32.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define VXD __host__ __device__ __forceinline__

namespace vxm {

constexpr float kEps = 1e-6f;                          // src/math.zig:4
constexpr float kPi = 3.14159265358979323846f;         // src/math.zig:3

VXD float fmin_(float a, float b) { return fminf(a, b); }
VXD float fmax_(float a, float b) { return fmaxf(a, b); }
VXD float clampz(float x, float lo, float hi) { return fmaxf(lo, fminf(x, hi)); }
VXD float sin32(float x) { return (float)sin((double)x); }
VXD float cos32(float x) { return (float)cos((double)x); }

struct V3 {
  float x, y, z;
};
struct Q4 {
  float x, y, z, w;
};
struct M3 {
  float m[9];  // row major
};
struct Box {
  V3 lo, hi;
};

VXD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
VXD V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
VXD V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
VXD V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
VXD V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
VXD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VXD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
VXD float len(V3 a) { return sqrtf(dot(a, a)); }
// src/math.zig:101-105
VXD V3 normalize(V3 a) {
  float l = len(a);
  if (l < kEps) return V3{0.f, 0.f, 0.f};
  return scale(a, 1.0f / l);
}
// src/math.zig:107-111
VXD V3 normalizeOrZero(V3 a) {
  float ls = dot(a, a);
  if (ls < kEps * kEps) return V3{0.f, 0.f, 0.f};
  return scale(a, 1.0f / sqrtf(ls));
}
VXD V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
VXD V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
VXD V3 vclamp(V3 a, V3 lo, V3 hi) {
  return V3{clampz(a.x, lo.x, hi.x), clampz(a.y, lo.y, hi.y), clampz(a.z, lo.z, hi.z)};
}
VXD bool finite3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// src/math.zig:298-308
VXD Q4 qnormalize(Q4 q) {
  float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  if (l < kEps) return Q4{0.f, 0.f, 0.f, 1.f};
  float inv = 1.0f / l;
  return Q4{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
// src/math.zig:267-275
VXD V3 qrotate(Q4 q, V3 v) {
  float ls = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (ls < 0.99f || ls > 1.01f) q = qnormalize(q);
  V3 qv{q.x, q.y, q.z};
  V3 uv = cross(qv, v);
  V3 uuv = cross(qv, uv);
  return add(add(v, scale(uv, 2.0f * q.w)), scale(uuv, 2.0f));
}
VXD Q4 qconj(Q4 q) { return Q4{-q.x, -q.y, -q.z, q.w}; }
// src/math.zig:258-265
VXD Q4 qmul(Q4 a, Q4 o) {
  return Q4{a.w * o.x + a.x * o.w + a.y * o.z - a.z * o.y, a.w * o.y - a.x * o.z + a.y * o.w + a.z * o.x,
            a.w * o.z + a.x * o.y - a.y * o.x + a.z * o.w, a.w * o.w - a.x * o.x - a.y * o.y - a.z * o.z};
}
// src/math.zig:230-240
VXD Q4 qFromAxisAngle(V3 axis, float angle) {
  float h = angle * 0.5f;
  float s = sin32(h);
  V3 na = normalize(axis);
  return Q4{na.x * s, na.y * s, na.z * s, cos32(h)};
}
// src/math.zig:755-768
VXD Q4 integrateRotation(Q4 q, V3 w, float dt) {
  V3 om = scale(w, dt);
  float l = len(om);
  if (l < kEps) return q;
  float half = l * 0.5f;
  V3 axis = scale(om, 1.0f / l);
  Q4 d = qFromAxisAngle(axis, half * 2.0f);
  return qnormalize(qmul(d, q));
}
// src/solver.zig:984-1005
VXD Q4 integrateRotationSmall(Q4 r, V3 d) {
  Q4 hw{d.x * 0.5f, d.y * 0.5f, d.z * 0.5f, 0.0f};
  Q4 dq{hw.w * r.x + hw.x * r.w + hw.y * r.z - hw.z * r.y, hw.w * r.y - hw.x * r.z + hw.y * r.w + hw.z * r.x,
        hw.w * r.z + hw.x * r.y - hw.y * r.x + hw.z * r.w, hw.w * r.w - hw.x * r.x - hw.y * r.y - hw.z * r.z};
  return qnormalize(Q4{r.x + dq.x, r.y + dq.y, r.z + dq.z, r.w + dq.w});
}

// src/math.zig:350-372
VXD M3 qToMat3(Q4 q) {
  float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  return M3{{1.0f - (yy + zz), xy + wz, xz - wy, xy - wz, 1.0f - (xx + zz), yz + wx, xz + wy, yz - wx,
             1.0f - (xx + yy)}};
}
// src/math.zig:446-452
VXD V3 mulVec(const M3& a, V3 v) {
  return V3{a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
// src/math.zig:454-464
VXD M3 mmul(const M3& a, const M3& b) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r.m[3 * i + j] = a.m[3 * i] * b.m[j] + a.m[3 * i + 1] * b.m[3 + j] + a.m[3 * i + 2] * b.m[6 + j];
  return r;
}
VXD M3 mtranspose(const M3& a) {
  return M3{{a.m[0], a.m[3], a.m[6], a.m[1], a.m[4], a.m[7], a.m[2], a.m[5], a.m[8]}};
}
// src/body.zig:386-392 and src/solver.zig:351-354: (R * I) * R^T
VXD M3 worldInvInertia(Q4 rot, const M3& iil) {
  M3 R = qToMat3(rot);
  return mmul(mmul(R, iil), mtranspose(R));
}
// src/math.zig:482-507
VXD M3 minverse(const M3& a) {
  const float* m = a.m;
  float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
              m[2] * (m[3] * m[7] - m[4] * m[6]);
  if (fabsf(det) < kEps) return M3{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
  float id = 1.0f / det;
  return M3{{(m[4] * m[8] - m[5] * m[7]) * id, (m[2] * m[7] - m[1] * m[8]) * id,
             (m[1] * m[5] - m[2] * m[4]) * id, (m[5] * m[6] - m[3] * m[8]) * id,
             (m[0] * m[8] - m[2] * m[6]) * id, (m[2] * m[3] - m[0] * m[5]) * id,
             (m[3] * m[7] - m[4] * m[6]) * id, (m[1] * m[6] - m[0] * m[7]) * id,
             (m[0] * m[4] - m[1] * m[3]) * id}};
}

// Transform.transformPoint (src/math.zig:543-546): normalize, then rotate
VXD V3 xformPoint(V3 pos, Q4 rot, V3 p) { return add(qrotate(qnormalize(rot), p), pos); }
// Transform.inverseTransformPoint (src/math.zig:552-554): conjugate, no normalize
VXD V3 invXformPoint(V3 pos, Q4 rot, V3 p) { return qrotate(qconj(rot), sub(p, pos)); }

// AABB.transform (src/math.zig:686-693) over getVertices order (src/math.zig:671-683)
VXD Box boxTransform(Box b, V3 pos, Q4 rot) {
  Box r{V3{INFINITY, INFINITY, INFINITY}, V3{-INFINITY, -INFINITY, -INFINITY}};
  for (int i = 0; i < 8; i++) {
    V3 c{(i & 1) ? b.hi.x : b.lo.x, (i & 2) ? b.hi.y : b.lo.y, (i & 4) ? b.hi.z : b.lo.z};
    V3 p = xformPoint(pos, rot, c);
    r.lo = vmin(r.lo, p);
    r.hi = vmax(r.hi, p);
  }
  return r;
}
VXD Box boxInflate(Box b, float m) {
  V3 mm{m, m, m};
  return Box{sub(b.lo, mm), add(b.hi, mm)};
}
VXD bool boxContains(Box a, Box o) {
  return o.lo.x >= a.lo.x && o.hi.x <= a.hi.x && o.lo.y >= a.lo.y && o.hi.y <= a.hi.y &&
         o.lo.z >= a.lo.z && o.hi.z <= a.hi.z;
}
VXD bool boxOverlaps(Box a, Box o) {
  return a.lo.x <= o.hi.x && a.hi.x >= o.lo.x && a.lo.y <= o.hi.y && a.hi.y >= o.lo.y &&
         a.lo.z <= o.hi.z && a.hi.z >= o.lo.z;
}

// OBB of a collider: OBB.fromAABB (src/math.zig:715-721) + getAxes/getVertices
struct Obb {
  V3 c, h;
  Q4 r;
};
VXD Obb obbFromBox(Box b, V3 pos, Q4 rot) {
  V3 center = scale(add(b.lo, b.hi), 0.5f);
  V3 ext = scale(sub(b.hi, b.lo), 0.5f);
  return Obb{xformPoint(pos, rot, center), ext, rot};
}
VXD void obbAxes(const Obb& o, V3 ax[3]) {
  ax[0] = qrotate(o.r, V3{1.f, 0.f, 0.f});
  ax[1] = qrotate(o.r, V3{0.f, 1.f, 0.f});
  ax[2] = qrotate(o.r, V3{0.f, 0.f, 1.f});
}
// src/math.zig:730-748
VXD void obbVertices(const Obb& o, V3 v[8]) {
  V3 ax[3];
  obbAxes(o, ax);
  V3 ex = scale(ax[0], o.h.x), ey = scale(ax[1], o.h.y), ez = scale(ax[2], o.h.z);
  V3 c = o.c;
  v[0] = sub(sub(sub(c, ex), ey), ez);
  v[1] = sub(sub(add(c, ex), ey), ez);
  v[2] = sub(add(sub(c, ex), ey), ez);
  v[3] = sub(add(add(c, ex), ey), ez);
  v[4] = add(sub(sub(c, ex), ey), ez);
  v[5] = add(sub(add(c, ex), ey), ez);
  v[6] = add(add(sub(c, ex), ey), ez);
  v[7] = add(add(add(c, ex), ey), ez);
}

// src/collision_voxel.zig:47-83: normalized LUT vectors
VXD V3 lutNormal(uint32_t idx) {
  const signed char L[26][3] = {
      {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},  {0, 0, 1},   {0, 0, -1},  {1, 1, 0},
      {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},  {-1, 0, -1},
      {0, 1, 1},   {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {1, 1, 1},   {1, 1, -1},  {1, -1, 1},
      {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  if (idx >= 26) return V3{0.f, 0.f, 0.f};
  return normalize(V3{(float)L[idx][0], (float)L[idx][1], (float)L[idx][2]});
}

// computeNormalIndex (src/collision_voxel.zig:103-130) of a 6-bit empty-side
// mask (bits -x, +x, -y, +y, -z, +z): the table entry equal to the summed
// directions of the empty sides.  Every nonzero sum of unit steps is in the
// table; a zero sum (opposite sides empty) goes to findClosestNormalIndex
// (src/collision_voxel.zig:85-99), where every dot product is zero: entry 0.
VXD uint32_t normalIndexOf(uint32_t empty) {
  const signed char L[26][3] = {
      {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},  {0, 0, 1},   {0, 0, -1},  {1, 1, 0},
      {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},  {-1, 0, -1},
      {0, 1, 1},   {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {1, 1, 1},   {1, 1, -1},  {1, -1, 1},
      {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  const int d[3] = {(int)((empty >> 1) & 1u) - (int)(empty & 1u),
                    (int)((empty >> 3) & 1u) - (int)((empty >> 2) & 1u),
                    (int)((empty >> 5) & 1u) - (int)((empty >> 4) & 1u)};
  for (uint32_t i = 0; i < 26; i++)
    if (L[i][0] == d[0] && L[i][1] == d[1] && L[i][2] == d[2]) return i;
  return 0;
}
// classifyFromNeighbors -> classifyVoxel (src/collision_voxel.zig:194-235):
// the voxel byte of a solid voxel, type by the number of empty sides.
VXD uint8_t classifyByte(uint32_t empty) {
  const uint32_t c = (uint32_t)__builtin_popcount(empty & 63u);
  const uint32_t t = c == 0 ? 4u : c == 1 ? 3u : c == 2 ? 2u : 1u;  // inside, face, edge, corner
  return (uint8_t)(t | (normalIndexOf(empty & 63u) << 3));
}

// src/collider.zig:187-192
VXD uint64_t hashBrickPos(int32_t x, int32_t y, int32_t z) {
  return (uint64_t)(int64_t)x ^ ((uint64_t)(int64_t)y * 0x517cc1b727220a95ull) ^
         ((uint64_t)(int64_t)z * 0x9e3779b97f4a7c15ull);
}
VXD int32_t floorDiv8(int32_t a) { return a >= 0 ? a / 8 : -((-a + 7) / 8); }

// Solver colouring priority: a 64-bit finalizer mix of the pair key (a<<32|b).
VXD uint32_t colorPrio(uint64_t key) {
  uint64_t h = key;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return (uint32_t)(h >> 32);
}

}  // namespace vxm
