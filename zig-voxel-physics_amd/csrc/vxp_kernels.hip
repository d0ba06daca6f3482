// vxp_kernels.hip — gfx950 kernels of the step pipeline of zig-voxel-physics.
//
// One step (src/system.zig:537-555) runs as:
//   k_step_begin      allDynamicBodiesSleeping per world (src/system.zig:538-567)
//   k_fat_aabb        broadPhase fat-AABB rule (src/system.zig:570-583, src/broad_phase.zig:175-212)
//   k_pairs_*         findOverlappingPairs pair set (src/broad_phase.zig:488-526)
//   k_pair_process    narrowPhase filter + processPair up to detection (src/system.zig:586-655,
//                     src/narrow_phase.zig:1414-1449, 644-698)
//   k_cache_replay    BrickPairCache get/put in pair order (src/narrow_phase.zig:228-296,1451-1513)
//   k_probe           detectVoxelCollisions/checkNeighborhood/generateContact (src/narrow_phase.zig:722-922)
//   k_reduce          ContactManifold.addContact replay + warm start (src/narrow_phase.zig:451-481,1049-1059)
//   k_reduce_tail, k_scatter, k_copy_skipped  cache values + counters, ContactManager.manifolds compaction
//   k_islands         buildIslandsWithBodyTypes labels (src/islands.zig:318-441)
//   k_ref_coloring    GraphColoring.build artifact (src/islands.zig:668-736)
//   k_solve           ContactSolver.solve per world, per-colour Gauss-Seidel (src/solver.zig:494-958),
//                     write-back and updateSleepStates (src/system.zig:786-792)
#include <hip/hip_runtime.h>

#include <cstring>
#include "vxp_sort.hpp"

#include "vxp_dev.hpp"
#include "vxp_kernels.hpp"

using namespace vxm;
namespace vxd {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
// The item count of a launch: `n` (the grid's size), or min(*n_dev, n) when a
// device word is given (launches recorded before the count exists: the grid
// is sized for a capacity n).
__device__ __forceinline__ uint32_t devCount(uint32_t n, const uint32_t* n_dev) {
  return n_dev ? min(*n_dev, n) : n;
}
__device__ __forceinline__ V3 f4v(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ Q4 f4q(float4 a) { return Q4{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ float4 v4(V3 a) { return make_float4(a.x, a.y, a.z, 0.0f); }
__device__ __forceinline__ float4 q4(Q4 a) { return make_float4(a.x, a.y, a.z, a.w); }

__device__ __forceinline__ uint8_t voxelAt(const DevColliders& C, const DevCollider& h, int32_t x,
                                           int32_t y, int32_t z) {
  int32_t bx = (x >> 3) - h.org[0], by = (y >> 3) - h.org[1], bz = (z >> 3) - h.org[2];
  if ((uint32_t)bx >= (uint32_t)h.dim[0] || (uint32_t)by >= (uint32_t)h.dim[1] ||
      (uint32_t)bz >= (uint32_t)h.dim[2])
    return 0;
  int32_t b = C.dir[h.dir_off + ((uint32_t)bz * h.dim[1] + by) * h.dim[0] + bx];
  if (b < 0) return 0;
  return C.vox[(size_t)b * 512 + (x & 7) + 8 * (y & 7) + 64 * (z & 7)];
}

// One collider's brick directory geometry, held in registers for a probe loop.
struct VoxMap {
  int32_t org[3], dim[3];
  uint32_t dir_off;
};
__device__ __forceinline__ VoxMap voxMap(const DevCollider& h) {
  VoxMap m;
  for (int k = 0; k < 3; k++) {
    m.org[k] = h.org[k];
    m.dim[k] = h.dim[k];
  }
  m.dir_off = h.dir_off;
  return m;
}
// voxelAt with the directory and brick pool passed as non-aliasing pointers.
__device__ __forceinline__ uint8_t voxelAt(const int32_t* __restrict__ dir,
                                           const uint8_t* __restrict__ vox, const VoxMap& m,
                                           int32_t x, int32_t y, int32_t z) {
  int32_t bx = (x >> 3) - m.org[0], by = (y >> 3) - m.org[1], bz = (z >> 3) - m.org[2];
  if ((uint32_t)bx >= (uint32_t)m.dim[0] || (uint32_t)by >= (uint32_t)m.dim[1] ||
      (uint32_t)bz >= (uint32_t)m.dim[2])
    return 0;
  int32_t b = dir[m.dir_off + ((uint32_t)bz * m.dim[1] + by) * m.dim[0] + bx];
  if (b < 0) return 0;
  return vox[(size_t)b * 512 + (x & 7) + 8 * (y & 7) + 64 * (z & 7)];
}

__device__ __forceinline__ Box colliderBox(const DevCollider& h) {
  return Box{V3{h.lo[0], h.lo[1], h.lo[2]}, V3{h.hi[0], h.hi[1], h.hi[2]}};
}

// src/collider.zig:213-220
__device__ __forceinline__ V3 voxelToLocal(float vs, int32_t x, int32_t y, int32_t z) {
  float half = vs * 0.5f;
  return V3{(float)x * vs + half, (float)y * vs + half, (float)z * vs + half};
}
// src/collider.zig:223-229
__device__ __forceinline__ void localToVoxel(float vs, V3 p, int32_t o[3]) {
  o[0] = (int32_t)floorf(p.x / vs);
  o[1] = (int32_t)floorf(p.y / vs);
  o[2] = (int32_t)floorf(p.z / vs);
}

__device__ __forceinline__ bool shouldCheck(uint32_t ta, uint32_t tb) {
  if (ta == 0 || tb == 0) return false;
  if (ta == VXP_VOXEL_CORNER || tb == VXP_VOXEL_CORNER) return true;
  return ta == VXP_VOXEL_EDGE && tb == VXP_VOXEL_EDGE;
}

// src/narrow_phase.zig:384-392
__device__ __forceinline__ uint64_t featureId(int32_t ax, int32_t ay, int32_t az, int32_t bx,
                                              int32_t by, int32_t bz) {
  return (uint64_t)((int64_t)ax & 0xFFFF) | ((uint64_t)((int64_t)ay & 0xFFFF) << 16) |
         ((uint64_t)((int64_t)az & 0xFFFF) << 32) | ((uint64_t)((int64_t)bx & 0xF) << 48) |
         ((uint64_t)((int64_t)by & 0xF) << 52) | ((uint64_t)((int64_t)bz & 0xF) << 56);
}

__device__ __forceinline__ int waveInclusiveScan(int x, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int t = __shfl_up(x, d, kWave);
    if (lane >= d) x += t;
  }
  return x;
}
__device__ __forceinline__ uint32_t waveSum(uint32_t x) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

__device__ __forceinline__ uint64_t shflXor64(uint64_t v, uint32_t d) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, (int)d, kWave);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), (int)d, kWave);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, kWave);
  uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ float rdlane(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ uint32_t rdlane(uint32_t v, int k) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
}

// First index in the sorted a[0..n) whose value exceeds v.
__device__ __forceinline__ uint32_t upperBound(const uint32_t* a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Exclusive scan of data[0..n) by the whole block (tmp holds blockDim.x
// entries); *total receives the sum when non-null.
__device__ void blockScanExclusive(uint32_t* data, uint32_t n, uint32_t* tmp /*blockDim*/,
                                   uint32_t* total) {
  // per block-sized chunk: inclusive scans within each wave by shuffles, the
  // wave totals through tmp, and a running carry kept in registers (the same
  // in every thread)
  constexpr uint32_t W = kWave;
  const uint32_t tid = threadIdx.x, bs = blockDim.x, lane = tid & (W - 1), wave = tid / W;
  const uint32_t nwv = (bs + W - 1) / W, last = min(W - 1, bs - 1 - wave * W);
  uint32_t carry = 0;
  __syncthreads();  // the data may have just been written by other threads
  for (uint32_t base = 0; base < n; base += bs) {
    const uint32_t i = base + tid;
    const uint32_t v = i < n ? data[i] : 0;
    uint32_t incl = v;
#pragma unroll
    for (uint32_t d = 1; d < W; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)incl, d, kWave);
      if (lane >= d) incl += t;
    }
    if (lane == last) tmp[wave] = incl;
    __syncthreads();
    uint32_t before = 0, chunk = 0;
    for (uint32_t k = 0; k < nwv; k++) {
      const uint32_t x = tmp[k];
      before += k < wave ? x : 0u;
      chunk += x;
    }
    if (i < n) data[i] = carry + before + incl - v;
    carry += chunk;
    __syncthreads();
  }
  if (total && tid == 0) *total = carry;
  __syncthreads();
}

// atomicAdd(addr, v) aggregated over the wave when every lane with v != 0
// targets the same counter.  Call from converged code (all 64 lanes).
__device__ __forceinline__ void aggAdd(uint32_t* addr, uint32_t v) {
  uint64_t act = __ballot(v != 0);
  if (!act) return;
  int leader = __ffsll((long long)act) - 1;
  uint64_t a = (uint64_t)(uintptr_t)addr;
  uint64_t la = shfl64(a, leader);
  if (__ballot(v != 0 && a != la) == 0) {
    uint32_t sum = waveSum(v);
    if ((int)(threadIdx.x & (kWave - 1)) == leader) atomicAdd(addr, sum);
  } else if (v) {
    atomicAdd(addr, v);
  }
}

// Per-world counters summed over the block: thread t adds v[k] to
// base[k][world * stride] (k < N).  Worlds must be non-decreasing in thread
// order among contributing threads (pairs are sorted by world).  One atomic
// per run of equal worlds per counter; every thread of the block must call.
template <int N>
__device__ void blockAggAdd(uint32_t world, const uint32_t (&v)[N], uint32_t* const (&base)[N],
                            uint32_t stride) {
  __shared__ uint32_t s_w[32], s_v[32][N];
  const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  bool any = false;
#pragma unroll
  for (int k = 0; k < N; k++) any |= v[k] != 0;
  const uint64_t act = __ballot(any);
  uint32_t lw = kNone;
  if (act) {
    const int leader = __ffsll((long long)act) - 1;
    lw = rdlane(world, leader);
    if (__ballot(any && world != lw) == 0) {
#pragma unroll
      for (int k = 0; k < N; k++) {
        uint32_t sum = waveSum(v[k]);
        if (lane == 0) s_v[wave][k] = sum;
      }
    } else {
      if (any)
#pragma unroll
        for (int k = 0; k < N; k++)
          if (v[k]) atomicAdd(base[k] + (size_t)world * stride, v[k]);
      lw = kNone;
    }
  }
  if (lane == 0) s_w[wave] = lw;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t nwv = (blockDim.x + kWave - 1) / kWave;
    uint32_t cw = kNone, acc[N];
    for (uint32_t q = 0; q <= nwv; q++) {
      uint32_t qw = q < nwv ? s_w[q] : kNone;
      if (q < nwv && qw == kNone) continue;
      if (qw != cw || q == nwv) {
        if (cw != kNone)
#pragma unroll
          for (int k = 0; k < N; k++)
            if (acc[k]) atomicAdd(base[k] + (size_t)cw * stride, acc[k]);
        cw = qw;
#pragma unroll
        for (int k = 0; k < N; k++) acc[k] = 0;
      }
      if (q < nwv)
#pragma unroll
        for (int k = 0; k < N; k++) acc[k] += s_v[q][k];
    }
  }
}

// ---------------------------------------------------------------------------
// step begin: allDynamicBodiesSleeping, per-step stat reset
// ---------------------------------------------------------------------------
__global__ void k_step_begin(DevCfg cfg, DevBodies B, WorldState* ws, vxp_stats* st,
                             uint32_t n_worlds, uint32_t* step_ctr) {
  uint32_t w = blockIdx.x;
  if (w >= n_worlds) return;
  if (w == 0 && threadIdx.x < kStepCounters) step_ctr[threadIdx.x] = 0;
  __shared__ uint32_t awake;
  if (threadIdx.x == 0) awake = 0;
  __syncthreads();
  uint32_t g0 = w * cfg.max_bodies, mine = 0;
  for (uint32_t s = threadIdx.x; s < cfg.max_bodies; s += blockDim.x) {
    uint32_t g = g0 + s;
    if (B.active[g] && B.type[g] == VXP_BODY_DYNAMIC && !B.sleeping[g]) mine++;
  }
  if (mine) atomicAdd(&awake, mine);
  __syncthreads();
  if (threadIdx.x == 0) {
    ws[w].awake = awake;
    ws[w].skip = (cfg.enable_sleeping && awake == 0) ? 1u : 0u;
    ws[w].pair_lo = ws[w].pair_hi = 0;
    ws[w].produced = 0;
    // ContactManager.beginFrame: the current segment becomes the previous one
    ws[w].prev_start = ws[w].man_start;
    ws[w].prev_count = ws[w].man_count;
    if (!ws[w].skip) {
      vxp_stats& s = st[w];
      s.steps_executed++;
      s.awake_bodies_stepped += awake;
      s.broad_phase_pairs = s.narrow_phase_tested = s.narrow_phase_colliding = 0;
      s.contacts_generated = s.voxels_checked = s.pairs_skipped_sleeping = 0;
      s.cache_hits = s.cache_misses = s.manifolds_skipped_sleeping = 0;
      s.contact_candidates = 0;
      s.manifolds = s.solver_manifolds = s.colors = s.position_iterations_run = 0;
      s.max_penetration = 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// broad phase: fat AABB rule (src/broad_phase.zig:175-212) for awake dynamic bodies
// ---------------------------------------------------------------------------
__device__ Box worldBoxOf(const DevBodies& B, const DevColliders& C, uint32_t g) {
  V3 p = f4v(B.pos[g]);
  int32_t c = B.collider[g];
  if (c < 0) return Box{p, p};
  return boxTransform(colliderBox(C.hdr[c]), p, f4q(B.rot[g]));
}

// moved (tree order only): 1 where the fat box changed, the bodies the tree
// re-inserts (BroadPhase.updateBody returning true).
__global__ void k_fat_aabb(DevCfg cfg, DevBodies B, DevColliders C, const WorldState* ws,
                           uint32_t n_slots, uint8_t* moved) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_slots) return;
  if (moved) moved[g] = 0;
  if (ws[g / cfg.max_bodies].skip) return;
  if (!B.active[g] || B.type[g] != VXP_BODY_DYNAMIC || B.sleeping[g]) return;
  Box aabb = worldBoxOf(B, C, g);
  Box fat = boxInflate(aabb, 0.1f);
  Box old{f4v(B.fat_lo[g]), f4v(B.fat_hi[g])};
  if (boxContains(old, aabb)) {
    Box large = boxInflate(fat, 0.1f * 4.0f);
    if (boxContains(large, old)) return;
  }
  V3 d = scale(f4v(B.lv[g]), 2.0f);
  Box e = fat;
  if (d.x < 0) e.lo.x += d.x; else e.hi.x += d.x;
  if (d.y < 0) e.lo.y += d.y; else e.hi.y += d.y;
  if (d.z < 0) e.lo.z += d.z; else e.hi.z += d.z;
  B.fat_lo[g] = v4(e.lo);
  B.fat_hi[g] = v4(e.hi);
  if (moved) moved[g] = 1;
}

// ---------------------------------------------------------------------------
// pair set: every inclusive overlap of fat AABBs within a world, except
// static-static; key (world, min id, max id).  Sort-and-sweep: bodies are
// ordered by (world, fat min.x) and each body scans forward while the next
// body's min.x <= its max.x.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t orderedFloat(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float comp(const float4& v, uint32_t a) {
  return a == 0 ? v.x : (a == 1 ? v.y : v.z);
}

// Sweep axis of each world: the axis with the fewest expected overlaps, i.e.
// the smallest mean fat-box extent relative to the spread of the box centres
// (the pair set does not depend on the axis).  Also the world's active slots
// and the bounds of its doubled box centres (the LBVH's Morton grid).
__global__ void k_world_axis(DevCfg cfg, DevBodies B, WorldState* ws, uint32_t n_worlds) {
  uint32_t w = blockIdx.x;
  if (w >= n_worlds) return;
  __shared__ float lo[3][1024], hi[3][1024], ext[3][1024];
  __shared__ uint32_t n_act;
  if (threadIdx.x == 0) n_act = 0;
  __syncthreads();
  float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
  float e[3] = {0.f, 0.f, 0.f};
  uint32_t na = 0;
  for (uint32_t s = threadIdx.x; s < cfg.max_bodies; s += blockDim.x) {
    uint32_t g = w * cfg.max_bodies + s;
    if (!B.active[g]) continue;
    na++;
    float4 a = B.fat_lo[g], b = B.fat_hi[g];
    float c[3] = {a.x + b.x, a.y + b.y, a.z + b.z};
    float x[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
    for (int k = 0; k < 3; k++) {
      l[k] = fminf(l[k], c[k]);
      h[k] = fmaxf(h[k], c[k]);
      e[k] += fminf(x[k], 1e30f);
    }
  }
  if (na) atomicAdd(&n_act, na);
  for (int k = 0; k < 3; k++) {
    lo[k][threadIdx.x] = l[k];
    hi[k][threadIdx.x] = h[k];
    ext[k][threadIdx.x] = e[k];
  }
  __syncthreads();
  for (uint32_t d = blockDim.x / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d)
      for (int k = 0; k < 3; k++) {
        lo[k][threadIdx.x] = fminf(lo[k][threadIdx.x], lo[k][threadIdx.x + d]);
        hi[k][threadIdx.x] = fmaxf(hi[k][threadIdx.x], hi[k][threadIdx.x + d]);
        ext[k][threadIdx.x] += ext[k][threadIdx.x + d];
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // overlap density ~ 2 * total extent / spread of centres (centres are doubled)
    float r[3];
    for (int k = 0; k < 3; k++) {
      float spread = hi[k][0] - lo[k][0];
      r[k] = spread > 0.0f ? ext[k][0] / spread : INFINITY;
    }
    uint32_t ax = 0;
    if (r[1] < r[ax]) ax = 1;
    if (r[2] < r[ax]) ax = 2;
    ws[w].axis = ax;
    ws[w].np_emit = 0;
    ws[w].n_leaves = n_act;
    for (int k = 0; k < 3; k++) {
      // a spread that is zero or not finite puts every centre in cell 0
      const float spread = hi[k][0] - lo[k][0];
      ws[w].c_lo[k] = lo[k][0];
      ws[w].c_scale[k] = spread > 0.0f && spread < INFINITY ? 1023.0f / spread : 0.0f;
    }
  }
}

// Sweep order of each world in one block: its slots sorted by the sweep
// coordinate of the fat box's lower corner (inactive slots and padding last,
// ties in slot order, as the device-wide sort of (world, coordinate) keys).
template <uint32_t kBS, uint32_t kIPT>
__global__ __launch_bounds__(kBS) void k_sweep_sort(DevCfg cfg, DevBodies B, const WorldState* ws,
                                                    uint32_t* order) {
  __shared__ uint32_t s_k[kBS * kIPT], s_v[kBS * kIPT];
  __shared__ uint32_t s_wcnt[2 * (kBS / 64) * kRadixBins], s_run[2 * kRadixBins];
  const uint32_t w = blockIdx.x, MB = cfg.max_bodies, ax = ws[w].axis;
  for (uint32_t d = threadIdx.x; d < 2 * (kBS / 64) * kRadixBins; d += kBS) s_wcnt[d] = 0;
  uint32_t keys[kIPT], vals[kIPT];
#pragma unroll
  for (uint32_t k = 0; k < kIPT; k++) {  // striped: slot k * kBS + tid
    const uint32_t sl = k * kBS + threadIdx.x;
    const uint32_t g = w * MB + sl;
    keys[k] = 0xFFFFFFFFu;
    vals[k] = g;
    if (sl < MB && B.active[g]) keys[k] = orderedFloat(comp(B.fat_lo[g], ax));
  }
  blockRadixSort<kBS, kIPT, true>(keys, vals, MB, 32, s_k, s_v, s_wcnt, s_run);
#pragma unroll
  for (uint32_t k = 0; k < kIPT; k++) {
    const uint32_t r = k * kBS + threadIdx.x;
    if (r < MB) order[w * MB + r] = vals[k];
  }
}

__global__ void k_sweep_keys(DevCfg cfg, DevBodies B, const WorldState* ws, uint64_t* keys,
                             uint32_t* vals, uint32_t n_slots) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_slots) return;
  uint32_t w = g / cfg.max_bodies;
  // inactive slots sort to the end of their world
  uint32_t k = B.active[g] ? orderedFloat(comp(B.fat_lo[g], ws[w].axis)) : 0xFFFFFFFFu;
  keys[g] = ((uint64_t)w << 32) | k;
  vals[g] = g;
}

// Fat boxes gathered into sweep order, field-major: lo.xyz + id, hi.xyz +
// (world << 2 | active << 1 | static).
__global__ void k_sweep_gather(DevCfg cfg, DevBodies B, const uint32_t* order, uint32_t n_slots,
                               float4* slo, float4* shi) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_slots) return;
  uint32_t g = order[i];
  float4 lo = B.fat_lo[g], hi = B.fat_hi[g];
  uint32_t meta = ((g / cfg.max_bodies) << 2) | (B.active[g] ? 2u : 0u) |
                  (B.type[g] == VXP_BODY_STATIC ? 1u : 0u);
  lo.w = __uint_as_float(B.id[g]);
  hi.w = __uint_as_float(meta);
  slo[i] = lo;
  shi[i] = hi;
}

// Sweep length of body i: its candidates are the bodies after it in sweep
// order, up to the first one that starts past i's upper bound on the sweep
// axis, is inactive or belongs to the next world (a monotone predicate over
// the world's sorted slots [w * max_bodies, (w + 1) * max_bodies)).  send[i]
// = end of the range, wcnt[i] = its 64-candidate work items; wcnt[n] = 0.
__global__ void k_sweep_count(DevCfg cfg, const WorldState* ws, const float4* slo,
                              const float4* shi, uint32_t n_slots, uint32_t* send,
                              uint32_t* wcnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_slots) return;
  if (i == n_slots - 1) wcnt[n_slots] = 0;
  const float4 hi4 = shi[i];
  const uint32_t mi = __float_as_uint(hi4.w);
  uint32_t e = i + 1, c = 0;
  if (mi & 2u) {
    const uint32_t w = mi >> 2, ax = ws[w].axis;
    const float hax = comp(hi4, ax);
    uint32_t lo = i + 1, hi = min(n_slots, (w + 1) * cfg.max_bodies);
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      bool past = !(__float_as_uint(shi[mid].w) & 2u) || comp(slo[mid], ax) > hax;
      if (past) hi = mid; else lo = mid + 1;
    }
    e = lo;
    c = (e - (i + 1) + kWave - 1) / kWave;
  }
  send[i] = e;
  wcnt[i] = c;
}

// Work item = (body i, 64 consecutive candidates).  Wave gw of the fixed grid
// owns the contiguous items [gw * q, (gw + 1) * q), q = ceil(total / waves).
// kCount: wcount[gw] = its overlapping pairs (wcount[waves] = 0).  Otherwise
// the pairs (inclusive overlap, src/broad_phase.zig:488-526; static-static
// excluded) are written from woff[gw] on, in work-item order, as packed keys
// of 32 bits when narrow (the packed width 2 * id_bits + world bits fits) and
// of 64 otherwise.
constexpr uint32_t kSweepWaves = 8192;

template <bool kCount>
__global__ __launch_bounds__(256) void k_sweep_pairs(WorldState* ws, const float4* slo,
                                                     const float4* shi, uint32_t n_slots,
                                                     const uint32_t* send, const uint32_t* woff_i,
                                                     uint32_t* wcount, const uint32_t* woff_w,
                                                     uint64_t* out, uint32_t id_bits, bool narrow,
                                                     bool per_world, uint32_t cap) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t gw = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint32_t total = woff_i[n_slots];
  const uint32_t q = (total + kSweepWaves - 1) / kSweepWaves;
  const uint32_t g0 = min(total, gw * q), g1 = min(total, g0 + q);
  uint32_t n = 0, wn = 0, wcur = kNone;  // wn: pairs of world wcur not yet added to it
  uint32_t o = kCount ? 0 : woff_w[gw];
  if (g0 < g1) {
    uint32_t i = upperBound(woff_i, n_slots, g0) - 1, cur = kNone, end = 0, w = 0, idi = 0;
    bool si = false;
    Box bi;
    for (uint32_t wid = g0; wid < g1; wid++) {
      // next body with work: usually adjacent; across a world's inactive slots
      // (and bodies without candidates) search instead of walking
      if (woff_i[i + 1] <= wid) {
        if (woff_i[i + 2] > wid) i++;
        else i = upperBound(woff_i, n_slots, wid) - 1;
      }
      if (i != cur) {
        cur = i;
        const float4 li = slo[i], hi4 = shi[i];
        const uint32_t mi = __float_as_uint(hi4.w);
        w = mi >> 2;
        idi = __float_as_uint(li.w);
        si = mi & 1u;
        bi = Box{f4v(li), f4v(hi4)};
        end = send[i];
      }
      const uint32_t j = i + 1 + (wid - woff_i[i]) * kWave + lane;
      bool emit = false;
      uint32_t idj = 0;
      if (j < end) {
        const float4 lj = slo[j], hj = shi[j];
        emit = boxOverlaps(bi, Box{f4v(lj), f4v(hj)}) && !(si && (__float_as_uint(hj.w) & 1u));
        idj = __float_as_uint(lj.w);
      }
      const uint64_t em = __ballot(emit);
      if (kCount) {
        n += (uint32_t)__popcll(em);
        if (per_world) {
          if (w != wcur) {
            if (wn && lane == 0) atomicAdd(&ws[wcur].np_emit, wn);
            wcur = w;
            wn = 0;
          }
          wn += (uint32_t)__popcll(em);
        }
      } else {
        // packed key (world, a, b) with ids of id_bits bits, expanded after sorting
        if (emit) {
          const uint64_t k =
              ((uint64_t)w << (2 * id_bits)) | ((uint64_t)min(idi, idj) << id_bits) | max(idi, idj);
          const uint32_t at = o + (uint32_t)__popcll(em & ((1ull << lane) - 1ull));
          if (at >= cap) {
            // past the output's capacity (a recorded step sized from an earlier
            // count): the overflow is detected from the count and the step redone
          } else if (narrow)
            reinterpret_cast<uint32_t*>(out)[at] = (uint32_t)k;
          else
            out[at] = k;
        }
        o += (uint32_t)__popcll(em);
      }
    }
  }
  if (kCount && lane == 0) {
    wcount[gw] = n;
    if (gw == kSweepWaves - 1) wcount[kSweepWaves] = 0;
    if (wn) atomicAdd(&ws[wcur].np_emit, wn);
  }
}

// ---------------------------------------------------------------------------
// broad phase, LBVH: a Karras radix tree per world over the Morton codes of
// the fat-box centres, refit bottom-up, queried once per leaf.  It finds the
// same pair set as the sweep (every pair of active boxes that overlap,
// inclusive, and are not both static; src/broad_phase.zig:488-526) and emits
// it in the same packed form, so the pair sort after it is shared.
//
// Per world w the leaves are its active slots in Morton order, at
// [w * max_bodies, + n_leaves) of the leaf arrays, and the n_leaves - 1
// internal nodes at the same offsets of the node arrays (node 0 the root).
// A node reference is a leaf index | kLbLeaf, an internal index, or kLbEnd.
// Every node carries its escape ("rope"): the first node after its subtree in
// depth-first order, so a query needs no stack, and a query that starts at
// leaf i's escape meets exactly the leaves after i, each pair once.
//   leaf k:     llo = lo.xyz, id     lhi = hi.xyz, escape | kLbStatic if static
//   internal i: ilo = lo.xyz, left   ihi = hi.xyz, escape
// ---------------------------------------------------------------------------
constexpr uint32_t kLbLeaf = 0x80000000u, kLbEnd = 0x40000000u, kLbStatic = 0x20000000u;
constexpr uint32_t kLbIdx = 0x1FFFFFFFu;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
// 30-bit Morton code of a fat box's doubled centre on its world's 1024^3 grid.
__device__ __forceinline__ uint32_t mortonOf(float4 lo, float4 hi, const WorldState& w) {
  const float c[3] = {lo.x + hi.x, lo.y + hi.y, lo.z + hi.z};
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float t = fminf(fmaxf((c[k] - w.c_lo[k]) * w.c_scale[k], 0.0f), 1023.0f);  // NaN -> 0
    q[k] = (uint32_t)t;
  }
  return (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
}

// Morton order of each world in one block (inactive slots last, ties in slot
// order, as the device-wide sort of (world, code) keys).
template <uint32_t kBS, uint32_t kIPT>
__global__ __launch_bounds__(kBS) void k_lbvh_sort(DevCfg cfg, DevBodies B, const WorldState* ws,
                                                   uint32_t* order) {
  __shared__ uint32_t s_k[kBS * kIPT], s_v[kBS * kIPT];
  __shared__ uint32_t s_wcnt[2 * (kBS / 64) * kRadixBins], s_run[2 * kRadixBins];
  const uint32_t w = blockIdx.x, MB = cfg.max_bodies;
  const WorldState W = ws[w];
  for (uint32_t d = threadIdx.x; d < 2 * (kBS / 64) * kRadixBins; d += kBS) s_wcnt[d] = 0;
  uint32_t keys[kIPT], vals[kIPT];
#pragma unroll
  for (uint32_t k = 0; k < kIPT; k++) {  // striped: slot k * kBS + tid
    const uint32_t sl = k * kBS + threadIdx.x;
    const uint32_t g = w * MB + sl;
    keys[k] = 0xFFFFFFFFu;
    vals[k] = g;
    if (sl < MB && B.active[g]) keys[k] = mortonOf(B.fat_lo[g], B.fat_hi[g], W);
  }
  blockRadixSort<kBS, kIPT, true>(keys, vals, MB, 32, s_k, s_v, s_wcnt, s_run);
#pragma unroll
  for (uint32_t k = 0; k < kIPT; k++) {
    const uint32_t r = k * kBS + threadIdx.x;
    if (r < MB) order[w * MB + r] = vals[k];
  }
}

__global__ void k_lbvh_keys(DevCfg cfg, DevBodies B, const WorldState* ws, uint64_t* keys,
                            uint32_t* vals, uint32_t n_slots) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_slots) return;
  uint32_t w = g / cfg.max_bodies;
  uint32_t k = B.active[g] ? mortonOf(B.fat_lo[g], B.fat_hi[g], ws[w]) : 0xFFFFFFFFu;
  keys[g] = ((uint64_t)w << 32) | k;
  vals[g] = g;
}

// Leaves in Morton order: boxes, ids, static flags and codes.
__global__ void k_lbvh_gather(DevCfg cfg, DevBodies B, const WorldState* ws,
                              const uint32_t* order, uint32_t n_slots, float4* llo, float4* lhi,
                              uint32_t* code) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_slots) return;
  const uint32_t w = i / cfg.max_bodies;
  if (i - w * cfg.max_bodies >= ws[w].n_leaves) return;
  const uint32_t g = order[i];
  float4 lo = B.fat_lo[g], hi = B.fat_hi[g];
  code[i] = mortonOf(lo, hi, ws[w]);
  lo.w = __uint_as_float(B.id[g]);
  hi.w = __uint_as_float(B.type[g] == VXP_BODY_STATIC ? kLbStatic : 0u);
  llo[i] = lo;
  lhi[i] = hi;
}

// Common-prefix length of leaves a and b of a world (codes extended by the
// leaf index so that equal codes still split); -1 outside [0, n).
__device__ __forceinline__ int lbDelta(const uint32_t* code, int n, int a, int b) {
  if (b < 0 || b >= n) return -1;
  const uint32_t ca = code[a], cb = code[b];
  return ca != cb ? __clz(ca ^ cb) : 32 + __clz((uint32_t)a ^ (uint32_t)b);
}

// Internal node i of each world (Karras 2012, "Maximizing parallelism in the
// construction of BVHs, octrees, and k-d trees", section 4): its key range
// [first, last], split and children; parents of both children; the refit
// counter cleared.
__global__ void k_lbvh_build(DevCfg cfg, const WorldState* ws, uint32_t n_slots,
                             const uint32_t* code_all, uint2* child, uint2* range,
                             uint32_t* iparent, uint32_t* lparent, uint32_t* flag) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_slots) return;
  const uint32_t w = t / cfg.max_bodies, base = w * cfg.max_bodies;
  const int n = (int)ws[w].n_leaves, i = (int)(t - base);
  if (i + 1 >= n) return;
  const uint32_t* code = code_all + base;
  const int d = lbDelta(code, n, i, i + 1) > lbDelta(code, n, i, i - 1) ? 1 : -1;
  const int dmin = lbDelta(code, n, i, i - d);
  int lmax = 2;
  while (lbDelta(code, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int s = lmax >> 1; s >= 1; s >>= 1)
    if (lbDelta(code, n, i, i + (l + s) * d) > dmin) l += s;
  const int j = i + l * d;
  const int dnode = lbDelta(code, n, i, j);
  int s = 0;
  for (int div = 2;; div <<= 1) {
    const int st = (l + div - 1) / div;
    if (lbDelta(code, n, i, i + (s + st) * d) > dnode) s += st;
    if (st == 1) break;
  }
  const int gamma = i + s * d + min(d, 0);
  const int lo = min(i, j), hi = max(i, j);
  const uint32_t left = lo == gamma ? (kLbLeaf | (uint32_t)gamma) : (uint32_t)gamma;
  const uint32_t right = hi == gamma + 1 ? (kLbLeaf | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
  child[t] = make_uint2(left, right);
  range[t] = make_uint2((uint32_t)lo, (uint32_t)hi);
  flag[t] = 0;
  if (left & kLbLeaf) lparent[base + gamma] = (uint32_t)i;
  else iparent[base + gamma] = (uint32_t)i;
  if (right & kLbLeaf) lparent[base + gamma + 1] = (uint32_t)i;
  else iparent[base + gamma + 1] = (uint32_t)i;
}

// Escape of a node whose key range ends at leaf `last`: the subtree that
// starts at leaf last + 1 (internal node last + 1 when its range starts
// there, else that leaf), or kLbEnd after the world's last leaf.
__device__ __forceinline__ uint32_t lbEscape(const uint2* range, int n, int last) {
  if (last + 1 >= n) return kLbEnd;
  if (last + 1 <= n - 2 && (int)range[last + 1].x == last + 1) return (uint32_t)(last + 1);
  return kLbLeaf | (uint32_t)(last + 1);
}

// Leaf k's escape, then node boxes bottom-up from it: the second of a node's
// two children to finish computes the node (the counter orders the writes of
// the first before the reads of the second), up to the root.  kWg: the
// world's leaves are all handled by one workgroup, counters in LDS and
// workgroup-scope ordering; otherwise device-scope counters and fences.
template <bool kWg>
__device__ __forceinline__ void lbRefitLeaf(uint32_t base, int n, int k, float4* llo, float4* lhi,
                                            const uint2* child, const uint2* range,
                                            const uint32_t* iparent, const uint32_t* lparent,
                                            uint32_t* flag, float4* ilo, float4* ihi) {
  {
    float4 h = lhi[base + k];
    h.w = __uint_as_float((__float_as_uint(h.w) & kLbStatic) | lbEscape(range + base, n, k));
    lhi[base + k] = h;
  }
  if (n < 2) return;
  uint32_t p = lparent[base + k];
  for (int up = 0; up < n; up++) {  // at most n - 1 levels
    if (kWg) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (atomicAdd(&flag[p], 1u) == 0u) break;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
      __threadfence();
      if (atomicAdd(&flag[base + p], 1u) == 0u) break;
      __threadfence();
    }
    const uint2 ch = child[base + p];
    const uint32_t a = ch.x & kLbIdx, b = ch.y & kLbIdx;
    const float4 alo = (ch.x & kLbLeaf) ? llo[base + a] : ilo[base + a];
    const float4 ahi = (ch.x & kLbLeaf) ? lhi[base + a] : ihi[base + a];
    const float4 blo = (ch.y & kLbLeaf) ? llo[base + b] : ilo[base + b];
    const float4 bhi = (ch.y & kLbLeaf) ? lhi[base + b] : ihi[base + b];
    const float4 lo = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z),
                                  __uint_as_float(ch.x));
    const float4 hi = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z),
                                  __uint_as_float(lbEscape(range + base, n, (int)range[base + p].y)));
    ilo[base + p] = lo;
    ihi[base + p] = hi;
    if (p == 0) break;
    p = iparent[base + p];
  }
}

// One thread per leaf slot over every world (counters cleared by the build).
__global__ void k_lbvh_refit(DevCfg cfg, const WorldState* ws, uint32_t n_slots, float4* llo,
                             float4* lhi, const uint2* child, const uint2* range,
                             const uint32_t* iparent, const uint32_t* lparent, uint32_t* flag,
                             float4* ilo, float4* ihi) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_slots) return;
  const uint32_t w = t / cfg.max_bodies, base = w * cfg.max_bodies;
  const int n = (int)ws[w].n_leaves, k = (int)(t - base);
  if (k >= n) return;
  lbRefitLeaf<false>(base, n, k, llo, lhi, child, range, iparent, lparent, flag, ilo, ihi);
}

// One workgroup per world, counters in LDS (max_bodies words): the fences
// stay inside the workgroup's compute unit instead of reaching every XCD's L2.
__global__ void k_lbvh_refit_wg(DevCfg cfg, const WorldState* ws, float4* llo, float4* lhi,
                                const uint2* child, const uint2* range, const uint32_t* iparent,
                                const uint32_t* lparent, float4* ilo, float4* ihi) {
  extern __shared__ uint32_t s_flag[];
  const uint32_t w = blockIdx.x, base = w * cfg.max_bodies;
  const int n = (int)ws[w].n_leaves;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s_flag[i] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x)
    lbRefitLeaf<true>(base, n, k, llo, lhi, child, range, iparent, lparent, s_flag, ilo, ihi);
}

__device__ __forceinline__ bool lbOverlap(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  return alo.x <= bhi.x && ahi.x >= blo.x && alo.y <= bhi.y && ahi.y >= blo.y &&
         alo.z <= bhi.z && ahi.z >= blo.z;
}

// One query per leaf from its escape: count its pairs (cnt[i]; cnt[n] = 0),
// or write them from off[i] as packed (world, min id, max id) keys, 32-bit
// when `narrow`.
template <bool kCount>
__global__ __launch_bounds__(256) void k_lbvh_pairs(DevCfg cfg, WorldState* ws,
                                                    uint32_t n_slots, const float4* llo,
                                                    const float4* lhi, const float4* ilo,
                                                    const float4* ihi, uint32_t* cnt,
                                                    const uint32_t* off, uint64_t* out,
                                                    uint32_t id_bits, bool narrow) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_slots) return;
  if (kCount && t == n_slots - 1) cnt[n_slots] = 0;
  const uint32_t w = t / cfg.max_bodies, base = w * cfg.max_bodies;
  const uint32_t n = ws[w].n_leaves;
  uint32_t c = 0;
  if (t - base < n) {
    const float4 qlo = llo[t], qhi = lhi[t];
    const uint32_t idi = __float_as_uint(qlo.w), qm = __float_as_uint(qhi.w);
    const bool si = qm & kLbStatic;
    uint32_t node = qm & ~kLbStatic, o = kCount ? 0u : off[t];
    // a walk visits each of the 2n - 1 nodes at most once
    for (uint32_t steps = 0; !(node & kLbEnd); steps++) {
      if (steps >= 2 * n) {
        atomicOr(&ws[w].error, kErrLbvh);
        break;
      }
      const uint32_t x = base + (node & kLbIdx);
      if (node & kLbLeaf) {
        const float4 lo = llo[x], hi = lhi[x];
        const uint32_t m = __float_as_uint(hi.w);
        if (lbOverlap(qlo, qhi, lo, hi) && !(si && (m & kLbStatic))) {
          if (kCount) {
            c++;
          } else {
            const uint32_t idj = __float_as_uint(lo.w);
            const uint64_t key = ((uint64_t)w << (2 * id_bits)) |
                                 ((uint64_t)min(idi, idj) << id_bits) | max(idi, idj);
            if (narrow) reinterpret_cast<uint32_t*>(out)[o] = (uint32_t)key;
            else out[o] = key;
            o++;
          }
        }
        node = m & ~kLbStatic;
      } else {
        const float4 lo = ilo[x], hi = ihi[x];
        node = lbOverlap(qlo, qhi, lo, hi) ? __float_as_uint(lo.w) : __float_as_uint(hi.w);
      }
    }
  }
  if (kCount) cnt[t] = c;
}

// ---------------------------------------------------------------------------
// narrow phase part 1: filter, SAT, cache key, overlap region
// ---------------------------------------------------------------------------
__device__ __forceinline__ void projObb(const V3 v[8], V3 ax, float& mn, float& mx) {
  mn = INFINITY;
  mx = -INFINITY;
  for (int i = 0; i < 8; i++) {
    float p = dot(v[i], ax);
    mn = fminf(mn, p);
    mx = fmaxf(mx, p);
  }
}
__device__ __forceinline__ bool sepAxis(const V3 va[8], const V3 vb[8], V3 ax) {
  float a0, a1, b0, b1;
  projObb(va, ax, a0, a1);
  projObb(vb, ax, b0, b1);
  return a1 < b0 || b1 < a0;
}
// satEarlyOut (src/narrow_phase.zig:532-569): true when separated
__device__ bool satSeparated(const Obb& A, const Obb& Bo) {
  V3 va[8], vb[8], aa[3], bb[3];
  obbVertices(A, va);
  obbVertices(Bo, vb);
  obbAxes(A, aa);
  obbAxes(Bo, bb);
  for (int i = 0; i < 3; i++)
    if (sepAxis(va, vb, aa[i])) return true;
  for (int i = 0; i < 3; i++)
    if (sepAxis(va, vb, bb[i])) return true;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      V3 cr = cross(aa[i], bb[j]);
      float l = len(cr);
      if (l < 0.001f) continue;
      if (sepAxis(va, vb, scale(cr, 1.0f / l))) return true;
    }
  return false;
}

// src/narrow_phase.zig:148-169
__device__ uint64_t transformHash(V3 rp, Q4 rq) {
  int32_t qx = (int32_t)(rp.x * 100.0f), qy = (int32_t)(rp.y * 100.0f), qz = (int32_t)(rp.z * 100.0f);
  int16_t rw = (int16_t)(int32_t)clampz(rq.w * 1000.0f, -32767.0f, 32767.0f);
  int16_t rx = (int16_t)(int32_t)clampz(rq.x * 1000.0f, -32767.0f, 32767.0f);
  int16_t ry = (int16_t)(int32_t)clampz(rq.y * 1000.0f, -32767.0f, 32767.0f);
  int16_t rz = (int16_t)(int32_t)clampz(rq.z * 1000.0f, -32767.0f, 32767.0f);
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

__device__ int32_t findPrev(const DevManifold* prev, uint32_t lo, uint32_t hi, uint64_t key) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint64_t k = prev[mid].key;
    if (k == key) return (int32_t)mid;
    if (k < key) lo = mid + 1; else hi = mid;
  }
  return -1;
}

struct PairCounts {
  uint32_t bp, skipped, tested;
};

__device__ PairCounts pairProcess(const DevCfg& cfg, const DevBodies& B, const DevColliders& C,
                                  WorldState* ws, const uint64_t* keys, uint32_t n_pairs,
                                  uint32_t p, PairRec& r, const DevManifold* prev,
                                  const uint8_t* drop) {
  PairCounts cnt{0, 0, 0};
  uint64_t key = keys[p];
  uint32_t w = keyWorld(key), ida = keyA(key), idb = keyB(key);
  if (p == 0 || keyWorld(keys[p - 1]) != w) ws[w].pair_lo = p;
  if (p == n_pairs - 1 || keyWorld(keys[p + 1]) != w) ws[w].pair_hi = p + 1;
  r.world = w;
  r.ida = ida;
  r.idb = idb;
  r.flag = PF_NONE;
  r.nchunks = 0;
  r.vchk = 0;
  r.aux = kNone;
  r.th = 0;
  r.mn[0] = r.mn[1] = r.mn[2] = 0;
  r.nx = r.ny = r.nz = 0;
  int32_t sa = B.idmap[(size_t)w * cfg.id_cap + ida];
  int32_t sb = B.idmap[(size_t)w * cfg.id_cap + idb];
  r.sa = (uint32_t)sa;
  r.sb = (uint32_t)sb;
  if (ws[w].skip || sa < 0 || sb < 0) {
    return cnt;
  }
  cnt.bp = 1;
  uint32_t ga = w * cfg.max_bodies + sa, gb = w * cfg.max_bodies + sb;
  bool asl = B.sleeping[ga], bsl = B.sleeping[gb];
  if (asl && bsl) {
    cnt.skipped = 1;
    int32_t j = findPrev(prev, ws[w].prev_start, ws[w].prev_start + ws[w].prev_count, key);
    if (j >= 0) {
      r.flag = PF_CARRY;
      r.aux = (uint32_t)j;
    }
    return cnt;
  }
  if (drop && drop[p]) {  // past the filtered-pair array (src/system.zig:594,611-614)
    return cnt;
  }
  if (cfg.threaded) {
    // the thread-pool narrow phase sees the first 256 active bodies only; its
    // parallel branch counts every pair as tested before looking the bodies
    // up (src/system.zig:618-629, src/narrow_phase.zig:1188-1199)
    const WorldState& W = ws[w];
    if ((uint32_t)sa >= W.thr_cut || (uint32_t)sb >= W.thr_cut) {
      cnt.tested = W.thr_par;
      return cnt;
    }
  }
  cnt.tested = 1;
  int32_t ca = B.collider[ga], cb = B.collider[gb];
  if (ca < 0 || cb < 0) {
    return cnt;
  }
  const DevCollider& HA = C.hdr[ca];
  const DevCollider& HB = C.hdr[cb];
  V3 pa = f4v(B.pos[ga]), pb = f4v(B.pos[gb]);
  Q4 qa = f4q(B.rot[ga]), qb = f4q(B.rot[gb]);
  if (cfg.use_sat) {
    Obb oa = obbFromBox(colliderBox(HA), pa, qa);
    Obb ob = obbFromBox(colliderBox(HB), pb, qb);
    if (satSeparated(oa, ob)) {
      return cnt;
    }
  }
  uint8_t ta = B.type[ga], tb = B.type[gb];
  bool use_cache = cfg.use_cache && (ta == VXP_BODY_STATIC || tb == VXP_BODY_STATIC || asl || bsl);
  if (use_cache) {
    r.th = transformHash(sub(pb, pa), qmul(qconj(qa), qb));
    r.flag = PF_CACHE;
  } else {
    r.flag = PF_DETECT;
  }
  // getOverlapRegion (src/narrow_phase.zig:644-698)
  Obb obb = obbFromBox(colliderBox(HB), pb, qb);
  V3 vb[8];
  obbVertices(obb, vb);
  Q4 inv = qconj(qa);
  V3 tmin{INFINITY, INFINITY, INFINITY}, tmax{-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < 8; i++) {
    V3 rel = qrotate(inv, sub(vb[i], pa));
    tmin = vmin(tmin, rel);
    tmax = vmax(tmax, rel);
  }
  float vs = HA.vs;
  tmin = sub(tmin, V3{vs, vs, vs});
  tmax = add(tmax, V3{vs, vs, vs});
  V3 alo{HA.lo[0], HA.lo[1], HA.lo[2]}, ahi{HA.hi[0], HA.hi[1], HA.hi[2]};
  V3 mnc = vclamp(tmin, alo, ahi), mxc = vclamp(tmax, alo, ahi);
  if (!(mnc.x >= mxc.x || mnc.y >= mxc.y || mnc.z >= mxc.z)) {
    int32_t a[3], b[3];
    localToVoxel(vs, mnc, a);
    localToVoxel(vs, mxc, b);
    for (int k = 0; k < 3; k++) {
      a[k] = max(a[k], HA.minb[k]);
      b[k] = min(b[k], HA.maxb[k]);
    }
    if (b[0] >= a[0] && b[1] >= a[1] && b[2] >= a[2]) {
      r.mn[0] = a[0];
      r.mn[1] = a[1];
      r.mn[2] = a[2];
      r.nx = (uint32_t)(b[0] - a[0] + 1);
      r.ny = (uint32_t)(b[1] - a[1] + 1);
      r.nz = (uint32_t)(b[2] - a[2] + 1);
      uint64_t nv = (uint64_t)r.nx * r.ny * r.nz;
      r.nchunks = (uint32_t)((nv + kWave - 1) / kWave);
    }
  }
  return cnt;
}

// Also flags the cache-eligible pairs (cflag[0..n_pairs], last entry 0) for
// their ordered compaction, clears the pair's reduction outputs and appends
// the pairs that the reduction must visit to wlist (any order).
__global__ __launch_bounds__(512) void k_pair_process(DevCfg cfg, DevBodies B, DevColliders C, WorldState* ws,
                               vxp_stats* st, const uint64_t* keys, uint32_t n_pairs, PairRec* pr,
                               const DevManifold* prev, const uint8_t* drop, uint32_t* cflag,
                               uint32_t* produce, uint32_t* p_vchk, uint32_t* wlist,
                               uint32_t* step_ctr, uint32_t* counts, bool chunk_bound,
                               const uint32_t* n_dev) {
  __shared__ uint32_t s_n, s_base, s_chunks;
  n_pairs = devCount(n_pairs, n_dev);
  // block-strided (a grid smaller than the pairs, as in a recorded step, loops)
  for (uint32_t base = blockIdx.x * blockDim.x; base < n_pairs; base += gridDim.x * blockDim.x) {
  const uint32_t p = base + threadIdx.x;
  PairCounts c{0, 0, 0};
  uint32_t w = 0;
  bool work = false;
  if (threadIdx.x == 0) s_n = s_chunks = 0;
  __syncthreads();
  uint32_t pos = 0;
  if (p < n_pairs) {
    PairRec r;
    c = pairProcess(cfg, B, C, ws, keys, n_pairs, p, r, prev, drop);
    w = r.world;
    const uint32_t f = r.flag;
    // only the pairs with work have readers of their record (the reduction's
    // work list, the cache replay, the probe)
    if (f != PF_NONE) pr[p] = r;
    cflag[p] = f == PF_CACHE;
    // work items of the voxel probe: the region's chunks when the pair runs the
    // detection (a cache pair's count is set by the replay when it misses)
    counts[p] = f == PF_DETECT ? r.nchunks : 0u;
    if (p == n_pairs - 1) cflag[n_pairs] = 0;
    produce[p] = 0;
    p_vchk[p] = 0;  // the probe's count pass adds the pair's checked voxels
    work = f == PF_CARRY || f == PF_DETECT || f == PF_CACHE;
    if (work) pos = atomicAdd(&s_n, 1u);
    // work items of this pair if it runs the detection (a cache pair only on a
    // miss): the step's bound on the probe's work items
    if (chunk_bound && (f == PF_DETECT || f == PF_CACHE) && r.nchunks)
      atomicAdd(&s_chunks, r.nchunks);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_base = s_n ? atomicAdd(&step_ctr[kCtrReduceWork], s_n) : 0;
    if (s_chunks) atomicAdd(&step_ctr[kCtrChunkBound], s_chunks);
  }
  __syncthreads();
  if (work) wlist[s_base + pos] = p;
  const uint32_t v[3] = {c.bp, c.skipped, c.tested};
  uint32_t* const ctr[3] = {&st->broad_phase_pairs, &st->pairs_skipped_sleeping,
                            &st->narrow_phase_tested};
  blockAggAdd<3>(w, v, ctr, sizeof(vxp_stats) / 4);
  __syncthreads();  // the next round reuses the block's LDS counters
  }
}



// ---------------------------------------------------------------------------
// BrickPairCache replay: one thread per world walks its cache-eligible pairs in
// pair order (get, then put on a miss).  Puts reserve their entry at miss time
// and record (frame, pair) as owner; values are written by k_reduce_tail after
// the hits have been read.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t cacheHash(uint32_t a, uint32_t b, uint64_t th) {
  uint64_t h = 0;
  h = h * 31 + a;
  h = h * 31 + b;
  for (int i = 0; i < 6; i++) h = h * 31;
  return h * 31 + th;
}

__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// The fields of a cache entry that a lookup reads.
struct CacheWin {
  uint32_t a, b, valid;
  uint64_t th, lu;
};
__device__ __forceinline__ CacheWin cacheWin(const CacheKeyRec& e) {
  return CacheWin{e.a, e.b, e.valid, e.th, e.last_used};
}
// Minimum of v over each row of 16 lanes, in every lane of the row (DPP: quad
// swaps, then rotations by 4 and 8 within the row).  The whole wave must be active.
__device__ __forceinline__ uint32_t rowMin16(uint32_t v) {
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false));   // [1,0,3,2]
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false));   // [2,3,0,1]
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, false));  // row_ror:4
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false));  // row_ror:8
  return v;
}

// Entry of probe lane `lane` (< 16 <= cap) of the window that starts at `start`.
__device__ __forceinline__ uint32_t cacheSlot(uint32_t start, uint32_t lane, uint32_t cap) {
  uint32_t i = start + lane;
  return i >= cap ? i - cap : i;
}

// Cache-eligible pairs in the narrow phase's order: position i of that order
// holds pair order[i] (i itself in key order, order == nullptr); cflag and
// cidx are indexed by position: clist[cidx[i]] = pair.
__global__ void k_cache_scatter(const uint32_t* cflag, const uint32_t* cidx, const uint32_t* order,
                                uint32_t n_pairs, uint32_t* clist, const uint32_t* n_dev) {
  n_pairs = devCount(n_pairs, n_dev);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_pairs; i += gridDim.x * blockDim.x)
    if (cflag[i]) clist[cidx[i]] = order ? order[i] : i;
}

// One block per world: its kCacheThreads threads stage the key table in LDS
// and write it back; wave 0 replays the gets/puts of the world's
// cache-eligible pairs (clist[cidx[pair_lo] .. cidx[pair_hi])) strictly in
// pair order: lanes 0..15 probe the 16-entry window of one get/put in
// parallel and a ballot picks the first stopping entry.  A world without
// gets/puts this step leaves its table untouched.
constexpr uint32_t kCacheThreads = 256;
__device__ __forceinline__ void copyWords(uint2* dst, const uint2* src, uint32_t n) {
  const uint32_t tid = threadIdx.x;
  uint32_t i = tid;
  for (; i + 3 * kCacheThreads < n; i += 4 * kCacheThreads) {
    const uint2 a = src[i], b = src[i + kCacheThreads], c = src[i + 2 * kCacheThreads],
                d = src[i + 3 * kCacheThreads];
    dst[i] = a;
    dst[i + kCacheThreads] = b;
    dst[i + 2 * kCacheThreads] = c;
    dst[i + 3 * kCacheThreads] = d;
  }
  for (; i < n; i += kCacheThreads) dst[i] = src[i];
}
__global__ __launch_bounds__(kCacheThreads) void k_cache_replay(DevCfg cfg, WorldState* ws,
                                                                vxp_stats* st, PairRec* pr,
                                                                CacheKeyRec* keys,
                                                                const uint32_t* cidx,
                                                                const uint32_t* clist,
                                                                uint32_t* counts,
                                                                uint32_t n_worlds) {
  extern __shared__ CacheKeyRec E[];
  const uint32_t w = blockIdx.x, tid = threadIdx.x, lane = tid & (kWave - 1);
  if (w >= n_worlds || ws[w].skip) return;
  const uint32_t frame = ws[w].frame + 1;
  __syncthreads();
  if (tid == 0) ws[w].frame = frame;
  if (!cfg.use_cache) return;
  const uint32_t lo = ws[w].pair_lo, hi = ws[w].pair_hi;
  const uint32_t c0 = lo < hi ? cidx[lo] : 0, total = lo < hi ? cidx[hi] - c0 : 0;
  if (!total) return;
  const uint32_t cap = cfg.cache_cap, maxp = cap < 16 ? cap : 16;
  CacheKeyRec* G = keys + (size_t)w * cap;
  // stage the key table as 8-byte words (an entry is 40 bytes)
  static_assert(sizeof(CacheKeyRec) % 8 == 0, "cache key record");
  const uint32_t nq = cap * (uint32_t)sizeof(CacheKeyRec) / 8;
  copyWords(reinterpret_cast<uint2*>(E), reinterpret_cast<const uint2*>(G), nq);
  __syncthreads();
  uint32_t hits = 0, misses = 0;
  // the thread-pool narrow phase's parallel branch only reads the cache: a
  // miss stores nothing (src/narrow_phase.zig:1228-1260)
  const bool store = !(cfg.threaded && ws[w].thr_par);
  if (tid < kWave) {
    const bool probe = lane < maxp;
    // pair data of a batch of 64 pairs; the next batch's is loaded (and each
    // lane hashes its own pair) while the current batch is walked
    auto fetch = [&](uint32_t base, uint32_t& p, uint32_t& ida, uint32_t& idb, uint64_t& th) {
      const bool inb = base + lane < total;
      p = inb ? clist[c0 + base + lane] : 0;
      ida = inb ? pr[p].ida : 0;
      idb = inb ? pr[p].idb : 0;
      th = inb ? pr[p].th : 0;
    };
    uint32_t np = 0, nida = 0, nidb = 0;
    uint64_t nth = 0;
    if (total) fetch(0, np, nida, nidb, nth);
    for (uint32_t base = 0; base < total; base += kWave) {
      const bool inb = base + lane < total;
      const uint32_t p = np, ida = nida, idb = nidb;
      const uint64_t th = nth;
      if (base + kWave < total) fetch(base + kWave, np, nida, nidb, nth);
      const uint32_t my_start = (uint32_t)(cacheHash(ida, idb, th) % cap);
      uint32_t my_flag = PF_CACHE, my_aux = kNone;
      const uint32_t nb = min(total - base, (uint32_t)kWave);
      // The window of the pair being resolved is held in registers (lane l <
      // maxp: entry start + l).  The next pair's window is loaded before this
      // pair's update is stored; LDS operations of a wave complete in order, so
      // that load sees every earlier update and only this pair's is patched in.
      uint32_t idx = cacheSlot(rdlane(my_start, 0), lane, cap);
      CacheWin cur = probe ? cacheWin(E[idx]) : CacheWin{};
      for (uint32_t j = 0; j < nb; j++) {
        const uint32_t a = rdlane(ida, (int)j), b = rdlane(idb, (int)j);
        const uint64_t t = ((uint64_t)rdlane((uint32_t)(th >> 32), (int)j) << 32) |
                           rdlane((uint32_t)th, (int)j);
        const bool more = j + 1 < nb;
        uint32_t nidx = 0;
        CacheWin nxt{};
        if (more) {
          nidx = cacheSlot(rdlane(my_start, (int)j + 1), lane, cap);
          if (probe) nxt = cacheWin(E[nidx]);
        }
        const bool eq = cur.valid && cur.a == a && cur.b == b && cur.th == t;
        const uint64_t stop = __ballot(probe && (!cur.valid || eq));
        uint32_t rflag, raux;
        if (stop) {
          const int f = __ffsll((long long)stop) - 1;
          raux = rdlane(idx, f);
          rflag = rdlane((uint32_t)eq, f) ? PF_HIT : PF_MISS;
        } else {
          // window full: evict the first entry with the smallest last_used
          // (last_used only ever holds a 32-bit frame number; the probe lanes
          // are lanes 0..maxp-1, all in the first row of 16)
          const uint32_t lu = probe ? (uint32_t)cur.lu : 0xFFFFFFFFu;
          const uint32_t lo = rdlane(rowMin16(lu), 0);
          const uint64_t at = __ballot(probe && lu == lo);
          raux = rdlane(idx, __ffsll((long long)at) - 1);
          rflag = PF_MISS;
        }
        if (lane == 0) {
          if (rflag == PF_HIT) {
            E[raux].last_used = frame;
            hits++;
          } else if (store) {
            CacheKeyRec& e = E[raux];
            e.a = a;
            e.b = b;
            e.th = t;
            e.last_used = frame;
            e.valid = 1;
            e.owner_frame = frame;
            e.owner_pair = rdlane(p, (int)j);
            misses++;
          } else {
            misses++;
          }
        }
        if (more && probe && nidx == raux && (store || rflag == PF_HIT)) {
          nxt.lu = frame;
          if (rflag == PF_MISS) {
            nxt.a = a;
            nxt.b = b;
            nxt.th = t;
            nxt.valid = 1;
          }
        }
        if (lane == j) {
          my_flag = rflag;
          my_aux = raux;
        }
        cur = nxt;
        idx = nidx;
      }
      if (inb) {
        pr[p].flag = my_flag;
        pr[p].aux = my_aux;
        if (my_flag == PF_MISS) counts[p] = pr[p].nchunks;  // a miss runs the detection
      }
    }
  }
  __syncthreads();
  copyWords(reinterpret_cast<uint2*>(G), reinterpret_cast<const uint2*>(E), nq);
  if (tid == 0) {
    st[w].cache_hits += hits;
    st[w].cache_misses += misses;
  }
}


// The pair of every work item: wi_pair[wi_off[p] + c] = p for c < counts[p].
__global__ void k_wi_fill(const uint32_t* counts, const uint32_t* wi_off, uint32_t n_pairs,
                          uint32_t* wi_pair, uint32_t* wi_ccnt, const uint32_t* n_dev) {
  n_pairs = devCount(n_pairs, n_dev);
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    const uint32_t c = counts[p], b = wi_off[p];
    for (uint32_t k = 0; k < c; k++) {
      wi_pair[b + k] = p;
      wi_ccnt[b + k] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// voxel probe: one wave per 64-voxel chunk of a pair's overlap region (region
// order z, y, x as src/narrow_phase.zig:759-764; lane = position in the chunk).
// Each lane runs the 27 probes of checkNeighborhood in (dz, dy, dx) order.
// Candidates of a chunk are written contiguously in (lane, probe) order, so
// chunk order + in-chunk order is the reference's candidate sequence.
// ---------------------------------------------------------------------------
// A rotation as qrotate applies it (src/math.zig:267-275): the quaternion
// after qrotate's norm check, and 2w.  Derived once per pair, xformBy and
// invXformBy give the bits of xformPoint / invXformPoint (src/math.zig) with
// the body's quaternion, without the per-point normalisation (a square root
// and a division of uniform values).
struct Rot {
  V3 v;
  float w2;
};
__device__ __forceinline__ Rot rotOf(Q4 q) {
  float ls = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (ls < 0.99f || ls > 1.01f) q = qnormalize(q);
  return Rot{V3{q.x, q.y, q.z}, 2.0f * q.w};
}
__device__ __forceinline__ V3 rotApply(const Rot& r, V3 v) {
  V3 uv = cross(r.v, v);
  V3 uuv = cross(r.v, uv);
  return add(add(v, scale(uv, r.w2)), scale(uuv, 2.0f));
}
// xformPoint(pos, rot, p) = qrotate(qnormalize(rot), p) + pos
__device__ __forceinline__ V3 xformBy(const Rot& r, V3 pos, V3 p) {
  return add(rotApply(r, p), pos);
}
// invXformPoint(pos, rot, p) = qrotate(qconj(rot), p - pos)
__device__ __forceinline__ V3 invXformBy(const Rot& r, V3 pos, V3 p) {
  return rotApply(r, sub(p, pos));
}

struct ProbeCtx {
  V3 pa, pb;
  Q4 qa, qb;
  Rot xa, xb, ia, ib;  // rotations of xformPoint and invXformPoint for A and B
  float vsa, vsb, spec;
  int32_t bminb[3], bmaxb[3];
};
__device__ __forceinline__ void ctxRots(ProbeCtx& c) {
  c.xa = rotOf(qnormalize(c.qa));
  c.xb = rotOf(qnormalize(c.qb));
  c.ia = rotOf(qconj(c.qa));
  c.ib = rotOf(qconj(c.qb));
}

// Probe order of a neighbour (dz, dy, dx in -1..1): (dz+1)*9 + (dy+1)*3 + (dx+1)
// (checkNeighborhood, src/narrow_phase.zig:796-857).  The write pass records
// each candidate as (penetration, feature id, lane << 5 | probe); the full
// contact of the few that survive the top-4 is rebuilt by contactAt.

// checkNeighborhood's probes of A voxel (x, y, z) from the brick masks: the
// probes whose B voxel is in bounds, non-empty and passes shouldCheck with
// the A voxel's type (any non-empty one for an A corner, corner or edge for
// an A edge, corner otherwise), as bit (dz+1)*9 + (dy+1)*3 + (dx+1) of the
// result, so that ascending bits are the probe order; *vchk = the non-empty
// in-bounds B voxels.  The window spans at most two bricks per axis: one
// directory read per brick and one read per mask plane per z slice, which
// covers the slice's rows (bit x + 8y of the slice's word).
// A voxel (x, y, z) in world space (wa) and the B voxel around it (cb);
// false when no probe of its neighbourhood is in B's voxel bounds.
__device__ __forceinline__ bool probeWindow(const ProbeCtx& c, int32_t x, int32_t y, int32_t z,
                                            V3& wa, int32_t cb[3]) {
  wa = xformBy(c.xa, c.pa, voxelToLocal(c.vsa, x, y, z));
  V3 lb = invXformBy(c.ib, c.pb, wa);
  localToVoxel(c.vsb, lb, cb);
  bool inb = true;
  for (int k = 0; k < 3; k++)
    inb = inb && max(cb[k] - 1, c.bminb[k]) <= min(cb[k] + 1, c.bmaxb[k] - 1);
  return inb;
}
// The probe set of probeMask from the window around cb (in bounds).
template <bool kInside>
__device__ __forceinline__ uint32_t probeMaskAt(const int32_t* __restrict__ dir,
                                                const uint64_t* __restrict__ mask,
                                                const VoxMap& mb, const ProbeCtx& c,
                                                const int32_t cb[3], uint32_t ta, uint32_t* vchk,
                                                uint32_t* inside) {
  // in-bounds window of the neighbourhood per axis
  int32_t lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = max(cb[k] - 1, c.bminb[k]);
    hi[k] = min(cb[k] + 1, c.bmaxb[k] - 1);
  }
  uint32_t pm = 0, vc = 0, im = 0;
  for (int32_t zb = lo[2] >> 3; zb <= hi[2] >> 3; zb++)
    for (int32_t yb = lo[1] >> 3; yb <= hi[1] >> 3; yb++)
      for (int32_t xb8 = lo[0] >> 3; xb8 <= hi[0] >> 3; xb8++) {
        const int32_t gx = xb8 - mb.org[0], gy = yb - mb.org[1], gz = zb - mb.org[2];
        if ((uint32_t)gx >= (uint32_t)mb.dim[0] || (uint32_t)gy >= (uint32_t)mb.dim[1] ||
            (uint32_t)gz >= (uint32_t)mb.dim[2])
          continue;
        const int32_t bi = dir[mb.dir_off + ((uint32_t)gz * mb.dim[1] + gy) * mb.dim[0] + gx];
        if (bi < 0) continue;
        const int32_t xa = max(lo[0], xb8 * 8), xe = min(hi[0], xb8 * 8 + 7);
        const int32_t ya = max(lo[1], yb * 8), ye = min(hi[1], yb * 8 + 7);
        const int32_t za = max(lo[2], zb * 8), ze = min(hi[2], zb * 8 + 7);
        const uint32_t wmask = (1u << (xe - xa + 1)) - 1u;
        const uint64_t* mz = mask + (size_t)bi * kMaskWords;
        for (int32_t bz = za; bz <= ze; bz++) {
          const uint64_t w0 = mz[bz & 7];
          uint64_t wt = ~0ull, wi = 0;
          if (ta != VXP_VOXEL_CORNER) {
            wt = mz[8 + (bz & 7)];
            if (ta == VXP_VOXEL_EDGE) wt |= mz[16 + (bz & 7)];
          } else if (kInside) {
            wi = mz[24 + (bz & 7)];  // only an A corner takes inside B voxels
          }
          for (int32_t by = ya; by <= ye; by++) {
            const uint32_t sh = 8u * (uint32_t)(by & 7) + (uint32_t)(xa & 7);
            const uint32_t occ = (uint32_t)(w0 >> sh) & wmask;
            const uint32_t at = (bz - cb[2] + 1) * 9 + (by - cb[1] + 1) * 3 + (xa - cb[0] + 1);
            vc += (uint32_t)__builtin_popcount(occ);
            pm |= (occ & (uint32_t)(wt >> sh)) << at;
            if (kInside) im |= (occ & (uint32_t)(wi >> sh)) << at;
          }
        }
      }
  *vchk = vc;
  if (kInside) *inside = im;
  return pm;
}
template <bool kInside>
__device__ __forceinline__ uint32_t probeMask(const int32_t* __restrict__ dir,
                                              const uint64_t* __restrict__ mask, const VoxMap& mb,
                                              const ProbeCtx& c, int32_t x, int32_t y, int32_t z,
                                              uint32_t ta, V3& wa, int32_t cb[3],
                                              uint32_t* vchk, uint32_t* inside = nullptr) {
  if (!probeWindow(c, x, y, z, wa, cb)) {
    *vchk = 0;
    if (kInside) *inside = 0;
    return 0;
  }
  return probeMaskAt<kInside>(dir, mask, mb, c, cb, ta, vchk, inside);
}
// B voxel of probe k around the B voxel cb.
__device__ __forceinline__ void probeB(const int32_t cb[3], uint32_t k, int32_t& bx, int32_t& by,
                                       int32_t& bz) {
  bz = cb[2] + (int32_t)(k / 9u) - 1;
  by = cb[1] + (int32_t)(k / 3u % 3u) - 1;
  bx = cb[0] + (int32_t)(k % 3u) - 1;
}

// The counts of checkNeighborhood (candidates n, non-empty probed B voxels
// vchk): the probes of probeMask that pass the distance test of
// checkNeighborhood (src/narrow_phase.zig:831-840).
__device__ __forceinline__ uint32_t probeCountAt(const int32_t* __restrict__ dir,
                                                 const uint64_t* __restrict__ mask,
                                                 const VoxMap& mb, const ProbeCtx& c, V3 wa,
                                                 const int32_t cb[3], uint8_t va,
                                                 uint32_t* vchk) {
  uint32_t pm = probeMaskAt<false>(dir, mask, mb, c, cb, va & 7u, vchk, nullptr);
  uint32_t n = 0;
  while (pm) {
    const uint32_t k = (uint32_t)__builtin_ctz(pm);
    pm &= pm - 1;
    int32_t bx, by, bz;
    probeB(cb, k, bx, by, bz);
    V3 wb = xformBy(c.xb, c.pb, voxelToLocal(c.vsb, bx, by, bz));
    V3 diff = sub(wb, wa);
    float dsq = dot(diff, diff);
    if (dsq < c.spec) n++;
  }
  return n;
}

// The candidates of checkNeighborhood: the first `limit` of A voxel
// (x, y, z) in probe order, as (penetration, feature id, lane << 5 | probe).
// The penetration rule's INSIDE test of the B voxel reads the inside plane.
__device__ __forceinline__ void probeWrite(const int32_t* __restrict__ dir,
                                           const uint64_t* __restrict__ mask, const VoxMap& mb,
                                           const ProbeCtx& c, int32_t x, int32_t y, int32_t z,
                                           uint8_t va, uint32_t lane, Cand* __restrict__ out,
                                           uint32_t limit) {
  V3 wa;
  int32_t cb[3];
  uint32_t vc, im;
  const uint32_t ta = va & 7u;
  uint32_t pm = probeMask<true>(dir, mask, mb, c, x, y, z, ta, wa, cb, &vc, &im);
  uint32_t n = 0;
  while (pm && n < limit) {
    const uint32_t k = (uint32_t)__builtin_ctz(pm);
    pm &= pm - 1;
    int32_t bx, by, bz;
    probeB(cb, k, bx, by, bz);
    V3 wb = xformBy(c.xb, c.pb, voxelToLocal(c.vsb, bx, by, bz));
    V3 diff = sub(wb, wa);
    float dsq = dot(diff, diff);
    if (!(dsq < c.spec)) continue;
    // the penetration of generateContact (src/narrow_phase.zig:879-906)
    float dist = sqrtf(dsq);
    float pen = c.vsa - dist;
    if (ta == VXP_VOXEL_INSIDE || ((im >> k) & 1u)) pen = fmaxf(pen, c.vsa);
    uint64_t fid = featureId(x, y, z, bx, by, bz);
    Cand& o = out[n];
    o.pen = pen;
    o.fid_lo = (uint32_t)fid;
    o.fid_hi = (uint32_t)(fid >> 32);
    o.code = (lane << 5) | k;
    n++;
  }
}

// generateContact (src/narrow_phase.zig:860-922) for A voxel (x, y, z) against
// the B voxel of probe `probe` around it, with the same arithmetic as the probe.
__device__ vxp_contact contactAt(const DevColliders& C, const DevCollider& HA,
                                 const DevCollider& HB, const ProbeCtx& c, int32_t x, int32_t y,
                                 int32_t z, uint32_t probe) {
  const uint8_t va = voxelAt(C, HA, x, y, z);
  V3 wa = xformBy(c.xa, c.pa, voxelToLocal(c.vsa, x, y, z));
  V3 lb = invXformBy(c.ib, c.pb, wa);
  int32_t cb[3];
  localToVoxel(c.vsb, lb, cb);
  const int32_t dz = (int32_t)(probe / 9) - 1, dy = (int32_t)(probe / 3 % 3) - 1,
                dx = (int32_t)(probe % 3) - 1;
  const int32_t bx = cb[0] + dx, by = cb[1] + dy, bz = cb[2] + dz;
  const uint8_t vb = voxelAt(C, HB, bx, by, bz);
  const uint32_t ta = va & 7u, tb = vb & 7u;
  V3 wb = xformBy(c.xb, c.pb, voxelToLocal(c.vsb, bx, by, bz));
  V3 diff = sub(wb, wa);
  float dsq = dot(diff, diff);
  float dist = sqrtf(dsq);
  V3 cp = scale(add(wa, wb), 0.5f);
  float pen = c.vsa - dist;
  V3 na = qrotate(c.qa, lutNormal(va >> 3));
  V3 nb = qrotate(c.qb, lutNormal(vb >> 3));
  V3 nrm;
  if (ta == VXP_VOXEL_FACE) {
    nrm = na;
  } else if (tb == VXP_VOXEL_FACE) {
    nrm = neg(nb);
  } else if (dist > 0.001f) {
    V3 dn = scale(diff, 1.0f / dist);
    if (dot(dn, na) > 0.3f) nrm = na;
    else if (dot(dn, neg(nb)) > 0.3f) nrm = neg(nb);
    else nrm = dn;
  } else {
    nrm = na;
  }
  if (ta == VXP_VOXEL_INSIDE || tb == VXP_VOXEL_INSIDE) pen = fmaxf(pen, c.vsa);
  V3 la = invXformBy(c.ia, c.pa, cp);
  V3 lbb = invXformBy(c.ib, c.pb, cp);
  vxp_contact o;
  o.point[0] = cp.x; o.point[1] = cp.y; o.point[2] = cp.z;
  o.normal[0] = nrm.x; o.normal[1] = nrm.y; o.normal[2] = nrm.z;
  o.penetration = pen;
  o.local_anchor_a[0] = la.x; o.local_anchor_a[1] = la.y; o.local_anchor_a[2] = la.z;
  o.local_anchor_b[0] = lbb.x; o.local_anchor_b[1] = lbb.y; o.local_anchor_b[2] = lbb.z;
  o.normal_impulse = 0.0f;
  o.tangent_impulse[0] = o.tangent_impulse[1] = 0.0f;
  o.normal_mass = 0.0f;
  o.tangent_mass[0] = o.tangent_mass[1] = 0.0f;
  o.pad = 0;
  o.feature_id = featureId(x, y, z, bx, by, bz);
  return o;
}

// Probe context of pair r (body poses, voxel sizes, B's voxel bounds).
__device__ __forceinline__ ProbeCtx probeCtx(const DevBodies& B, const DevCollider& HA,
                                             const DevCollider& HB, uint32_t ga, uint32_t gb) {
  ProbeCtx c;
  c.pa = f4v(B.pos[ga]);
  c.pb = f4v(B.pos[gb]);
  c.qa = f4q(B.rot[ga]);
  c.qb = f4q(B.rot[gb]);
  c.vsa = HA.vs;
  c.vsb = HB.vs;
  float cds = c.vsa * c.vsa * 1.0f * 1.0f;
  c.spec = cds + 0.02f * 0.02f + 2.0f * sqrtf(cds) * 0.02f;
  for (int k = 0; k < 3; k++) {
    c.bminb[k] = HB.minb[k];
    c.bmaxb[k] = HB.maxb[k];
  }
  ctxRots(c);
  return c;
}

// Region voxel `lin` (z, y, x order) of pair r.
__device__ __forceinline__ void regionVoxel(const PairRec& r, uint32_t lin, int32_t& x, int32_t& y,
                                            int32_t& z) {
  uint32_t nxy = r.nx * r.ny;
  x = r.mn[0] + (int32_t)(lin % r.nx);
  y = r.mn[1] + (int32_t)((lin / r.nx) % r.ny);
  z = r.mn[2] + (int32_t)(lin / nxy);
}

// Two passes over the same work items.  kCount: per-lane candidate counts
// (lane_n), per-work-item totals wi_ccnt (wi_ccnt[total_wi] = 0) and
// voxels_checked sums.  Otherwise: candidates written from wi_cbase[wid] (the
// exclusive scan of wi_ccnt) in (lane, probe) order.
// Workgroups go to the 8 XCDs round-robin; this maps them so that runs of
// `run` consecutive logical blocks execute on one XCD (consecutive work items
// of a pair then share that XCD's L2) while the runs still rotate over the
// XCDs.  A bijection on [0, gridDim.x) for any run >= 1: blocks past the last
// whole round of runs keep their index.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcdBlock(uint32_t run) {
  const uint32_t b = blockIdx.x, full = gridDim.x / (kXcds * run) * (kXcds * run);
  if (b >= full) return b;
  const uint32_t x = b % kXcds, r = b / kXcds;
  return (r / run) * (kXcds * run) + x * run + r % run;
}

// A load of launch-invariant data through the scalar cache: the address is
// wave-uniform and nothing in the launch writes it.
template <class T>
__device__ __forceinline__ T sload(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  typedef const __attribute__((address_space(4))) uint32_t* CW;
  const CW w = (CW)(p);
  uint32_t v[sizeof(T) / 4];
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; i++) v[i] = w[i];
  T t;
  __builtin_memcpy(&t, v, sizeof(T));
  return t;
}

// A pair's probe context, loaded through the scalar cache by the waves that
// take its work items.
struct ProbePair {
  uint32_t woff, nx, ny, n;  // first work item, region extent, region voxels
  uint32_t nxy, mx, mxy;     // nx * ny; division multipliers of nx and nxy (0: divide)
  int32_t mn[3];
  DevCollider HA, HB;
  ProbeCtx c;
};
// Multiplier of the exact division by d of every n < count: floor(n / d) =
// umulhi(n, m) with m = ceil(2^32 / d) when (count - 1) * d <= 2^32 (with
// e = m d - 2^32 < d, n m / 2^32 = n / d + n e / (d 2^32) and n e < 2^32
// keeps the quotient's floor).  0 where that does not hold or d = 1.
__device__ __forceinline__ uint32_t divMagic(uint32_t d, uint32_t count) {
  if (d <= 1u || (uint64_t)(count ? count - 1u : 0u) * d > (1ull << 32)) return 0u;
  return (uint32_t)(((1ull << 32) + d - 1u) / d);
}
__device__ __forceinline__ uint32_t divBy(uint32_t n, uint32_t d, uint32_t m) {
  return m ? __umulhi(n, m) : n / d;
}
__device__ __forceinline__ void loadProbePair(ProbePair& q, const DevCfg& cfg, const DevBodies& B,
                                              const DevColliders& C, const PairRec* pr,
                                              const uint32_t* wi_off, uint32_t p) {
  const PairRec r = sload(pr + p);
  q.woff = sload(wi_off + p);
  q.nx = r.nx;
  q.ny = r.ny;
  q.n = r.nx * r.ny * r.nz;
  q.nxy = r.nx * r.ny;
  q.mx = divMagic(q.nx, q.n);
  q.mxy = divMagic(q.nxy, q.n);
  for (int k = 0; k < 3; k++) q.mn[k] = r.mn[k];
  const uint32_t ga = r.world * cfg.max_bodies + r.sa, gb = r.world * cfg.max_bodies + r.sb;
  q.HA = sload(C.hdr + sload(B.collider + ga));
  q.HB = sload(C.hdr + sload(B.collider + gb));
  ProbeCtx& c = q.c;
  c.pa = f4v(sload(B.pos + ga));
  c.pb = f4v(sload(B.pos + gb));
  c.qa = f4q(sload(B.rot + ga));
  c.qb = f4q(sload(B.rot + gb));
  c.vsa = q.HA.vs;
  c.vsb = q.HB.vs;
  float cds = c.vsa * c.vsa * 1.0f * 1.0f;
  c.spec = cds + 0.02f * 0.02f + 2.0f * sqrtf(cds) * 0.02f;
  for (int k = 0; k < 3; k++) {
    c.bminb[k] = q.HB.minb[k];
    c.bmaxb[k] = q.HB.maxb[k];
  }
  ctxRots(c);
}
// Region voxel of lane `lane` of work item `wid` (z, y, x order); false past
// the region's end.
__device__ __forceinline__ bool probeVoxelXYZ(const ProbePair& q, uint32_t wid, uint32_t lane,
                                              int32_t& x, int32_t& y, int32_t& z) {
  const uint32_t lin = (wid - q.woff) * kWave + lane;
  if (lin >= q.n) return false;
  const uint32_t r = divBy(lin, q.nx, q.mx), zz = divBy(lin, q.nxy, q.mxy);
  x = q.mn[0] + (int32_t)(lin - r * q.nx);
  y = q.mn[1] + (int32_t)(r - zz * q.ny);
  z = q.mn[2] + (int32_t)zz;
  return true;
}

// The waves' share of the work items: contiguous ranges, so a pair's
// consecutive chunks reuse its context; runs of `xcd_run` blocks stay on one
// XCD, so neighbouring ranges share that XCD's L2.
__device__ __forceinline__ void probeRange(uint32_t total_wi, uint32_t xcd_run, uint32_t& w_lo,
                                           uint32_t& w_hi) {
  const uint32_t wpb = blockDim.x / kWave, nwaves = gridDim.x * wpb;
  const uint32_t g =
      __builtin_amdgcn_readfirstlane(xcdBlock(xcd_run) * wpb + threadIdx.x / kWave);
  const uint32_t per = (total_wi + nwaves - 1) / nwaves;
  w_lo = min(g * per, total_wi);
  w_hi = min(w_lo + per, total_wi);
}

constexpr uint32_t kProbeLdsWords = 15 * kWave;  // count pass LDS per wave (3.75 KiB)

// Count pass: per-lane candidate counts (lane_n), per-work-item totals
// wi_ccnt (zeroed by k_wi_fill), each pair's voxels_checked added to p_vchk
// (zeroed by k_pair_process),
// wi_ccnt[total_wi] = 0, and the lanes with candidates (work item * 64 +
// lane) on lane_list.  Most region voxels of A are solid but a work item's
// lanes past the region or on empty voxels would idle, so the lanes are
// compacted before the neighbourhood test: each wave queues its non-empty A
// voxels (work item, lane, type) in LDS and runs probeCount on 64 of them at a
// time (fewer at a pair's end), adding each work item's sums with one atomic
// per work item and batch.
__global__ __launch_bounds__(256, 5) void k_probe_count(DevCfg cfg, DevBodies B, DevColliders C,
                                                     const PairRec* pr, const uint32_t* wi_pair,
                                                     const uint32_t* wi_off, uint32_t total_wi,
                                                     uint32_t* wi_ccnt, uint32_t* p_vchk,
                                                     uint8_t* lane_n, uint32_t* lane_list,
                                                     uint32_t* list_n, uint32_t xcd_run,
                                                     const uint32_t* n_dev) {
  // per wave in dynamic LDS (kProbeLdsWords words): the queue of non-empty A
  // voxels (work item, lane | type << 8, and x, y, z), the lanes with
  // candidates, and the queue of voxels whose window lies in B's bounds (work
  // item, lane | type, B voxel cb, world position wa)
  extern __shared__ uint32_t probe_lds[];
  const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  uint32_t* const qw = probe_lds + wv * kProbeLdsWords;
  uint32_t* const ql = qw + kWave;
  uint32_t* const le = ql + kWave;            // 2 * kWave
  uint32_t* const rw = le + 2 * kWave;
  uint32_t* const rl = rw + kWave;
  uint32_t* const rcb = rl + kWave;           // 3 * kWave
  float* const rwa = reinterpret_cast<float*>(rcb + 3 * kWave);  // 3 * kWave
  uint32_t* const qxyz = reinterpret_cast<uint32_t*>(rwa + 3 * kWave);  // 3 * kWave: x, y, z
  // the lanes with candidates gather in le and go to lane_list once 64 are
  // there (or at the end), with one returning atomic per flush
  uint32_t ln = 0;
  auto listFlush = [&]() {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(list_n, ln);
    base = rdlane(base, 0);
    for (uint32_t i = lane; i < ln; i += kWave) lane_list[base + i] = le[i];
    ln = 0;
  };
  total_wi = devCount(total_wi, n_dev);
  uint32_t w_lo, w_hi;
  probeRange(total_wi, xcd_run, w_lo, w_hi);
  if (lane == 0 && w_lo < w_hi && w_hi == total_wi) wi_ccnt[total_wi] = 0;
  ProbePair q{};
  uint32_t pcur = 0xFFFFFFFFu, qn = 0, rn = 0;
  // stage 2: the neighbourhood test of the in-window entries [0, cnt)
  auto flush2 = [&](uint32_t cnt) {
    const bool act = lane < cnt;
    uint32_t ew = 0, el = 0, n = 0, vc = 0;
    if (act) {
      ew = rw[lane];
      el = rl[lane];
      const int32_t cb[3] = {(int32_t)rcb[lane], (int32_t)rcb[kWave + lane],
                             (int32_t)rcb[2 * kWave + lane]};
      const V3 wa{rwa[lane], rwa[kWave + lane], rwa[2 * kWave + lane]};
      n = probeCountAt(C.dir, C.mask, voxMap(q.HB), q.c, wa, cb, (uint8_t)(el >> 8), &vc);
      lane_n[(size_t)ew * kWave + (el & (kWave - 1))] = (uint8_t)n;
    }
    const uint64_t hm = __ballot(n != 0u);
    if (hm) {
      if (n) le[ln + (uint32_t)__popcll(hm & ((1ull << lane) - 1ull))] = ew * kWave + (el & (kWave - 1));
      ln += (uint32_t)__popcll(hm);
      if (ln >= kWave) listFlush();
    }
    // entries are in work-item order: one candidate sum per work item of the
    // batch; the batch is one pair's, so its checked voxels are one sum
    uint64_t rem = __ballot(act);
    uint32_t bv = 0;
    while (rem) {
      const int f = __ffsll((long long)rem) - 1;
      const uint32_t w0 = rdlane(ew, f);
      const bool mine = act && ew == w0;
      // a work item's sums are below 64 * 27: both in one reduction
      const uint32_t s2 = waveSum(mine ? (n << 16) | vc : 0u);
      const uint32_t sn = s2 >> 16;
      bv += s2 & 0xFFFFu;
      if ((int)lane == f && sn) atomicAdd(wi_ccnt + w0, sn);
      rem &= ~__ballot(mine);
    }
    if (lane == 0 && bv) atomicAdd(p_vchk + pcur, bv);
  };
  // stage 1: the window of the queued non-empty voxels [0, cnt); those with
  // no probe in B's bounds count nothing, the others move to stage 2
  auto flush1 = [&](uint32_t cnt) {
    const bool act = lane < cnt;
    const uint32_t ew = act ? qw[lane] : 0u, el = act ? ql[lane] : 0u;
    bool inb = false;
    V3 wa{0.f, 0.f, 0.f};
    int32_t cb[3] = {0, 0, 0};
    if (act) {
      const int32_t x = (int32_t)qxyz[lane], y = (int32_t)qxyz[kWave + lane],
                    z = (int32_t)qxyz[2 * kWave + lane];
      inb = probeWindow(q.c, x, y, z, wa, cb);
      if (!inb) lane_n[(size_t)ew * kWave + (el & (kWave - 1))] = 0;
    }
    const uint64_t m = __ballot(inb);
    if (!m) return;
    const uint32_t k = (uint32_t)__popcll(m);
    const uint32_t slot = rn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    auto put = [&](uint32_t at) {
      rw[at] = ew;
      rl[at] = el;
      for (int j = 0; j < 3; j++) rcb[j * kWave + at] = (uint32_t)cb[j];
      rwa[at] = wa.x;
      rwa[kWave + at] = wa.y;
      rwa[2 * kWave + at] = wa.z;
    };
    if (inb && slot < kWave) put(slot);
    if (rn + k < kWave) {
      rn += k;
      return;
    }
    flush2(kWave);
    if (inb && slot >= kWave) put(slot - kWave);
    rn = rn + k - kWave;
  };
  for (uint32_t wid = w_lo; wid < w_hi; wid++) {
    const uint32_t p = sload(wi_pair + wid);
    if (p != pcur) {
      if (qn) flush1(qn);
      if (rn) flush2(rn);
      qn = rn = 0;
      pcur = p;
      loadProbePair(q, cfg, B, C, pr, wi_off, p);
    }
    int32_t x, y, z;
    uint8_t va = 0;
    if (probeVoxelXYZ(q, wid, lane, x, y, z)) va = voxelAt(C, q.HA, x, y, z);
    const bool ne = (va & 7u) != 0;
    if (!ne) lane_n[(size_t)wid * kWave + lane] = 0;
    const uint64_t m = __ballot(ne);
    if (!m) continue;
    const uint32_t k = (uint32_t)__popcll(m);
    const uint32_t slot = qn + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    const uint32_t ent = lane | ((uint32_t)va << 8);
    auto queue = [&](uint32_t at) {
      qw[at] = wid;
      ql[at] = ent;
      qxyz[at] = (uint32_t)x;
      qxyz[kWave + at] = (uint32_t)y;
      qxyz[2 * kWave + at] = (uint32_t)z;
    };
    if (ne && slot < kWave) queue(slot);
    if (qn + k < kWave) {
      qn += k;
      continue;
    }
    flush1(kWave);
    if (ne && slot >= kWave) queue(slot - kWave);
    qn = qn + k - kWave;
  }
  if (qn) flush1(qn);
  if (rn) flush2(rn);
  if (ln) listFlush();
}

// Write pass: one thread per listed lane (work item * 64 + lane, a lane with
// candidates): its pair's context, its offset wi_cbase[wid] plus the counts
// of the work item's lower lanes, and its candidates in probe order.
__global__ __launch_bounds__(256) void k_probe_write(DevCfg cfg, DevBodies B, DevColliders C,
                                                     const PairRec* pr, const uint32_t* wi_pair,
                                                     const uint32_t* wi_off,
                                                     const uint32_t* lane_list, uint32_t n_items,
                                                     const uint32_t* wi_cbase,
                                                     const uint8_t* lane_n, Cand* cand,
                                                     const uint32_t* n_dev) {
  n_items = devCount(n_items, n_dev);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items;
       i += gridDim.x * blockDim.x) {
    const uint32_t e = lane_list[i], wid = e / kWave, lane = e % kWave;
    // counts of the lanes below this one: bytes of the work item's 64 count
    // bytes, four per word (sums stay below 2^8 per word: 4 * 27)
    const uint4* row = reinterpret_cast<const uint4*>(lane_n + (size_t)wid * kWave);
    uint32_t below = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      const uint4 v = row[k];
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (uint32_t j = 0; j < 4; j++) {
        const uint32_t first = 16 * k + 4 * j;  // lane of the word's low byte
        const uint32_t keep = lane <= first ? 0u : lane - first >= 4 ? 4u : lane - first;
        const uint32_t m = keep == 4u ? 0xFFFFFFFFu : (1u << (8 * keep)) - 1u;
        below += ((w[j] & m) * 0x01010101u) >> 24;
      }
    }
    const uint32_t n = lane_n[e];
    ProbePair q;
    loadProbePair(q, cfg, B, C, pr, wi_off, wi_pair[wid]);
    int32_t x, y, z;
    probeVoxelXYZ(q, wid, lane, x, y, z);
    const uint8_t va = voxelAt(C, q.HA, x, y, z);
    probeWrite(C.dir, C.mask, voxMap(q.HB), q.c, x, y, z, va, lane,
               cand + wi_cbase[wid] + below, n);
  }
}

// ---------------------------------------------------------------------------
// manifold reduction: ContactManifold.addContact replay over the candidates
// in sequence order, warm start against the previous frame, cache hit/carry.
// ---------------------------------------------------------------------------

// ContactManager.addManifold warm start (src/narrow_phase.zig:1011-1021,1049-1059)
__device__ void warmStartFrom(vxp_manifold& m, const DevManifold* prev, int32_t j) {
  if (j < 0) return;
  const vxp_manifold& o = prev[j].m;
  for (uint32_t i = 0; i < m.contact_count; i++)
    for (uint32_t k = 0; k < o.contact_count; k++)
      if (m.contacts[i].feature_id == o.contacts[k].feature_id) {
        m.contacts[i].normal_impulse = o.contacts[k].normal_impulse;
        m.contacts[i].tangent_impulse[0] = o.contacts[k].tangent_impulse[0];
        m.contacts[i].tangent_impulse[1] = o.contacts[k].tangent_impulse[1];
        break;
      }
}


// One wave per pair.  The streaming top-4 of ContactManifold.addContact
// (src/narrow_phase.zig:451-481) is replayed exactly in candidate order: the
// wave loads 64 candidates at a time and a lane is visited only if it could
// change the slots at its turn (fewer than 4 slots, its feature id is in a
// slot, or some slot is strictly shallower); every other candidate leaves the
// state unchanged in the sequential order too.  The slots are wave-uniform.
// Per-pair counters go to rstat and are summed by k_reduce_tail.
__device__ __forceinline__ void reducePairWave(const DevCfg& cfg, const DevBodies& B,
                                               const DevColliders& C,
                                               const WorldState* ws, const PairRec* pr,
                                               const uint64_t* keys, uint32_t p, uint32_t f,
                                               const uint32_t* wi_off, const uint32_t* wi_cbase,
                                               const uint32_t* wi_ccnt, const uint32_t* p_vchk,
                                               const Cand* cand, const DevManifold* prev,
                                               const CacheVal* cvals, CacheVal* cstage,
                                               DevManifold* staging, uint32_t* produce,
                                               RStat* rstat) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t aux = pr[p].aux;
  if (f == PF_CARRY) {
    const uint2* src = reinterpret_cast<const uint2*>(prev + aux);
    uint2* dst = reinterpret_cast<uint2*>(staging + p);
    for (uint32_t i = lane; i < sizeof(DevManifold) / 8; i += kWave) dst[i] = src[i];
    if (lane == 0) {
      produce[p] = 1;
      rstat[p] = RStat{1u, 0u, 0u, 0u, 0u, 0u};
    }
    return;
  }
  const uint32_t w = pr[p].world, sa = pr[p].sa, sb = pr[p].sb;
  const uint32_t ga = w * cfg.max_bodies + sa, gb = w * cfg.max_bodies + sb;
  uint32_t ncont = 0, vchk = 0, ncand = 0;
  vxp_contact mine;  // contact `lane` of the manifold (lanes 0..3)
  if (f == PF_HIT) {
    const CacheVal& v = cvals[(size_t)w * cfg.cache_cap + aux];
    if (!v.has_collision) {
      if (lane == 0) {
        produce[p] = 0;
        rstat[p] = RStat{0u, 0u, 0u, 0u, 0u, 0u};
      }
      return;
    }
    ncont = v.count;
    vchk = v.vchk;
    if (lane < ncont) mine = v.c[lane];
  } else {
    const uint32_t w0 = wi_off[p], w1 = w0 + pr[p].nchunks;
    vchk = p_vchk[p];
    // no candidates (the work items' counts sum to the difference of their
    // scanned offsets): no contact, nothing to reduce
    const bool none = wi_cbase[w1] == wi_cbase[w0];
    uint32_t sidx0 = 0, sidx1 = 0, sidx2 = 0, sidx3 = 0;
    uint32_t swi0 = 0, swi1 = 0, swi2 = 0, swi3 = 0;  // work item of each slot's candidate
    float spen0 = 0.f, spen1 = 0.f, spen2 = 0.f, spen3 = 0.f;
    uint64_t sfid0 = 0, sfid1 = 0, sfid2 = 0, sfid3 = 0;
    uint32_t ns = 0, ncl = 0;
    // work items 64 at a time: lanes read the counts, and only the non-empty
    // items are visited, in ascending order
    for (uint32_t wb = w0; !none && wb < w1; wb += kWave) {
      const uint32_t wl = wb + lane;
      const uint32_t cl = wl < w1 ? wi_ccnt[wl] : 0u, bl = wl < w1 ? wi_cbase[wl] : 0u;
      ncl += cl;
      uint64_t nz = __ballot(cl != 0u);
      while (nz) {
      const int kw = __ffsll((long long)nz) - 1;
      nz &= nz - 1;
      const uint32_t wi = wb + (uint32_t)kw;
      const uint32_t cnt = rdlane(cl, kw), base = rdlane(bl, kw);
      for (uint32_t off = 0; off < cnt; off += kWave) {
        const bool valid = off + lane < cnt;
        float pen = 0.0f;
        uint64_t fid = 0;
        if (valid) {
          const Cand& c = cand[base + off + lane];
          pen = c.pen;
          fid = ((uint64_t)c.fid_hi << 32) | c.fid_lo;
        }
        int32_t last = -1;
        while (true) {
          bool cand_lane = valid && (int32_t)lane > last;
          if (cand_lane && ns == 4) {
            bool same = fid == sfid0 || fid == sfid1 || fid == sfid2 || fid == sfid3;
            bool deeper = spen0 < pen || spen1 < pen || spen2 < pen || spen3 < pen;
            cand_lane = same || deeper;
          }
          const uint64_t mask = __ballot(cand_lane);
          if (!mask) break;
          const int k = __ffsll((long long)mask) - 1;
          const float pk = rdlane(pen, k);
          const uint64_t fk = ((uint64_t)rdlane((uint32_t)(fid >> 32), k) << 32) |
                              rdlane((uint32_t)fid, k);
          const uint32_t ik = base + off + (uint32_t)k;
          last = k;
          // addContact on the uniform slots
          if (ns > 0 && sfid0 == fk) { sidx0 = ik; swi0 = wi; spen0 = pk; continue; }
          if (ns > 1 && sfid1 == fk) { sidx1 = ik; swi1 = wi; spen1 = pk; continue; }
          if (ns > 2 && sfid2 == fk) { sidx2 = ik; swi2 = wi; spen2 = pk; continue; }
          if (ns > 3 && sfid3 == fk) { sidx3 = ik; swi3 = wi; spen3 = pk; continue; }
          if (ns < 4) {
            if (ns == 0) { sidx0 = ik; swi0 = wi; spen0 = pk; sfid0 = fk; }
            else if (ns == 1) { sidx1 = ik; swi1 = wi; spen1 = pk; sfid1 = fk; }
            else if (ns == 2) { sidx2 = ik; swi2 = wi; spen2 = pk; sfid2 = fk; }
            else { sidx3 = ik; swi3 = wi; spen3 = pk; sfid3 = fk; }
            ns++;
            continue;
          }
          float minp = pk;
          int mi = -1;
          if (spen0 < minp) { minp = spen0; mi = 0; }
          if (spen1 < minp) { minp = spen1; mi = 1; }
          if (spen2 < minp) { minp = spen2; mi = 2; }
          if (spen3 < minp) { minp = spen3; mi = 3; }
          if (mi == 0) { sidx0 = ik; swi0 = wi; spen0 = pk; sfid0 = fk; }
          else if (mi == 1) { sidx1 = ik; swi1 = wi; spen1 = pk; sfid1 = fk; }
          else if (mi == 2) { sidx2 = ik; swi2 = wi; spen2 = pk; sfid2 = fk; }
          else if (mi == 3) { sidx3 = ik; swi3 = wi; spen3 = pk; sfid3 = fk; }
        }
      }
      }
    }
    ncand = waveSum(ncl);
    ncont = ns;
    if (lane < ncont) {
      // rebuild the winner's contact from its A voxel (lane of its work item)
      // and B neighbour (probe index)
      const uint32_t my = lane == 0 ? sidx0 : lane == 1 ? sidx1 : lane == 2 ? sidx2 : sidx3;
      const uint32_t mwi = lane == 0 ? swi0 : lane == 1 ? swi1 : lane == 2 ? swi2 : swi3;
      const uint32_t code = cand[my].code;
      const DevCollider& HA = C.hdr[B.collider[ga]];
      const DevCollider& HB = C.hdr[B.collider[gb]];
      const ProbeCtx c = probeCtx(B, HA, HB, ga, gb);
      int32_t x, y, z;
      regionVoxel(pr[p], (mwi - w0) * kWave + (code >> 5), x, y, z);
      mine = contactAt(C, HA, HB, c, x, y, z, code & 31u);
    }
    if (f == PF_MISS) {
      // cached value: contacts before warm starting (src/narrow_phase.zig:1501-1509)
      CacheVal& cv = cstage[p];
      if (lane < ncont) cv.c[lane] = mine;
      if (lane == 0) {
        cv.count = ncont;
        cv.has_collision = ncont > 0;
        cv.vchk = vchk;
        cv.pad = 0;
      }
    }
  }
  if (ncont == 0) {
    if (lane == 0) {
      produce[p] = 0;
      rstat[p] = RStat{0u, vchk, 0u, 0u, ncand, 0u};
    }
    return;
  }
  // ContactManager.addManifold warm start (src/narrow_phase.zig:1011-1021,1049-1059)
  const uint64_t key = keys[p];
  const int32_t j = findPrev(prev, ws[w].prev_start, ws[w].prev_start + ws[w].prev_count, key);
  if (j >= 0 && lane < ncont) {
    const vxp_manifold& o = prev[j].m;
    for (uint32_t k = 0; k < o.contact_count; k++)
      if (mine.feature_id == o.contacts[k].feature_id) {
        mine.normal_impulse = o.contacts[k].normal_impulse;
        mine.tangent_impulse[0] = o.contacts[k].tangent_impulse[0];
        mine.tangent_impulse[1] = o.contacts[k].tangent_impulse[1];
        break;
      }
  }
  DevManifold& d = staging[p];
  if (lane < 4) {
    if (lane >= ncont) memset(&mine, 0, sizeof(mine));
    d.m.contacts[lane] = mine;
  }
  if (lane == 0) {
    d.key = key;
    d.sa = sa;
    d.sb = sb;
    d.m.body_a_id = pr[p].ida;
    d.m.body_b_id = pr[p].idb;
    d.m.friction = sqrtf(B.fric[ga] * B.fric[gb]);
    d.m.restitution = fmaxf(B.rest[ga], B.rest[gb]);
    d.m.contact_count = ncont;
    d.m.color = kNone;
    produce[p] = 1;
    rstat[p] = RStat{1u, vchk, 1u, ncont, ncand, 0u};
  }
}

// The work list (built by k_pair_process; pairs not on it were cleared
// there), one thread per pair: a pair whose outcome needs no reduction is
// settled here — a detection or cache miss without candidates (the sum of
// its work items' counts, a difference of their scanned offsets, is zero),
// a cache hit without collision — and the others are listed for k_reduce's
// waves (one append per block).
__global__ __launch_bounds__(256) void k_reduce_settle(DevCfg cfg, const PairRec* pr,
                                                       const uint32_t* wlist, uint32_t* step_ctr,
                                                       const uint32_t* wi_off,
                                                       const uint32_t* wi_cbase,
                                                       const uint32_t* p_vchk,
                                                       const CacheVal* cvals, CacheVal* cstage,
                                                       uint32_t* produce, RStat* rstat,
                                                       uint32_t* coop, uint32_t cap) {
  __shared__ uint32_t s_n, s_base;
  const uint32_t total = min(step_ctr[kCtrReduceWork], cap);
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    uint32_t p = 0, pos = 0;
    bool keep = false;
    if (i < total) {
      p = wlist[i];
      const PairRec& r = pr[p];
      const uint32_t f = r.flag;
      if (f == PF_DETECT || f == PF_MISS) {
        const uint32_t w0 = wi_off[p], w1 = w0 + r.nchunks;
        if (wi_cbase[w1] == wi_cbase[w0]) {
          const uint32_t vchk = p_vchk[p];
          produce[p] = 0;
          rstat[p] = RStat{0u, vchk, 0u, 0u, 0u, 0u};
          if (f == PF_MISS) {  // the cached value of a miss without contacts
            CacheVal& cv = cstage[p];
            cv.count = 0;
            cv.has_collision = 0;
            cv.vchk = vchk;
            cv.pad = 0;
          }
        } else {
          keep = true;
        }
      } else if (f == PF_HIT) {
        if (!cvals[(size_t)r.world * cfg.cache_cap + r.aux].has_collision) {
          produce[p] = 0;
          rstat[p] = RStat{0u, 0u, 0u, 0u, 0u, 0u};
        } else {
          keep = true;
        }
      } else {
        keep = true;
      }
      if (keep) pos = atomicAdd(&s_n, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_n ? atomicAdd(&step_ctr[kCtrReduceCoop], s_n) : 0u;
    __syncthreads();
    if (keep) coop[s_base + pos] = p;
    __syncthreads();  // s_n and s_base are reused by the next round
  }
}

// Waves take the listed pairs one at a time in a grid-stride loop.
__global__ __launch_bounds__(256) void k_reduce(DevCfg cfg, DevBodies B, DevColliders C,
                                                const WorldState* ws,
                                                const PairRec* pr, const uint64_t* keys,
                                                const uint32_t* coop, const uint32_t* step_ctr,
                                                const uint32_t* wi_off,
                                                const uint32_t* wi_cbase, const uint32_t* wi_ccnt,
                                                const uint32_t* p_vchk, const Cand* cand,
                                                const DevManifold* prev, const CacheVal* cvals,
                                                CacheVal* cstage, DevManifold* staging,
                                                uint32_t* produce, RStat* rstat) {
  const uint32_t nwaves = gridDim.x * (blockDim.x / kWave);
  const uint32_t total = step_ctr[kCtrReduceCoop];
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < total;
       i += nwaves) {
    const uint32_t p = coop[i];
    reducePairWave(cfg, B, C, ws, pr, keys, p, pr[p].flag, wi_off, wi_cbase, wi_ccnt, p_vchk, cand,
                   prev, cvals, cstage, staging, produce, rstat);
  }
}

// BrickPairCache.put of this frame's misses (entry owner = the pair that
// reserved it during the replay), and the per-world narrow-phase counters.
__global__ void k_reduce_tail(DevCfg cfg, WorldState* ws, vxp_stats* st, const PairRec* pr,
                              uint32_t n_pairs, const uint32_t* wlist, const uint32_t* step_ctr,
                              const CacheKeyRec* keys, CacheVal* vals, const CacheVal* cstage,
                              const RStat* rstat) {
  // the reduction's work list (the pairs with a record and counters)
  const uint32_t n = min(step_ctr[kCtrReduceWork], n_pairs);
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
  const uint32_t i = base + threadIdx.x;
  RStat c{0u, 0u, 0u, 0u, 0u, 0u};
  uint32_t w = 0;
  if (i < n) {
    const uint32_t p = wlist[i];
    const PairRec& r = pr[p];
    w = r.world;
    c = rstat[p];
    if (r.flag == PF_MISS) {
      size_t e = (size_t)w * cfg.cache_cap + r.aux;
      if (keys[e].owner_frame == ws[w].frame && keys[e].owner_pair == p) vals[e] = cstage[p];
    }
  }
  const uint32_t v0[1] = {c.produced};
  uint32_t* const c0[1] = {&ws->produced};
  blockAggAdd<1>(w, v0, c0, sizeof(WorldState) / 4);
  __syncthreads();  // blockAggAdd's LDS slots are reused
  const uint32_t v1[4] = {c.vchk, c.colliding, c.contacts, c.cands};
  uint32_t* const c1[4] = {&st->voxels_checked, &st->narrow_phase_colliding,
                           &st->contacts_generated, &st->contact_candidates};
  blockAggAdd<4>(w, v1, c1, sizeof(vxp_stats) / 4);
  __syncthreads();
  }
}

// New manifold segments: skipped worlds keep their previous segment.
__global__ void k_world_segments(WorldState* ws, uint32_t n_worlds, uint32_t* total) {
  // single block: exclusive scan of per-world counts; total[1] = largest count
  __shared__ uint32_t carry, largest;
  __shared__ uint32_t buf[1024];
  if (threadIdx.x == 0) carry = largest = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n_worlds; base += blockDim.x) {
    uint32_t w = base + threadIdx.x;
    uint32_t c = 0;
    if (w < n_worlds) c = ws[w].skip ? ws[w].prev_count : ws[w].produced;
    buf[threadIdx.x] = c;
    __syncthreads();
    for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
      uint32_t t = threadIdx.x >= d ? buf[threadIdx.x - d] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (w < n_worlds) {
      ws[w].man_start = carry + buf[threadIdx.x] - c;
      ws[w].man_count = c;
      if (c) atomicMax(&largest, c);
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += buf[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    total[0] = carry;
    total[1] = largest;
  }
}

__global__ void k_scatter(const WorldState* ws, const PairRec* pr, uint32_t n_pairs,
                          const uint32_t* produce, const uint32_t* pidx,
                          const DevManifold* staging, DevManifold* out, const uint32_t* n_dev) {
  n_pairs = devCount(n_pairs, n_dev);
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    if (!produce[p]) continue;
    uint32_t w = pr[p].world;
    uint32_t dst = ws[w].man_start + pidx[p] - pidx[ws[w].pair_lo];
    out[dst] = staging[p];
  }
}

__global__ void k_copy_skipped(const WorldState* ws, const DevManifold* prev, DevManifold* out) {
  uint32_t w = blockIdx.x;
  if (!ws[w].skip) return;
  for (uint32_t i = threadIdx.x; i < ws[w].prev_count; i += blockDim.x)
    out[ws[w].man_start + i] = prev[ws[w].prev_start + i];
}

// ---------------------------------------------------------------------------
// islands: union of manifold endpoints across edges with no static body;
// label = minimum body id of the component (src/islands.zig:318-441).
// One block per world; hooking + pointer jumping on id-indexed parents.
// ---------------------------------------------------------------------------
// Root of r in a union-find forest whose parents only ever point to smaller
// indices (hooking to the smaller root), read while other threads hook.
__device__ __forceinline__ uint32_t findRoot(uint32_t* P, uint32_t r) {
  while (true) {
    uint32_t t = __atomic_load_n(&P[r], __ATOMIC_RELAXED);
    if (t == r) return r;
    r = t;
  }
}

// Joins the sets of a and b, the larger root hooked under the smaller, so a
// root is always its set's smallest element and one pass over the edges
// leaves the same roots as hooking rounds run to a fixed point.  Lock-free: a
// failed compare-and-swap means that root was hooked meanwhile; retry from
// the new roots (every failure is some other thread's progress).
__device__ __forceinline__ void unite(uint32_t* P, uint32_t a, uint32_t b) {
  while (true) {
    a = findRoot(P, a);
    b = findRoot(P, b);
    if (a == b) return;
    const uint32_t hi = max(a, b), lo = min(a, b);
    if (atomicCAS(&P[hi], hi, lo) == hi) return;
  }
}

__global__ __launch_bounds__(1024) void k_islands(DevCfg cfg, DevBodies B, WorldState* ws,
                                                  vxp_stats* st, const DevManifold* ml,
                                                  const DevManifold* prev, IslandArrays I,
                                                  size_t lds_bytes) {
  extern __shared__ uint32_t il[];
  uint32_t w = blockIdx.x;
  if (ws[w].skip) return;
  const uint32_t cap = cfg.id_cap, m0 = ws[w].man_start, M = ws[w].man_count;
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  // parent, endpoint flag and component size per body id, and each manifold's
  // union edge (kNone: a static endpoint), in LDS when they fit
  const bool L = ((size_t)3 * cap + 2 * (size_t)M) * 4 <= lds_bytes;
  uint32_t* P = L ? il : I.parent + (size_t)w * cap;
  uint32_t* E = L ? il + cap : I.endpoint + (size_t)w * cap;
  uint32_t* S = L ? il + 2 * cap : I.size + (size_t)w * cap;
  uint32_t* ea = L ? il + 3 * cap : nullptr;
  uint32_t* eb = L ? ea + M : nullptr;
  for (uint32_t i = tid; i < cap; i += bs) {
    P[i] = i;
    E[i] = 0;
    S[i] = 0;
  }
  __shared__ uint32_t differ;
  if (tid == 0) differ = M != ws[w].prev_count;
  __syncthreads();
  for (uint32_t i = tid; i < M; i += bs) {
    const DevManifold& d = ml[m0 + i];
    E[d.m.body_a_id] = 1;
    E[d.m.body_b_id] = 1;
    if (L) {
      const uint32_t ga = w * cfg.max_bodies + d.sa, gb = w * cfg.max_bodies + d.sb;
      const bool st_ = B.type[ga] == VXP_BODY_STATIC || B.type[gb] == VXP_BODY_STATIC;
      ea[i] = d.m.body_a_id;
      eb[i] = st_ ? kNone : d.m.body_b_id;
    }
    // the key sequence against the previous step's (same length only)
    if (i < ws[w].prev_count && d.key != prev[ws[w].prev_start + i].key) atomicOr(&differ, 1u);
  }
  __syncthreads();
  for (uint32_t i = tid; i < M; i += bs) {
    uint32_t ra, rb;
    if (L) {
      if (eb[i] == kNone) continue;
      ra = ea[i];
      rb = eb[i];
    } else {
      const DevManifold& d = ml[m0 + i];
      uint32_t ga = w * cfg.max_bodies + d.sa, gb = w * cfg.max_bodies + d.sb;
      if (B.type[ga] == VXP_BODY_STATIC || B.type[gb] == VXP_BODY_STATIC) continue;
      ra = d.m.body_a_id;
      rb = d.m.body_b_id;
    }
    unite(P, ra, rb);
  }
  __syncthreads();
  for (uint32_t i = tid; i < cap; i += bs) P[i] = findRoot(P, i);  // labels = roots
  __syncthreads();
  __shared__ uint32_t n_isl, largest, n_end, n_sleep, wsum[16], rbase;
  if (tid == 0) {
    n_isl = largest = n_end = n_sleep = rbase = 0;
  }
  __syncthreads();
  for (uint32_t i = tid; i < cap; i += bs)
    if (E[i]) atomicAdd(&S[P[i]], 1u);
  __syncthreads();
  for (uint32_t i = tid; i < cap; i += bs)
    if (S[i]) {
      atomicAdd(&n_isl, 1u);
      atomicAdd(&n_end, S[i]);
      atomicMax(&largest, S[i]);
    }
  // IslandManager rebuild (src/islands.zig:318-441): the graph hash over the
  // manifolds in key order changes exactly when the key sequence differs from
  // the previous step's, which (skips keeping it) is the one last built from
  WorldState& W = ws[w];
  const bool rebuild = W.isl_need || differ || W.isl.total_islands == 0;
  float* T = I.timer + (size_t)w * cap;
  uint32_t* Z = I.sleep + (size_t)w * cap;
  if (rebuild)  // Island.clear of every island (src/islands.zig:176-184)
    for (uint32_t i = tid; i < cap; i += bs) {
      T[i] = 0.0f;
      Z[i] = 0;
    }
  __syncthreads();
  if (tid == 0) {
    if (rebuild) {
      const uint32_t skipped = W.isl.rebuilds_skipped, performed = W.isl.rebuilds_performed + 1;
      W.isl = vxp_island_stats{};
      W.isl.rebuilds_skipped = skipped;
      W.isl.rebuilds_performed = performed;
      if (M) {
        W.isl.total_islands = W.isl.active_islands = n_isl;
        W.isl.largest_island_size = largest;
        W.isl.total_bodies = n_end;
        W.isl.total_contacts = M;
      }
      W.isl_need = 0;
    } else {
      W.isl.rebuilds_skipped++;
    }
    // the system's copy is taken before the sleep update (src/system.zig:739-743)
    st[w].island_count = W.isl.total_islands;
    st[w].active_islands = W.isl.active_islands;
    st[w].sleeping_islands = W.isl.sleeping_islands;
    st[w].largest_island_size = W.isl.largest_island_size;
  }
  // IslandManager.updateSleepStates (src/islands.zig:444-495) over the active
  // bodies copied in slot order (src/system.zig:747-756): island member id k
  // reads the k-th active slot's body.  A moving dynamic body at rank k marks
  // the island of body id k (bit 31 of its size word).
  const float lt = 0.1f * 0.1f, at = 0.1f * 0.1f;
  const uint32_t lane = tid & (kWave - 1), wave = tid / kWave;
  for (uint32_t b0 = 0; b0 < cfg.max_bodies; b0 += bs) {
    const uint32_t slot = b0 + tid;
    const size_t g = (size_t)w * cfg.max_bodies + slot;
    const bool act = slot < cfg.max_bodies && B.active[g];
    const uint64_t bal = __ballot(act);
    if (lane == 0) wsum[wave] = __popcll(bal);
    __syncthreads();
    uint32_t r = rbase + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    for (uint32_t k = 0; k < wave; k++) r += wsum[k];
    if (act && r < cap && E[r] && B.type[g] == VXP_BODY_DYNAMIC) {
      const float4 v = B.lv[g], a = B.av[g];
      const float lsq = v.x * v.x + v.y * v.y + v.z * v.z;  // Vec3.lengthSquared
      const float asq = a.x * a.x + a.y * a.y + a.z * a.z;
      if (lsq > lt || asq > at) atomicOr(&S[P[r]], 0x80000000u);
    }
    __syncthreads();
    if (tid == 0)
      for (uint32_t k = 0; k < bs / kWave; k++) rbase += wsum[k];
    __syncthreads();
  }
  for (uint32_t i = tid; i < cap; i += bs) {
    const uint32_t sz = S[i];
    if (!(sz & 0x7FFFFFFFu)) continue;
    uint32_t z;
    if (!(sz >> 31)) {
      const float t = T[i] + cfg.dt;
      T[i] = t;
      z = t >= 0.5f ? 1u : Z[i];  // ISLAND_SLEEP_TIME_THRESHOLD
    } else {
      T[i] = 0.0f;
      z = 0;
    }
    Z[i] = z;
    if (z) atomicAdd(&n_sleep, 1u);
  }
  if (L)  // labels and endpoint flags for the queries and the export
    for (uint32_t i = tid; i < cap; i += bs) {
      I.parent[(size_t)w * cap + i] = P[i];
      I.endpoint[(size_t)w * cap + i] = E[i];
    }
  __syncthreads();
  if (tid == 0) {
    W.isl.sleeping_islands = n_sleep;
    W.isl.active_islands = W.isl.total_islands - n_sleep;
  }
}

// GraphColoring.build (src/islands.zig:668-736): sequential greedy over the
// manifolds in pair-key order; static bodies count; colours saturate at 255.
__global__ void k_ref_coloring(DevCfg cfg, const WorldState* ws, const DevManifold* ml,
                               uint64_t* masks, uint8_t* colors) {
  uint32_t w = blockIdx.x;
  if (ws[w].skip || threadIdx.x != 0) return;
  uint64_t* Mk = masks + (size_t)w * cfg.id_cap * 4;
  for (uint32_t i = 0; i < cfg.id_cap * 4; i++) Mk[i] = 0;
  uint32_t m0 = ws[w].man_start, M = ws[w].man_count;
  for (uint32_t i = 0; i < M; i++) {
    uint32_t a = ml[m0 + i].m.body_a_id, b = ml[m0 + i].m.body_b_id;
    uint32_t c = 255;
    for (uint32_t q = 0; q < 4; q++) {
      uint64_t used = Mk[a * 4 + q] | Mk[b * 4 + q];
      if (~used) {
        c = q * 64 + (uint32_t)__builtin_ctzll(~used);
        break;
      }
    }
    if (c > 255) c = 255;
    colors[m0 + i] = (uint8_t)c;
    Mk[a * 4 + c / 64] |= 1ull << (c % 64);
    Mk[b * 4 + c / 64] |= 1ull << (c % 64);
  }
}

// ---------------------------------------------------------------------------
// serving: moving worlds between the resident pool and the batch slots
// ---------------------------------------------------------------------------
// Copies every descriptor's byte range; blockIdx.y selects the descriptor.
__global__ void k_copy_many(const CopyDesc* d, uint32_t n) {
  const uint32_t k = blockIdx.y;
  if (k >= n) return;
  const uint8_t* src = (const uint8_t*)d[k].src;
  uint8_t* dst = (uint8_t*)d[k].dst;
  const uint64_t bytes = d[k].bytes;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (((uintptr_t)src | (uintptr_t)dst | bytes) % 16 == 0) {
    const uint4* s4 = (const uint4*)src;
    uint4* d4 = (uint4*)dst;
    for (uint64_t i = t0; i < bytes / 16; i += stride) d4[i] = s4[i];
  } else {
    for (uint64_t i = t0; i < bytes; i += stride) dst[i] = src[i];
  }
}

// Manifolds moved into batch slots: pair keys carry the slot as world.
// ranges[3 * r ..] = (begin, count, slot); blockIdx.y = range.
__global__ void k_rekey_manifolds(DevManifold* ml, const uint32_t* ranges, uint32_t n_ranges) {
  const uint32_t r = blockIdx.y;
  if (r >= n_ranges) return;
  const uint32_t begin = ranges[3 * r], count = ranges[3 * r + 1], world = ranges[3 * r + 2];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    DevManifold& d = ml[begin + i];
    d.key = makeKey(world, keyA(d.key), keyB(d.key));
  }
}

// Empties the listed batch slots: no active bodies, no id mappings.
__global__ void k_clear_slots(DevCfg cfg, DevBodies B, const uint32_t* slots, uint32_t n) {
  const uint32_t k = blockIdx.y;
  if (k >= n) return;
  const size_t w = slots[k];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cfg.max_bodies;
       i += gridDim.x * blockDim.x)
    B.active[w * cfg.max_bodies + i] = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cfg.id_cap;
       i += gridDim.x * blockDim.x)
    B.idmap[w * cfg.id_cap + i] = -1;
}

// ---------------------------------------------------------------------------
// solver: ContactSolver.solve for one world per block (src/solver.zig:494-561)
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxColors = 1024;
constexpr uint32_t kSolveThreads = 1024;

// Solver records, indexed by position j of a solved manifold in colour order,
// stored field-major inside the world's slice (M = the world's manifold count):
// manifold field q of j at REC[q * M + j], point field q of (j, k) at
// PTS[q * 4M + k * M + j], so a wave's loads of one field are contiguous.
// Point record (8 x float4):
//   q0 normal.xyz nm | q1 point.xyz tm0 | q2 r_a.xyz tm1 | q3 r_b.xyz bias
//   q4 t1.xyz ni     | q5 t2.xyz ti0    | q6 anchor_a.xyz ti1 | q7 anchor_b.xyz base_sep
// Manifold record (9 x float4), refreshed by prepare every substep:
//   m0 (slot a, slot b, friction, count) | m1 pos_a.xyz im_a | m2 pos_b.xyz im_b
//   m3 (awake a, awake b, restitution, -) | m4..m8 inv_inertia_world a[9], b[9]
constexpr int kPtQ = 8;
constexpr int kMrQ = 9;

__device__ __forceinline__ float4 f4(float x, float y, float z, float w) {
  return make_float4(x, y, z, w);
}
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// A field's base (uniform: scalar registers) plus the manifold's byte offset as
// a 32-bit word, so every record load is a scalar-base + 32-bit vector-offset
// access with no 64-bit address arithmetic per load (a world's records stay
// far below 4 GiB per field).
struct SoaRec {
  float4* rec;
  float4* pts;
  size_t M;
  __device__ __forceinline__ static float4& at(float4* base, uint32_t j) {
    return *reinterpret_cast<float4*>(reinterpret_cast<char*>(base) + j * 16u);
  }
  __device__ __forceinline__ float4& r(uint32_t j, int q) const { return at(rec + q * M, j); }
  __device__ __forceinline__ float4& p(uint32_t j, uint32_t k, int q) const {
    return at(pts + (q * 4 + k) * M, j);
  }
};

struct MRec {
  uint32_t sa, sb, n;
  float friction;
  V3 pa, pb;
  float ima, imb;
  bool awa, awb;
  M3 iia, iib;
};

__device__ __forceinline__ MRec loadMRec(const SoaRec& S, uint32_t j) {
  float4 m0 = S.r(j, 0), m1 = S.r(j, 1), m2 = S.r(j, 2), m3 = S.r(j, 3), m4 = S.r(j, 4),
         m5 = S.r(j, 5), m6 = S.r(j, 6), m7 = S.r(j, 7), m8 = S.r(j, 8);
  MRec r;
  r.sa = __float_as_uint(m0.x);
  r.sb = __float_as_uint(m0.y);
  r.friction = m0.z;
  r.n = __float_as_uint(m0.w);
  r.pa = f4v(m1);
  r.ima = m1.w;
  r.pb = f4v(m2);
  r.imb = m2.w;
  r.awa = __float_as_uint(m3.x) != 0;
  r.awb = __float_as_uint(m3.y) != 0;
  r.iia = M3{{m4.x, m4.y, m4.z, m4.w, m5.x, m5.y, m5.z, m5.w, m6.x}};
  r.iib = M3{{m6.y, m6.z, m6.w, m7.x, m7.y, m7.z, m7.w, m8.x, m8.y}};
  return r;
}

// One constraint's two bodies held in registers for a whole manifold.
struct VPair {
  V3 la, aa, lb, ab;
};

__device__ __forceinline__ V3 velAt(V3 lv, V3 av, V3 pos, V3 pt) {
  return add(lv, cross(av, sub(pt, pos)));
}
// BodyState.applyImpulseAtPoint (src/solver.zig:316-322) on register state
__device__ __forceinline__ void impulseTo(bool awake, V3& lv, V3& av, float im, const M3& ii,
                                          V3 imp, V3 r) {
  if (!awake) return;
  lv = add(lv, scale(imp, im));
  av = add(av, mulVec(ii, cross(r, imp)));
}

// solveVelocityConstraints for one manifold (src/solver.zig:732-787) with
// solveFriction (src/solver.zig:790-838); kBias selects the soft biased pass
// or the hard relax pass (src/solver.zig:848-856).
// Every point record of the four contact slots is loaded before the first
// constraint row (slots past the count hold stale data inside the world's
// slice and are not used), so a pass waits for memory once per manifold.
template <bool kBias, class VelPtr>
__device__ __forceinline__ void solveManifoldVel(const DevCfg& cfg, VelPtr VL, VelPtr VA,
                                                 const SoaRec& S, uint32_t j) {
  MRec m = loadMRec(S, j);
  float4 Q[4][7];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int q = 0; q < 7; q++) Q[k][q] = S.p(j, k, q);
  V3 la = f4v(VL[m.sa]), aa = f4v(VA[m.sa]), lb = f4v(VL[m.sb]), ab = f4v(VA[m.sb]);
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    if (k >= m.n) break;
    const float4 q0 = Q[k][0], q1 = Q[k][1], q2 = Q[k][2], q3 = Q[k][3], q4 = Q[k][4],
                 q5 = Q[k][5], q6 = Q[k][6];
    V3 nrm = f4v(q0), pt = f4v(q1), ra = f4v(q2), rb = f4v(q3), t1 = f4v(q4), t2 = f4v(q5);
    float nm = q0.w, tm0 = q1.w, tm1 = q2.w, bias = q3.w;
    float ni = q4.w, ti0 = q5.w, ti1 = q6.w;
    V3 rel = sub(velAt(lb, ab, m.pb, pt), velAt(la, aa, m.pa, pt));
    float vn = dot(rel, nrm);
    float imp;
    if (kBias) imp = -cfg.mass_coeff * nm * (vn + bias) - cfg.impulse_coeff * ni;
    else imp = -nm * vn;
    float nw = fmaxf(ni + imp, 0.0f);
    imp = nw - ni;
    ni = nw;
    V3 Pn = scale(nrm, imp);
    impulseTo(m.awa, la, aa, m.ima, m.iia, neg(Pn), ra);
    impulseTo(m.awb, lb, ab, m.imb, m.iib, Pn, rb);
    if (cfg.enable_friction && m.friction > 0.0f) {
      V3 rel1 = sub(velAt(lb, ab, m.pb, pt), velAt(la, aa, m.pa, pt));
      float maxf = m.friction * ni;
      float vt = dot(rel1, t1);
      float fi = -tm0 * vt;
      float old = ti0;
      ti0 = clampz(old + fi, -maxf, maxf);
      fi = ti0 - old;
      V3 i1 = scale(t1, fi);
      impulseTo(m.awa, la, aa, m.ima, m.iia, neg(i1), ra);
      impulseTo(m.awb, lb, ab, m.imb, m.iib, i1, rb);
      V3 rel2 = sub(velAt(lb, ab, m.pb, pt), velAt(la, aa, m.pa, pt));
      float vt2 = dot(rel2, t2);
      float fi2 = -tm1 * vt2;
      float old2 = ti1;
      ti1 = clampz(old2 + fi2, -maxf, maxf);
      fi2 = ti1 - old2;
      V3 i2 = scale(t2, fi2);
      impulseTo(m.awa, la, aa, m.ima, m.iia, neg(i2), ra);
      impulseTo(m.awb, lb, ab, m.imb, m.iib, i2, rb);
      S.p(j, k, 5).w = ti0;
      S.p(j, k, 6).w = ti1;
    }
    S.p(j, k, 4).w = ni;
  }
  if (m.awa) {
    VL[m.sa] = v4(la);
    VA[m.sa] = v4(aa);
  }
  if (m.awb) {
    VL[m.sb] = v4(lb);
    VA[m.sb] = v4(ab);
  }
}

// The partner lane of a lane pair (2k, 2k + 1): DPP quad_perm(1, 0, 3, 2).
__device__ __forceinline__ float pairSwap(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ V3 pairSwap(V3 v) { return V3{pairSwap(v.x), pairSwap(v.y), pairSwap(v.z)}; }

// One row's velocity change of the lane's own body: BodyState.applyImpulse
// AtPoint (src/solver.zig:316-322) with the impulse this body receives (already
// negated for body A).  Lanes whose body is not awake skip it under the exec
// mask instead of computing and discarding it.
__device__ __forceinline__ void impulseMine(bool awake, V3& lv, V3& av, float im, const M3& ii,
                                            V3 imp, V3 r) {
  if (awake) {
    lv = add(lv, scale(imp, im));
    av = add(av, mulVec(ii, cross(r, imp)));
  }
}

// solveManifoldVel on a lane pair: lane `side` 0 holds body A, side 1 body B.
// Each lane evaluates the contact velocity of its own body and applies its own
// body's impulses; the relative velocity is assembled from the partner's value
// and every shared scalar (vn, impulse, clamp) is computed identically in both
// lanes, so the result equals the one-lane pass bit for bit.  Both lanes of a
// pair follow the same control flow (same manifold, count and friction).
//
// Signs without branches: each lane forms d = mine - partner, which is
// vb - va in lane B and its exact negation in lane A, so dot(d, axis) times
// sgn (+1 for B, -1 for A) is the reference's dot(vb - va, axis) (negating
// both operands of a rounded subtraction, or the terms of a rounded sum,
// negates the result exactly; only an exact zero can come out with the other
// sign, which changes no later non-zero value); the impulse a lane's body
// receives is the row impulse times sgn, which is the reference's neg(P) for
// body A bit for bit (-(x * y) == x * (-y)).
template <bool kBias, class VelPtr>
__device__ __forceinline__ void solveManifoldVel2(const DevCfg& cfg, VelPtr VL, VelPtr VA,
                                                  const SoaRec& S, uint32_t j, uint32_t side) {
  const float4 r0 = S.r(j, 0), r3 = S.r(j, 3);
  const uint32_t slot = __float_as_uint(side ? r0.y : r0.x);
  const uint32_t n = __float_as_uint(r0.w);
  const float friction = r0.z;
  const float4 rp = S.r(j, side ? 2 : 1);  // position, inverse mass
  const V3 pos = f4v(rp);
  const float im = rp.w;
  const bool aw = __float_as_uint(side ? r3.y : r3.x) != 0;
  const float4 m4 = S.r(j, 4), m5 = S.r(j, 5), m6 = S.r(j, 6), m7 = S.r(j, 7), m8 = S.r(j, 8);
  const M3 ii = side ? M3{{m6.y, m6.z, m6.w, m7.x, m7.y, m7.z, m7.w, m8.x, m8.y}}
                     : M3{{m4.x, m4.y, m4.z, m4.w, m5.x, m5.y, m5.z, m5.w, m6.x}};
  float4 Q[4][7];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int q = 0; q < 7; q++) Q[k][q] = S.p(j, k, q);
  V3 lv = f4v(VL[slot]), av = f4v(VA[slot]);
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    if (k >= n) break;
    const float4 q0 = Q[k][0], q1 = Q[k][1], q2 = Q[k][2], q3 = Q[k][3], q4 = Q[k][4],
                 q5 = Q[k][5], q6 = Q[k][6];
    V3 nrm = f4v(q0), pt = f4v(q1), t1 = f4v(q4), t2 = f4v(q5);
    V3 r = side ? f4v(q3) : f4v(q2);
    float nm = q0.w, tm0 = q1.w, tm1 = q2.w, bias = q3.w;
    float ni = q4.w, ti0 = q5.w, ti1 = q6.w;
    const float sgn = side ? 1.0f : -1.0f;  // A receives -P, B receives +P (exact negation)
    V3 mine = velAt(lv, av, pos, pt);
    V3 d = sub(mine, pairSwap(mine));
    float vn = dot(d, nrm) * sgn;
    float imp;
    if (kBias) imp = -cfg.mass_coeff * nm * (vn + bias) - cfg.impulse_coeff * ni;
    else imp = -nm * vn;
    float nw = fmaxf(ni + imp, 0.0f);
    imp = nw - ni;
    ni = nw;
    impulseMine(aw, lv, av, im, ii, scale(nrm, imp * sgn), r);
    if (cfg.enable_friction && friction > 0.0f) {
      mine = velAt(lv, av, pos, pt);
      d = sub(mine, pairSwap(mine));
      float maxf = friction * ni;
      float vt = dot(d, t1) * sgn;
      float fi = -tm0 * vt;
      float old = ti0;
      ti0 = clampz(old + fi, -maxf, maxf);
      fi = ti0 - old;
      impulseMine(aw, lv, av, im, ii, scale(t1, fi * sgn), r);
      mine = velAt(lv, av, pos, pt);
      d = sub(mine, pairSwap(mine));
      float vt2 = dot(d, t2) * sgn;
      float fi2 = -tm1 * vt2;
      float old2 = ti1;
      ti1 = clampz(old2 + fi2, -maxf, maxf);
      fi2 = ti1 - old2;
      impulseMine(aw, lv, av, im, ii, scale(t2, fi2 * sgn), r);
      S.p(j, k, 5).w = ti0;  // both lanes hold the same value
      S.p(j, k, 6).w = ti1;
    }
    S.p(j, k, 4).w = ni;
  }
  if (aw) {
    VL[slot] = v4(lv);
    VA[slot] = v4(av);
  }
}

// warmStart for one manifold (src/solver.zig:702-729)
template <class VelPtr>
__device__ __forceinline__ void warmManifold(float wss, VelPtr VL, VelPtr VA, const SoaRec& S,
                                             uint32_t j) {
  MRec m = loadMRec(S, j);
  V3 la = f4v(VL[m.sa]), aa = f4v(VA[m.sa]), lb = f4v(VL[m.sb]), ab = f4v(VA[m.sb]);
  for (uint32_t k = 0; k < m.n; k++) {
    float4 q0 = S.p(j, k, 0), q2 = S.p(j, k, 2), q3 = S.p(j, k, 3), q4 = S.p(j, k, 4),
           q5 = S.p(j, k, 5), q6 = S.p(j, k, 6);
    float sn = q4.w * wss, s1 = q5.w * wss, s2 = q6.w * wss;
    V3 imp = add(add(scale(f4v(q0), sn), scale(f4v(q4), s1)), scale(f4v(q5), s2));
    impulseTo(m.awa, la, aa, m.ima, m.iia, neg(imp), f4v(q2));
    impulseTo(m.awb, lb, ab, m.imb, m.iib, imp, f4v(q3));
    S.p(j, k, 4).w = sn;
    S.p(j, k, 5).w = s1;
    S.p(j, k, 6).w = s2;
  }
  if (m.awa) {
    VL[m.sa] = v4(la);
    VA[m.sa] = v4(aa);
  }
  if (m.awb) {
    VL[m.sb] = v4(lb);
    VA[m.sb] = v4(ab);
  }
}

__device__ __forceinline__ M3 loadM3(const float* p) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = p[i];
  return r;
}
__device__ __forceinline__ void storeM3(float* p, const M3& m) {
  for (int i = 0; i < 9; i++) p[i] = m.m[i];
}

// src/solver.zig:967-981
__device__ __forceinline__ V3 computeOrthogonal(V3 v) {
  if (fabsf(v.x) < fabsf(v.y)) {
    if (fabsf(v.x) < fabsf(v.z)) return normalizeOrZero(V3{0.0f, -v.z, v.y});
    return normalizeOrZero(V3{-v.y, v.x, 0.0f});
  }
  if (fabsf(v.y) < fabsf(v.z)) return normalizeOrZero(V3{-v.z, 0.0f, v.x});
  return normalizeOrZero(V3{-v.y, v.x, 0.0f});
}


__device__ __forceinline__ void stamp(uint64_t* t, int i) {
  if (t && threadIdx.x == 0) t[blockIdx.x * 16 + i] = wall_clock64();
}
// Timestamp i of world w, written by a block that is not the world's own.
__device__ __forceinline__ void stampW(uint64_t* t, uint32_t w, int i) {
  if (t && threadIdx.x == 0) t[w * 16 + i] = wall_clock64();
}

// ContactSolver.solve (src/solver.zig:494-561), split into solve units.
//
// A unit is a set of connected components of the graph whose nodes are the
// bodies the step moves (dynamic awake and kinematic) and whose edges are the
// solved manifolds between two such bodies.  Every body a unit writes belongs
// to that unit only, and every other body its manifolds touch (static,
// sleeping) is only read during the solve, so units are independent: solving
// each one in its own workgroup, with the colour order restricted to its
// manifolds, performs exactly the operations of the world-wide colour-ordered
// pass on every body and manifold.  The only world-wide quantities are the
// position-pass early exit (minimum separation over all manifolds, evaluated
// between launches) and the maximum penetration (a maximum, combined with
// atomics on the float bits of non-negative values).
//
//   k_solve_prep   block / world: body working state, colouring, units
//   k_solve_units  persistent grid: one unit at a time, all substeps and the
//                  first position pass, body velocities in LDS
//   k_solve_pos    position passes 2.. (one launch per pass), early exit
//   k_solve_final  block / world: writeBack, updateSleepStates, stats
constexpr uint32_t kUnitsMax = kUnitsPerWorld;

__device__ __forceinline__ bool mobileBody(const DevBodies& B, uint32_t g) {
  return B.active[g] && B.type[g] != VXP_BODY_STATIC && !B.sleeping[g];
}

// Greedy colouring of a world's solved manifolds in ascending (prio, key)
// order, run as a dataflow: a manifold takes the smallest colour unused by its
// lower neighbours once they are all coloured.  Each manifold counts its lower
// neighbours; rounds colour the ready ones and release their higher
// neighbours into the next round (one queue; each manifold enters once).
// Same colours as the sequential greedy.  A world's manifolds are stored in
// ascending key order, so the index breaks prio ties exactly as the key does.
// With kL the per-manifold state and the incidence lists live in LDS
// (kColourLdsPerManifold bytes per manifold).
constexpr uint32_t kColourLdsPerManifold = 40;
__device__ __forceinline__ bool colourBefore(uint32_t pn, uint32_t n, uint32_t pm, uint32_t m) {
  return pn < pm || (pn == pm && n < m);
}
template <bool kL>
__device__ void colourManifolds(uint32_t M, uint32_t nadj, uint32_t m0, uint32_t g0,
                                const DevBodies& B, const DevManifold* ml, const SolverScratch& S,
                                const uint32_t* prio_g, const uint32_t* adj_g, uint4* kdyn,
                                WorldState& x, uint32_t& s_tail, uint32_t& s_ncol) {
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  uint4* ce;          // per manifold: (offset, count) of each endpoint's incidence list
  uint32_t* cp;       // prio
  int32_t* cc;        // colour (-1 uncoloured, -2 not solved)
  uint32_t* cn;       // uncoloured lower neighbours
  uint32_t* cq;       // queue
  const uint32_t* cl; // incidence lists (manifold indices)
  if (kL) {
    ce = kdyn;
    cp = (uint32_t*)(kdyn + M);
    cc = (int32_t*)(cp + M);
    cn = (uint32_t*)(cc + M);
    cq = cn + M;
    uint32_t* l = cq + M;
    for (uint32_t k = tid; k < nadj; k += bs) l[k] = adj_g[k];
    cl = l;
  } else {
    ce = S.cend + m0;
    cp = (uint32_t*)prio_g + m0;
    cc = S.color + m0;
    cn = S.pcnt + m0;
    cq = S.fq + m0;
    cl = adj_g;
  }
  for (uint32_t i = tid; i < M; i += bs) {
    const DevManifold& d = ml[m0 + i];
    uint32_t ga = g0 + d.sa, gb = g0 + d.sb;
    bool la = B.type[ga] == VXP_BODY_DYNAMIC && !B.sleeping[ga];
    bool lb = B.type[gb] == VXP_BODY_DYNAMIC && !B.sleeping[gb];
    ce[i] = make_uint4(la ? S.adj_off[ga] : 0u, la ? S.adj_cnt[ga] : 0u, lb ? S.adj_off[gb] : 0u,
                       lb ? S.adj_cnt[gb] : 0u);
    if (kL) {
      cp[i] = prio_g[m0 + i];
      cc[i] = S.color[m0 + i];
    }
  }
  if (tid == 0) s_tail = 0;
  __syncthreads();
  // Neighbours are visited eight at a time so that each level of the
  // dependent loads (incidence entry, then its prio, then its colour or count)
  // is issued together; a padding entry names the manifold itself.
  constexpr uint32_t kB = 8;
  for (uint32_t i = tid; i < M; i += bs) {
    if (cc[i] != -1) continue;
    const uint4 e = ce[i];
    const uint32_t pm = cp[i];
    uint32_t lower = 0;
    for (int side = 0; side < 2; side++) {
      const uint32_t o = side ? e.z : e.x, c = side ? e.w : e.y;
      for (uint32_t k0 = 0; k0 < c; k0 += kB) {
        uint32_t nn[kB], pp[kB];
#pragma unroll
        for (uint32_t t = 0; t < kB; t++) nn[t] = k0 + t < c ? cl[o + k0 + t] : i;
#pragma unroll
        for (uint32_t t = 0; t < kB; t++) pp[t] = cp[nn[t]];
#pragma unroll
        for (uint32_t t = 0; t < kB; t++) lower += colourBefore(pp[t], nn[t], pm, i) ? 1u : 0u;
      }
    }
    cn[i] = lower;
    if (!lower) cq[atomicAdd(&s_tail, 1u)] = i;
  }
  __syncthreads();
  stamp(S.tstamp, 7);
  // Rounds: a group of kG lanes takes one ready manifold; its lanes split the
  // incidence lists, OR their lower neighbours' colours, and release the higher
  // neighbours whose last lower neighbour this was (enqueued per wave).
  constexpr uint32_t kG = 8;
  const uint32_t gl = tid % kG, ngroups = bs / kG;
  const uint32_t lane = tid & (kWave - 1);
  uint32_t head = 0, tail = s_tail;
  __syncthreads();
  while (head < tail) {
    for (uint32_t q = head + tid / kG; q < tail; q += ngroups) {
      const uint32_t i = cq[q];
      const uint4 e = ce[i];
      const uint32_t pm = cp[i];
      uint64_t used0 = 0;  // colours below 64 taken by lower neighbours
      for (int side = 0; side < 2; side++) {
        const uint32_t o = side ? e.z : e.x, c = side ? e.w : e.y;
        for (uint32_t k = gl; k < c; k += kG) {
          const uint32_t n = cl[o + k];
          const uint32_t pn = cp[n];
          const int32_t cv = __atomic_load_n(&cc[n], __ATOMIC_RELAXED);
          if (colourBefore(pn, n, pm, i) && cv >= 0 && cv < 64) used0 |= 1ull << cv;
        }
      }
#pragma unroll
      for (uint32_t d = 1; d < kG; d <<= 1) used0 |= shflXor64(used0, d);
      if (gl == 0) {
        uint32_t col;
        if (~used0) {
          col = (uint32_t)__builtin_ctzll(~used0);
        } else {  // all of 0..63 taken: the full mask, on the group's first lane
          uint64_t used[kMaxColors / 64];
          for (uint32_t u = 0; u < kMaxColors / 64; u++) used[u] = 0;
          for (int side = 0; side < 2; side++) {
            const uint32_t o = side ? e.z : e.x, c = side ? e.w : e.y;
            for (uint32_t k = 0; k < c; k++) {
              uint32_t n = cl[o + k];
              if (n == i || !colourBefore(cp[n], n, pm, i)) continue;
              int32_t cv = __atomic_load_n(&cc[n], __ATOMIC_RELAXED);
              if (cv >= 0 && cv < (int32_t)kMaxColors) used[cv >> 6] |= 1ull << (cv & 63);
            }
          }
          col = kMaxColors;
          for (uint32_t u = 0; u < kMaxColors / 64; u++)
            if (~used[u]) {
              col = u * 64 + (uint32_t)__builtin_ctzll(~used[u]);
              break;
            }
        }
        if (col >= kMaxColors) {
          col = kMaxColors - 1;
          atomicOr(&x.error, kErrColours);
        }
        __atomic_store_n(&cc[i], (int32_t)col, __ATOMIC_RELAXED);
        atomicMax(&s_ncol, col + 1);
      }
      // release the higher neighbours (they run in a later round, after the
      // barrier that makes this colour visible)
      for (int side = 0; side < 2; side++) {
        const uint32_t o = side ? e.z : e.x, c = side ? e.w : e.y;
        for (uint32_t k = gl; k < c; k += kG) {
          const uint32_t n = cl[o + k];
          const bool hi = colourBefore(pm, i, cp[n], n);
          const bool rel = hi && atomicSub(&cn[n], 1u) == 1u;
          const uint64_t m = __ballot(rel);
          if (m) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&s_tail, (uint32_t)__popcll(m));
            base = (uint32_t)__shfl((int)base, leader, kWave);
            if (rel) cq[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = n;
          }
        }
      }
    }
    __syncthreads();
    head = tail;
    tail = s_tail;
    __syncthreads();
  }
  if (kL)
    for (uint32_t i = tid; i < M; i += bs) S.color[m0 + i] = cc[i];
  __syncthreads();
}

// Body working state (BodyState.fromBody, src/solver.zig:231-243) of every
// slot of the worlds that step: pose, velocities and world inverse inertia,
// which only the solve units and later passes read.
__global__ void k_body_state(DevCfg cfg, DevBodies B, const WorldState* ws, SolverScratch S,
                             uint32_t n_slots) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_slots || ws[g / cfg.max_bodies].skip) return;
  if (!B.active[g]) {
    S.ws_lv[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    S.ws_av[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const float4 rot = B.rot[g];
  S.ws_pos[g] = B.pos[g];
  S.ws_rot[g] = rot;
  S.ws_lv[g] = B.lv[g];
  S.ws_av[g] = B.av[g];
  M3 ii;
  if (B.inv_mass[g] == 0.0f) {
    for (int i = 0; i < 9; i++) ii.m[i] = 0.0f;
  } else {
    ii = worldInvInertia(f4q(rot), loadM3(B.iil + (size_t)g * 9));
  }
  storeM3(S.ws_ii + (size_t)g * 9, ii);
}

__global__ __launch_bounds__(kSolveThreads) void k_solve_prep(DevCfg cfg, DevBodies B,
                                                              WorldState* ws, vxp_stats* st,
                                                              DevManifold* ml, SolverScratch S,
                                                              uint32_t* prio_buf,
                                                              uint32_t* step_ctr,
                                                              size_t lds_bytes,
                                                              const uint32_t* man_total) {
  const uint32_t w = blockIdx.x;
  if (ws[w].skip) return;
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  const uint32_t MB = cfg.max_bodies, g0 = w * MB;
  const uint32_t m0 = ws[w].man_start, M = ws[w].man_count;
  __shared__ uint32_t tmp[kSolveThreads];
  __shared__ uint32_t s_left, s_ncol, s_nsolve, s_skipped, s_points, s_nunits,
      s_points_tot;
  extern __shared__ uint4 kdyn[];
  __shared__ uint32_t cnt_m[kUnitsMax], cnt_b[kUnitsMax], off_m[kUnitsMax], off_b[kUnitsMax];
  __shared__ uint32_t uid[kUnitsMax];
  stamp(S.tstamp, 0);

  // (the body working state is k_body_state's, beside this kernel)
  // Awake-dynamic incidence counts AC and CSR offsets AO per body slot, in LDS
  // (the colouring's space, unused until then) when they fit.
  const bool csr_l = (size_t)MB * 8 <= lds_bytes;
  uint32_t* AC = csr_l ? reinterpret_cast<uint32_t*>(kdyn) : S.adj_cnt + g0;
  uint32_t* AO = csr_l ? AC + MB : S.adj_off + g0;
  for (uint32_t s = tid; s < MB; s += bs) AC[s] = 0;
  if (tid == 0) {
    s_nsolve = 0;
    s_skipped = 0;
    s_ncol = 0;
    s_points = 0;
    s_nunits = 1;
  }
  __syncthreads();
  // solvable manifolds (src/system.zig:687-702) and awake-dynamic incidence counts
  for (uint32_t i = tid; i < M; i += bs) {
    DevManifold& d = ml[m0 + i];
    uint32_t ga = g0 + d.sa, gb = g0 + d.sb;
    bool asl = B.sleeping[ga], bsl = B.sleeping[gb];
    d.m.color = kNone;
    if (asl && bsl) {
      S.color[m0 + i] = -2;
      atomicAdd(&s_skipped, 1u);
      continue;
    }
    atomicAdd(&s_nsolve, 1u);
    atomicAdd(&s_points, d.m.contact_count);
    S.color[m0 + i] = -1;
    prio_buf[m0 + i] = colorPrio(exportKey(d.key));
    if (B.type[ga] == VXP_BODY_DYNAMIC && !asl) atomicAdd(&AC[d.sa], 1u);
    if (B.type[gb] == VXP_BODY_DYNAMIC && !bsl) atomicAdd(&AC[d.sb], 1u);
  }
  __syncthreads();
  // CSR offsets (relative to 2 * m0)
  for (uint32_t s = tid; s < MB; s += bs) AO[s] = AC[s];
  __syncthreads();
  blockScanExclusive(AO, MB, tmp, &s_points_tot);
  const uint32_t nadj = s_points_tot;
  for (uint32_t s = tid; s < MB; s += bs) AC[s] = 0;
  __syncthreads();
  uint32_t* adj = S.adj + 2 * (size_t)m0;
  for (uint32_t i = tid; i < M; i += bs) {
    if (S.color[m0 + i] != -1) continue;
    const DevManifold& d = ml[m0 + i];
    uint32_t ga = g0 + d.sa, gb = g0 + d.sb;
    if (B.type[ga] == VXP_BODY_DYNAMIC && !B.sleeping[ga])
      adj[AO[d.sa] + atomicAdd(&AC[d.sa], 1u)] = i;
    if (B.type[gb] == VXP_BODY_DYNAMIC && !B.sleeping[gb])
      adj[AO[d.sb] + atomicAdd(&AC[d.sb], 1u)] = i;
  }
  __syncthreads();
  if (csr_l) {  // the colouring reads them from global memory
    for (uint32_t s = tid; s < MB; s += bs) {
      S.adj_off[g0 + s] = AO[s];
      S.adj_cnt[g0 + s] = AC[s];
    }
    __syncthreads();
  }
  stamp(S.tstamp, 1);
  if ((size_t)M * kColourLdsPerManifold <= lds_bytes)
    colourManifolds<true>(M, nadj, m0, g0, B, ml, S, prio_buf, adj, kdyn, ws[w], s_left, s_ncol);
  else
    colourManifolds<false>(M, nadj, m0, g0, B, ml, S, prio_buf, adj, kdyn, ws[w], s_left, s_ncol);
  __syncthreads();
  stamp(S.tstamp, 2);
  const uint32_t nsolve = s_nsolve;
  for (uint32_t i = tid; i < M; i += bs) {
    int32_t c = S.color[m0 + i];
    if (c >= 0) ml[m0 + i].m.color = (uint32_t)c;
  }
  // Solve units.  Working arrays: UE per manifold (slot a, slot b, bits:
  // 1 solved, 2 a mobile, 4 b mobile) and, per body slot, the union-find
  // parent P, solved manifolds per component root UC and their prefix UO;
  // in LDS (the colouring's space, free now) when they fit.
  const bool ul = (size_t)M * 16 + (size_t)MB * 12 <= lds_bytes;
  uint4* UE = ul ? kdyn : S.cend + m0;
  uint32_t* P = ul ? (uint32_t*)(kdyn + M) : S.upar + g0;
  uint32_t* UC = ul ? P + MB : S.ucnt + g0;
  uint32_t* UO = ul ? UC + MB : S.uoff + g0;
  __syncthreads();
  for (uint32_t s = tid; s < MB; s += bs) {
    P[s] = s;
    UC[s] = 0;
  }
  for (uint32_t i = tid; i < M; i += bs) {
    const DevManifold& d = ml[m0 + i];
    const uint32_t bits = (S.color[m0 + i] >= 0 ? 1u : 0u) | (mobileBody(B, g0 + d.sa) ? 2u : 0u) |
                          (mobileBody(B, g0 + d.sb) ? 4u : 0u);
    UE[i] = make_uint4(d.sa, d.sb, bits, 0u);
  }
  __syncthreads();
  // components over solved manifolds between two mobile bodies, each labelled
  // by its smallest slot
  for (uint32_t i = tid; i < M; i += bs) {
    const uint4 e = UE[i];
    if (e.z == 7u) unite(P, e.x, e.y);
  }
  __syncthreads();
  for (uint32_t s = tid; s < MB; s += bs) P[s] = findRoot(P, s);
  __syncthreads();
  // solved manifolds per component (indexed by the root slot), then units:
  // components in root order, cut every T manifolds (a component goes to the
  // unit where it starts)
  for (uint32_t i = tid; i < M; i += bs) {
    const uint4 e = UE[i];
    if ((e.z & 1u) && (e.z & 6u)) atomicAdd(&UC[P[(e.z & 2u) ? e.x : e.y]], 1u);
  }
  __syncthreads();
  for (uint32_t s = tid; s < MB; s += bs) UO[s] = UC[s];
  __syncthreads();
  blockScanExclusive(UO, MB, tmp, nullptr);
  // solve units of at least max(64, all manifolds / (2 CUs)) manifolds: about
  // two per CU over the whole batch (results do not depend on it)
  const uint32_t unit_min = max(64u, man_total[0] / (2u * max(cfg.n_cu, 1u)));
  const uint32_t T = max(unit_min, (nsolve + kUnitsMax - 1) / kUnitsMax);
  for (uint32_t u = tid; u < kUnitsMax; u += bs) cnt_m[u] = cnt_b[u] = uid[u] = 0;
  __syncthreads();
  for (uint32_t s = tid; s < MB; s += bs)
    if (UC[s]) uid[UO[s] / T] = 1;
  __syncthreads();
  if (tid == 0) {  // cut positions that start a component, numbered densely
    uint32_t n = 0;
    for (uint32_t u = 0; u < kUnitsMax; u++) {
      uint32_t used = uid[u];
      uid[u] = n;
      n += used;
    }
    s_nunits = max(n, 1u);
  }
  __syncthreads();
  const uint32_t nu = s_nunits;
  // unit of each mobile body (bodies without solved manifolds: by slot) and
  // of each solved manifold (one with no mobile body: unit 0)
  for (uint32_t s = tid; s < MB; s += bs) {
    uint32_t g = g0 + s;
    if (!mobileBody(B, g)) continue;
    uint32_t r = P[s];
    uint32_t u = UC[r] ? uid[UO[r] / T] : s % nu;
    S.adj_cnt[g] = u;
    atomicAdd(&cnt_b[u], 1u);
  }
  for (uint32_t i = tid; i < M; i += bs) {
    const uint4 e = UE[i];
    if (!(e.z & 1u)) continue;
    uint32_t u = (e.z & 6u) ? uid[UO[P[(e.z & 2u) ? e.x : e.y]] / T] : 0;
    prio_buf[m0 + i] = u;
    atomicAdd(&cnt_m[u], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t am = 0, ab = 0, big = 0;
    for (uint32_t u = 0; u < nu; u++) {
      big = max(big, cnt_m[u]);
      off_m[u] = am;
      off_b[u] = ab;
      S.units[(size_t)w * kUnitsMax + u] = make_uint4(am, cnt_m[u], ab, cnt_b[u]);
      am += cnt_m[u];
      ab += cnt_b[u];
    }
    uint32_t base = atomicAdd(&step_ctr[kCtrSolveUnits], nu);
    for (uint32_t u = 0; u < nu; u++) S.uqueue[base + u] = w * kUnitsMax + u;
    WorldState& x = ws[w];
    x.n_units = nu;
    x.nsolve = nsolve;
    x.max_unit_m = big;
    x.pos_run = (nsolve > 0 && cfg.pos_iters > 0) ? 1u : 0u;
    x.pos_done = 0;
    x.maxpen = 0;
    x.pos_min[0] = x.pos_min[1] = x.pos_min[2] = 0;
    vxp_stats& sx = st[w];
    sx.manifolds = M;
    sx.solver_manifolds = nsolve;
    sx.solver_points = s_points;
    sx.manifolds_skipped_sleeping = s_skipped;
    sx.colors = s_ncol;
  }
  __syncthreads();
  for (uint32_t u = tid; u < nu; u += bs) cnt_m[u] = cnt_b[u] = 0;
  __syncthreads();
  for (uint32_t s = tid; s < MB; s += bs) {
    uint32_t g = g0 + s;
    if (!mobileBody(B, g)) continue;
    uint32_t u = S.adj_cnt[g];
    S.ubody[g0 + off_b[u] + atomicAdd(&cnt_b[u], 1u)] = s;
  }
  for (uint32_t i = tid; i < M; i += bs) {
    if (S.color[m0 + i] < 0) continue;
    uint32_t u = prio_buf[m0 + i];
    S.uorder[m0 + off_m[u] + atomicAdd(&cnt_m[u], 1u)] = i;
  }
  stamp(S.tstamp, 3);
}

// The unit's manifolds grouped by colour: S.order[m0 + mb + j] for j in
// [ccount[c], ccount[c + 1]) has colour c.  Returns the unit's colour count.
__device__ uint32_t unitColours(const SolverScratch& S, uint32_t m0, uint32_t mb, uint32_t mn,
                                bool write_order, uint32_t* ccount, uint32_t* ccursor,
                                uint32_t* tmp, uint32_t* s_ncol) {
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  for (uint32_t c = tid; c <= kMaxColors; c += bs) ccount[c] = 0;
  if (tid == 0) *s_ncol = 0;
  __syncthreads();
  for (uint32_t k = tid; k < mn; k += bs) {
    uint32_t c = (uint32_t)S.color[m0 + S.uorder[m0 + mb + k]];
    atomicAdd(&ccount[c], 1u);
    atomicMax(s_ncol, c + 1);
  }
  __syncthreads();
  const uint32_t ncol = *s_ncol;
  blockScanExclusive(ccount, ncol + 1, tmp, nullptr);
  if (write_order) {
    for (uint32_t c = tid; c < ncol; c += bs) ccursor[c] = 0;
    __syncthreads();
    for (uint32_t k = tid; k < mn; k += bs) {
      uint32_t i = S.uorder[m0 + mb + k];
      uint32_t c = (uint32_t)S.color[m0 + i];
      S.order[m0 + mb + ccount[c] + atomicAdd(&ccursor[c], 1u)] = i;
    }
    __syncthreads();
  }
  return ncol;
}

// Minimum (kMax false) or maximum of v over the block, for any block size:
// the block is first folded onto the largest power of two below it.
template <bool kMax>
__device__ float blockReduce(float v, float* red) {
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  uint32_t p2 = 1;
  while (p2 * 2 <= bs) p2 *= 2;
  red[tid] = v;
  __syncthreads();
  if (tid < bs - p2) red[tid] = kMax ? fmaxf(red[tid], red[tid + p2]) : fminf(red[tid], red[tid + p2]);
  __syncthreads();
  for (uint32_t d = p2 / 2; d > 0; d >>= 1) {
    if (tid < d) red[tid] = kMax ? fmaxf(red[tid], red[tid + d]) : fminf(red[tid], red[tid + d]);
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

// solvePositionConstraints over one unit's manifolds (src/solver.zig:860-923);
// returns the unit's minimum separation (block-reduced) and raises maxpen.
__device__ float unitPositionPass(const DevCfg& cfg, const DevBodies& B, const SolverScratch& S,
                                  const SoaRec& SR, uint32_t g0, uint32_t mb, uint32_t ncol,
                                  const uint32_t* ccount, float* s_red, float& maxpen) {
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  float minsep = 0.0f;
  for (uint32_t c = 0; c < ncol; c++) {
    for (uint32_t j = mb + ccount[c] + tid; j < mb + ccount[c + 1]; j += bs) {
      float4 r0 = SR.r(j, 0);
      uint32_t ga = g0 + __float_as_uint(r0.x), gb = g0 + __float_as_uint(r0.y);
      uint32_t n = __float_as_uint(r0.w);
      uint8_t ta = B.type[ga], tb = B.type[gb];
      if (ta == VXP_BODY_STATIC && tb == VXP_BODY_STATIC) continue;
      float ima = B.inv_mass[ga], imb = B.inv_mass[gb];
      bool awa = ta == VXP_BODY_DYNAMIC && !B.sleeping[ga];
      bool awb = tb == VXP_BODY_DYNAMIC && !B.sleeping[gb];
      M3 iia = loadM3(S.ws_ii + (size_t)ga * 9), iib = loadM3(S.ws_ii + (size_t)gb * 9);
      V3 pa = f4v(S.ws_pos[ga]), pb = f4v(S.ws_pos[gb]);
      Q4 qa = f4q(S.ws_rot[ga]), qb = f4q(S.ws_rot[gb]);
      for (uint32_t k = 0; k < n; k++) {
        V3 nrm = f4v(SR.p(j, k, 0));
        float4 q6 = SR.p(j, k, 6), q7 = SR.p(j, k, 7);
        V3 wa = add(qrotate(qa, f4v(q6)), pa);
        V3 wb = add(qrotate(qb, f4v(q7)), pb);
        V3 ad = sub(wb, wa);
        float sep = q7.w + dot(ad, nrm);
        minsep = fminf(minsep, sep);
        maxpen = fmaxf(maxpen, -sep);
        if (sep >= -cfg.slop) continue;
        float corr = fmaxf(cfg.pcr * (sep + cfg.slop), -cfg.mpc);
        V3 ra = sub(wa, pa), rb = sub(wb, pb);
        V3 rna = cross(ra, nrm), rnb = cross(rb, nrm);
        float k_ = ima + imb;
        k_ += dot(rna, mulVec(iia, rna));
        k_ += dot(rnb, mulVec(iib, rnb));
        if (k_ < kEps) continue;
        V3 Pi = scale(nrm, -corr / k_);
        if (awa) {
          pa = sub(pa, scale(Pi, ima));
          qa = integrateRotationSmall(qa, mulVec(iia, cross(ra, neg(Pi))));
        }
        if (awb) {
          pb = add(pb, scale(Pi, imb));
          qb = integrateRotationSmall(qb, mulVec(iib, cross(rb, Pi)));
        }
      }
      if (awa) {
        S.ws_pos[ga] = v4(pa);
        S.ws_rot[ga] = q4(qa);
      }
      if (awb) {
        S.ws_pos[gb] = v4(pb);
        S.ws_rot[gb] = q4(qb);
      }
    }
    __syncthreads();
  }
  return blockReduce<false>(minsep, s_red);
}

// Combine a unit's position-pass results into its world: the minimum
// separation of pass `slot` (float bits of values <= 0 order as unsigned
// integers in reverse) and the maximum penetration (bits of values >= +0).
__device__ void unitPositionCombine(WorldState& x, uint32_t slot, float minsep, float maxpen,
                                    float* s_red) {
  const uint32_t tid = threadIdx.x;
  const float red = blockReduce<true>(maxpen, s_red);
  if (tid == 0) {
    float mp = red;
    if (mp == 0.0f) mp = 0.0f;
    float ms = minsep;
    if (ms == 0.0f) ms = 0.0f;
    atomicMax(&x.pos_min[slot], __float_as_uint(ms));
    atomicMax(&x.maxpen, __float_as_uint(mp));
  }
  __syncthreads();
}

// Unit blocks of kT threads: 512 lets one large unit run its colours on more
// lane pairs; 256 fits two blocks per CU (the passes keep a manifold's records
// in registers, about 210 VGPRs), for launches with many small units.
// A run of colour passes (src/solver.zig:526-548,848-856): colour c of cn
// manifolds goes to lane pairs (two lanes per manifold, bs / 2 manifolds at a
// time) when cn <= pair_color_max, to single lanes otherwise.
template <bool kBias, class VelPtr>
__device__ __forceinline__ void colourPasses(const DevCfg& cfg, VelPtr VL, VelPtr VA,
                                             const SoaRec& SR, const uint32_t* ccount, uint32_t mb,
                                             uint32_t ncol, uint32_t tid, uint32_t bs) {
  for (uint32_t c = 0; c < ncol; c++) {
    const uint32_t c0 = mb + ccount[c], cn = ccount[c + 1] - ccount[c];
    if (cn <= cfg.pair_color_max) {
      for (uint32_t t = tid >> 1; t < cn; t += bs >> 1)
        solveManifoldVel2<kBias>(cfg, VL, VA, SR, c0 + t, tid & 1u);
    } else {
      for (uint32_t j = c0 + tid; j < c0 + cn; j += bs)
        solveManifoldVel<kBias>(cfg, VL, VA, SR, j);
    }
    __syncthreads();
  }
}

template <bool kLds, uint32_t kT>
__global__ __launch_bounds__(kT) void k_solve_units(DevCfg cfg, DevBodies B, WorldState* ws,
                                                    DevManifold* ml, SolverScratch S,
                                                    uint32_t* step_ctr) {
  extern __shared__ float4 dyn_vel[];
  __shared__ uint32_t tmp[kT];
  __shared__ uint32_t ccount[kMaxColors + 1];
  __shared__ uint32_t ccursor[kMaxColors];
  __shared__ float s_red[kT];
  __shared__ uint32_t s_item, s_ncol;
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  const uint32_t total = step_ctr[kCtrSolveUnits];
  const uint32_t MB = cfg.max_bodies;
  for (;;) {
    if (tid == 0) s_item = atomicAdd(&step_ctr[kCtrSolveUnitPull], 1u);
    __syncthreads();
    const uint32_t item = s_item;
    __syncthreads();
    if (item >= total) break;
    const uint32_t q = S.uqueue[item];
    const uint32_t w = q / kUnitsMax;
    const uint32_t g0 = w * MB;
    const uint32_t m0 = ws[w].man_start, M = ws[w].man_count;
    const uint4 ud = S.units[q];
    const uint32_t mb = ud.x, mn = ud.y, bn = ud.w;
    const uint32_t* ub = S.ubody + g0 + ud.z;
    uint64_t* const ts = mn == ws[w].max_unit_m ? S.tstamp : nullptr;  // the largest unit
    stampW(ts, w, 8);
    float4* VL;
    float4* VA;
    if (kLds) {
      VL = dyn_vel;
      VA = dyn_vel + MB;
    } else {
      VL = S.ws_lv + g0;
      VA = S.ws_av + g0;
    }
    const SoaRec SR{S.rec + (size_t)m0 * kMrQ, S.pts + (size_t)m0 * 4 * kPtQ, M};
    const uint32_t ncol = unitColours(S, m0, mb, mn, true, ccount, ccursor, tmp, &s_ncol);
    // ContactSolver.addManifold (src/solver.zig:449-491), records in colour order
    for (uint32_t j = mb + tid; j < mb + mn; j += bs) {
      uint32_t i = S.order[m0 + j];
      const DevManifold& d = ml[m0 + i];
      uint32_t ga = g0 + d.sa, gb = g0 + d.sb;
      V3 pa = f4v(S.ws_pos[ga]), pb = f4v(S.ws_pos[gb]);
      Q4 qa = f4q(S.ws_rot[ga]), qb = f4q(S.ws_rot[gb]);
      SR.r(j, 0) = make_float4(__uint_as_float(d.sa), __uint_as_float(d.sb), d.m.friction,
                               __uint_as_float(d.m.contact_count));
      SR.r(j, 3) = make_float4(__uint_as_float(B.type[ga] == VXP_BODY_DYNAMIC && !B.sleeping[ga]),
                               __uint_as_float(B.type[gb] == VXP_BODY_DYNAMIC && !B.sleeping[gb]),
                               d.m.restitution, 0.0f);
      for (uint32_t k = 0; k < d.m.contact_count; k++) {
        const vxp_contact& c = d.m.contacts[k];
        V3 la{c.local_anchor_a[0], c.local_anchor_a[1], c.local_anchor_a[2]};
        V3 lb{c.local_anchor_b[0], c.local_anchor_b[1], c.local_anchor_b[2]};
        V3 n{c.normal[0], c.normal[1], c.normal[2]};
        V3 wa = add(qrotate(qa, la), pa);
        V3 wb = add(qrotate(qb, lb), pb);
        V3 ad = sub(wb, wa);
        SR.p(j, k, 0) = f4(n, 0.0f);
        SR.p(j, k, 1) = f4(V3{c.point[0], c.point[1], c.point[2]}, 0.0f);
        SR.p(j, k, 2) = f4(sub(wa, pa), 0.0f);
        SR.p(j, k, 3) = f4(sub(wb, pb), 0.0f);
        SR.p(j, k, 4) = f4(V3{0.f, 0.f, 0.f}, c.normal_impulse);
        SR.p(j, k, 5) = f4(V3{0.f, 0.f, 0.f}, c.tangent_impulse[0]);
        SR.p(j, k, 6) = f4(la, c.tangent_impulse[1]);
        SR.p(j, k, 7) = f4(lb, -c.penetration - dot(ad, n));
      }
      if (kLds) {
        VL[d.sa] = S.ws_lv[ga];
        VA[d.sa] = S.ws_av[ga];
        VL[d.sb] = S.ws_lv[gb];
        VA[d.sb] = S.ws_av[gb];
      }
    }
    if (kLds)
      for (uint32_t k = tid; k < bn; k += bs) {
        uint32_t s = ub[k];
        VL[s] = S.ws_lv[g0 + s];
        VA[s] = S.ws_av[g0 + s];
      }
    __syncthreads();
    stampW(ts, w, 9);
    const float h = cfg.h;
    const V3 grav{cfg.gravity[0], cfg.gravity[1], cfg.gravity[2]};
    const float dt_inv = h > kEps ? 1.0f / h : 0.0f;
    for (uint32_t sub_i = 0; sub_i < cfg.substeps; sub_i++) {
      // applyGravity (src/solver.zig:564-605)
      for (uint32_t k = tid; k < bn; k += bs) {
        uint32_t s = ub[k], g = g0 + s;
        if (B.type[g] != VXP_BODY_DYNAMIC) continue;
        V3 lv = add(f4v(VL[s]), scale(grav, h));
        V3 av = f4v(VA[s]);
        float ld = B.ldamp[g], ad = B.adamp[g];
        float ls = len(lv), as = len(av);
        if (ls < 0.5f) ld += 5.0f * (1.0f - (ls / 0.5f));
        if (as < 0.5f) ad += 5.0f * (1.0f - (as / 0.5f));
        float ldec = fmaxf(0.0f, 1.0f - ld * h), adec = fmaxf(0.0f, 1.0f - ad * h);
        lv = scale(lv, ldec);
        av = scale(av, adec);
        if (dot(lv, lv) < 0.001f * 0.001f) lv = V3{0.f, 0.f, 0.f};
        if (dot(av, av) < 0.001f * 0.001f) av = V3{0.f, 0.f, 0.f};
        VL[s] = v4(lv);
        VA[s] = v4(av);
      }
      __syncthreads();
      if (mn) {
        // prepareConstraints (src/solver.zig:608-699); also refreshes the
        // per-manifold copies of body position, mass and world inertia
        for (uint32_t j = mb + tid; j < mb + mn; j += bs) {
          float4 r0 = SR.r(j, 0), r3 = SR.r(j, 3);
          uint32_t sa = __float_as_uint(r0.x), sb = __float_as_uint(r0.y), n = __float_as_uint(r0.w);
          float restitution = r3.z;
          uint32_t ga = g0 + sa, gb = g0 + sb;
          V3 pa = f4v(S.ws_pos[ga]), pb = f4v(S.ws_pos[gb]);
          Q4 qa = f4q(S.ws_rot[ga]), qb = f4q(S.ws_rot[gb]);
          float ima = B.inv_mass[ga], imb = B.inv_mass[gb];
          M3 iia = loadM3(S.ws_ii + (size_t)ga * 9), iib = loadM3(S.ws_ii + (size_t)gb * 9);
          SR.r(j, 1) = f4(pa, ima);
          SR.r(j, 2) = f4(pb, imb);
          SR.r(j, 4) = f4(iia.m[0], iia.m[1], iia.m[2], iia.m[3]);
          SR.r(j, 5) = f4(iia.m[4], iia.m[5], iia.m[6], iia.m[7]);
          SR.r(j, 6) = f4(iia.m[8], iib.m[0], iib.m[1], iib.m[2]);
          SR.r(j, 7) = f4(iib.m[3], iib.m[4], iib.m[5], iib.m[6]);
          SR.r(j, 8) = f4(iib.m[7], iib.m[8], 0.0f, 0.0f);
          V3 lva = f4v(VL[sa]), ava = f4v(VA[sa]), lvb = f4v(VL[sb]), avb = f4v(VA[sb]);
          for (uint32_t k = 0; k < n; k++) {
            float4 q0 = SR.p(j, k, 0), q6 = SR.p(j, k, 6), q7 = SR.p(j, k, 7);
            V3 nrm = f4v(q0);
            V3 wa = add(qrotate(qa, f4v(q6)), pa);
            V3 wb = add(qrotate(qb, f4v(q7)), pb);
            V3 ra = sub(wa, pa), rb = sub(wb, pb);
            V3 pt = scale(add(wa, wb), 0.5f);
            V3 ad = sub(wb, wa);
            float sep = q7.w + dot(ad, nrm);
            V3 rna = cross(ra, nrm), rnb = cross(rb, nrm);
            float k_ = ima + imb;
            k_ += dot(rna, mulVec(iia, rna));
            k_ += dot(rnb, mulVec(iib, rnb));
            float nm = k_ > kEps ? 1.0f / k_ : 0.0f;
            V3 va = add(lva, cross(ava, sub(pt, pa)));
            V3 vb = add(lvb, cross(avb, sub(pt, pb)));
            V3 rel = sub(vb, va);
            float vn = dot(rel, nrm);
            float bias;
            if (sep > 0.0f) bias = sep * dt_inv;
            else bias = fmaxf(cfg.bias_rate * sep, -0.2f);
            if (cfg.enable_restitution && vn < -1.0f) bias = fmaxf(bias, -restitution * vn);
            float tm0 = SR.p(j, k, 1).w, tm1 = SR.p(j, k, 2).w;
            float4 q4 = SR.p(j, k, 4), q5 = SR.p(j, k, 5);
            V3 t1 = f4v(q4), t2 = f4v(q5);
            if (cfg.enable_friction) {
              t1 = computeOrthogonal(nrm);
              t2 = cross(nrm, t1);
              V3 a1 = cross(ra, t1), b1 = cross(rb, t1);
              float k1 = ima + imb;
              k1 += dot(a1, mulVec(iia, a1));
              k1 += dot(b1, mulVec(iib, b1));
              tm0 = k1 > kEps ? 1.0f / k1 : 0.0f;
              V3 a2 = cross(ra, t2), b2 = cross(rb, t2);
              float k2 = ima + imb;
              k2 += dot(a2, mulVec(iia, a2));
              k2 += dot(b2, mulVec(iib, b2));
              tm1 = k2 > kEps ? 1.0f / k2 : 0.0f;
            }
            SR.p(j, k, 0) = f4(nrm, nm);
            SR.p(j, k, 1) = f4(pt, tm0);
            SR.p(j, k, 2) = f4(ra, tm1);
            SR.p(j, k, 3) = f4(rb, bias);
            SR.p(j, k, 4) = f4(t1, q4.w);
            SR.p(j, k, 5) = f4(t2, q5.w);
          }
        }
        __syncthreads();
        if (sub_i == 0) stampW(ts, w, 10);
        // warmStart (src/solver.zig:702-729), substep 0 only
        if (cfg.warm_starting && sub_i == 0) {
          for (uint32_t c = 0; c < ncol; c++) {
            for (uint32_t j = mb + ccount[c] + tid; j < mb + ccount[c + 1]; j += bs)
              warmManifold(cfg.wss, VL, VA, SR, j);
            __syncthreads();
          }
        }
        // solveVelocityConstraints with bias
        for (uint32_t it = 0; it < cfg.vel_iters; it++)
          colourPasses<true>(cfg, VL, VA, SR, ccount, mb, ncol, tid, bs);
      }
      if (sub_i == 0) stampW(ts, w, 11);
      // integratePositions (src/solver.zig:341-356,841-845): the unit's
      // mobile bodies (active, not static, awake)
      for (uint32_t k = tid; k < bn; k += bs) {
        uint32_t s = ub[k], g = g0 + s;
        V3 p = add(f4v(S.ws_pos[g]), scale(f4v(VL[s]), h));
        S.ws_pos[g] = v4(p);
        if (!cfg.disable_rotation) {
          Q4 qq = integrateRotation(f4q(S.ws_rot[g]), f4v(VA[s]), h);
          S.ws_rot[g] = q4(qq);
        }
        if (B.type[g] == VXP_BODY_DYNAMIC)
          storeM3(S.ws_ii + (size_t)g * 9,
                  worldInvInertia(f4q(S.ws_rot[g]), loadM3(B.iil + (size_t)g * 9)));
      }
      __syncthreads();
      // relaxConstraints (src/solver.zig:848-856): hard, unbiased; the bodies'
      // positions and inertia moved in integrate, the point data did not
      if (mn) {
        for (uint32_t j = mb + tid; j < mb + mn; j += bs) {
          float4 r0 = SR.r(j, 0);
          uint32_t ga = g0 + __float_as_uint(r0.x), gb = g0 + __float_as_uint(r0.y);
          M3 iia = loadM3(S.ws_ii + (size_t)ga * 9), iib = loadM3(S.ws_ii + (size_t)gb * 9);
          SR.r(j, 1) = f4(f4v(S.ws_pos[ga]), SR.r(j, 1).w);
          SR.r(j, 2) = f4(f4v(S.ws_pos[gb]), SR.r(j, 2).w);
          SR.r(j, 4) = f4(iia.m[0], iia.m[1], iia.m[2], iia.m[3]);
          SR.r(j, 5) = f4(iia.m[4], iia.m[5], iia.m[6], iia.m[7]);
          SR.r(j, 6) = f4(iia.m[8], iib.m[0], iib.m[1], iib.m[2]);
          SR.r(j, 7) = f4(iib.m[3], iib.m[4], iib.m[5], iib.m[6]);
          SR.r(j, 8) = f4(iib.m[7], iib.m[8], 0.0f, 0.0f);
        }
        __syncthreads();
        colourPasses<false>(cfg, VL, VA, SR, ccount, mb, ncol, tid, bs);
      }
      if (sub_i == 0) stampW(ts, w, 12);
    }
    stampW(ts, w, 13);
    // first position pass (src/solver.zig:860-923); later passes run in
    // k_solve_pos once every unit of the world has finished this one
    if (mn && cfg.pos_iters > 0) {
      float maxpen = 0.0f;
      float minsep = unitPositionPass(cfg, B, S, SR, g0, mb, ncol, ccount, s_red, maxpen);
      unitPositionCombine(ws[w], 0, minsep, maxpen, s_red);
    }
    stampW(ts, w, 14);
    // storeImpulses (src/solver.zig:933-958)
    if (cfg.store_impulses) {
      for (uint32_t j = mb + tid; j < mb + mn; j += bs) {
        uint32_t i = S.order[m0 + j];
        DevManifold& d = ml[m0 + i];
        for (uint32_t k = 0; k < d.m.contact_count; k++) {
          vxp_contact& c = d.m.contacts[k];
          c.normal_impulse = SR.p(j, k, 4).w;
          c.tangent_impulse[0] = SR.p(j, k, 5).w;
          c.tangent_impulse[1] = SR.p(j, k, 6).w;
          c.normal_mass = SR.p(j, k, 0).w;
          c.tangent_mass[0] = SR.p(j, k, 1).w;
          c.tangent_mass[1] = SR.p(j, k, 2).w;
        }
      }
    }
    if (kLds)
      for (uint32_t k = tid; k < bn; k += bs) {
        uint32_t s = ub[k];
        S.ws_lv[g0 + s] = VL[s];
        S.ws_av[g0 + s] = VA[s];
      }
    __syncthreads();
    stampW(ts, w, 15);
  }
}

// Position pass `it` >= 1 for every unit whose world has not exited early:
// the world-wide minimum separation of pass it - 1 is complete when this
// launch starts.  Three rotating slots: pass it reads slot (it - 1) % 3,
// combines into slot it % 3 and clears slot (it + 1) % 3 for pass it + 1.
__global__ __launch_bounds__(256) void k_solve_pos(DevCfg cfg, DevBodies B, WorldState* ws,
                                                   SolverScratch S, const uint32_t* step_ctr,
                                                   uint32_t it) {
  __shared__ uint32_t tmp[256];
  __shared__ uint32_t ccount[kMaxColors + 1];
  __shared__ float s_red[256];
  __shared__ uint32_t s_ncol, s_run;
  const uint32_t tid = threadIdx.x;
  const uint32_t total = step_ctr[kCtrSolveUnits];
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    const uint32_t q = S.uqueue[item];
    const uint32_t w = q / kUnitsMax;
    WorldState& x = ws[w];
    // decided by one thread: pos_done may be set by another unit of the same
    // world during this launch, and then the slot test gives the same answer
    if (tid == 0) {
      uint32_t run = 0;
      if (x.nsolve != 0 && !x.pos_done) {
        if (__uint_as_float(x.pos_min[(it - 1) % 3]) >= -3.0f * cfg.slop) {
          x.pos_done = 1;
        } else {
          run = 1;
          x.pos_run = it + 1;
          x.pos_min[(it + 1) % 3] = 0;
        }
      }
      s_run = run;
    }
    __syncthreads();
    const uint32_t run = s_run;
    __syncthreads();
    if (!run) continue;
    const uint4 ud = S.units[q];
    const uint32_t mb = ud.x, mn = ud.y;
    if (!mn) continue;
    const uint32_t g0 = w * cfg.max_bodies, m0 = x.man_start;
    const SoaRec SR{S.rec + (size_t)m0 * kMrQ, S.pts + (size_t)m0 * 4 * kPtQ, x.man_count};
    const uint32_t ncol = unitColours(S, m0, mb, mn, false, ccount, nullptr, tmp, &s_ncol);
    float maxpen = 0.0f;
    float minsep = unitPositionPass(cfg, B, S, SR, g0, mb, ncol, ccount, s_red, maxpen);
    unitPositionCombine(x, it % 3, minsep, maxpen, s_red);
  }
}

// writeBackBodies (src/solver.zig:246-296,926-930), updateSleepStates
// (src/system.zig:786-792, src/body.zig:415-427) and the solver stats.
__global__ __launch_bounds__(kSolveThreads) void k_solve_final(DevCfg cfg, DevBodies B,
                                                               WorldState* ws, vxp_stats* st,
                                                               SolverScratch S) {
  const uint32_t w = blockIdx.x;
  if (ws[w].skip) return;
  const uint32_t tid = threadIdx.x, bs = blockDim.x;
  const uint32_t MB = cfg.max_bodies, g0 = w * MB;
  stamp(S.tstamp, 4);
  for (uint32_t s = tid; s < MB; s += bs) {
    uint32_t g = g0 + s;
    if (!B.active[g] || B.type[g] == VXP_BODY_STATIC) continue;
    V3 p = f4v(S.ws_pos[g]);
    Q4 q = f4q(S.ws_rot[g]);
    V3 lv = f4v(S.ws_lv[g]), av = f4v(S.ws_av[g]);
    bool fq = isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w);
    if (!finite3(p) || !fq) {
      V3 op = f4v(B.pos[g]);
      p = finite3(op) ? op : V3{0.0f, 2.0f, 0.0f};
      q = Q4{0.f, 0.f, 0.f, 1.f};
      lv = V3{0.f, 0.f, 0.f};
      av = V3{0.f, 0.f, 0.f};
    }
    if (!finite3(lv)) lv = V3{0.f, 0.f, 0.f};
    if (!finite3(av)) av = V3{0.f, 0.f, 0.f};
    float ls = dot(lv, lv), as = dot(av, av);
    if (ls > 50.0f * 50.0f) lv = scale(lv, 50.0f / sqrtf(ls));
    if (as > 30.0f * 30.0f) av = scale(av, 30.0f / sqrtf(as));
    if (ls < 0.002f) lv = V3{0.f, 0.f, 0.f};
    if (as < 0.002f) av = V3{0.f, 0.f, 0.f};
    q = qnormalize(q);
    B.pos[g] = v4(p);
    B.rot[g] = q4(q);
    B.lv[g] = v4(lv);
    B.av[g] = v4(av);
  }
  __syncthreads();
  stamp(S.tstamp, 5);
  if (cfg.enable_sleeping) {
    for (uint32_t s = tid; s < MB; s += bs) {
      uint32_t g = g0 + s;
      if (!B.active[g] || B.type[g] != VXP_BODY_DYNAMIC || B.sleeping[g]) continue;
      V3 lv = f4v(B.lv[g]), av = f4v(B.av[g]);
      if (len(lv) < 0.1f && len(av) < 0.1f) {
        float t = B.sleep_timer[g] + cfg.dt;
        B.sleep_timer[g] = t;
        if (t >= 0.5f) {
          B.sleeping[g] = 1;
          B.lv[g] = make_float4(0.f, 0.f, 0.f, 0.f);
          B.av[g] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
      } else {
        B.sleep_timer[g] = 0.0f;
      }
    }
  }
  stamp(S.tstamp, 6);
  if (tid == 0) {
    st[w].max_penetration = __uint_as_float(ws[w].maxpen);
    st[w].position_iterations_run = ws[w].pos_run;
  }
}

// ---------------------------------------------------------------------------
// collider build: classification, metadata and mass properties
// ---------------------------------------------------------------------------

// Per-brick voxel masks of the brick pool: words 0..7 non-empty, 8..15 corner,
// 16..23 edge, 24..31 inside voxels; bit x + 8y of word z.  Block per brick,
// thread per (y, z) row of 8 voxels.
__global__ __launch_bounds__(64) void k_brick_masks(const uint8_t* vox, uint64_t* mask) {
  __shared__ uint8_t rows[4][64];
  const uint32_t b = blockIdx.x, t = threadIdx.x;  // t = y + 8z
  const uint64_t w = reinterpret_cast<const uint64_t*>(vox + (size_t)b * 512)[t];
  uint8_t occ = 0, cor = 0, edg = 0, ins = 0;
  for (int x = 0; x < 8; x++) {
    const uint32_t ty = (uint32_t)(w >> (8 * x)) & 7u;
    occ |= ty != 0 ? 1u << x : 0u;
    cor |= ty == VXP_VOXEL_CORNER ? 1u << x : 0u;
    edg |= ty == VXP_VOXEL_EDGE ? 1u << x : 0u;
    ins |= ty == VXP_VOXEL_INSIDE ? 1u << x : 0u;
  }
  rows[0][t] = occ;
  rows[1][t] = cor;
  rows[2][t] = edg;
  rows[3][t] = ins;
  __syncthreads();
  if (t < kMaskWords) {
    const uint32_t k = t / 8, z = t % 8;
    uint64_t word = 0;
    for (int y = 0; y < 8; y++) word |= (uint64_t)rows[k][y + 8 * z] << (8 * y);
    mask[(size_t)b * kMaskWords + t] = word;
  }
}

// classifyFromNeighbors (src/collision_voxel.zig:213-235) for every voxel of
// the input bricks: block per brick, thread per voxel.  solid: 8 words per
// brick, bit x + 8y of word z; nbr: the -x, +x, -y, +y, -z, +z neighbour
// brick of each brick (-1: none, all empty).
__global__ __launch_bounds__(512) void k_collider_classify(const uint64_t* solid,
                                                           const int32_t* nbr, uint8_t* out) {
  const uint32_t b = blockIdx.x, i = threadIdx.x;
  const int x = i & 7, y = (i >> 3) & 7, z = i >> 6;
  auto at = [&](int32_t bb, int xx, int yy, int zz) -> bool {
    return bb >= 0 && ((solid[(size_t)bb * 8 + zz] >> (xx + 8 * yy)) & 1ull) != 0;
  };
  uint8_t v = 0;
  if (at((int32_t)b, x, y, z)) {
    const int32_t* nb = nbr + (size_t)b * 6;
    const bool side[6] = {x > 0 ? at((int32_t)b, x - 1, y, z) : at(nb[0], 7, y, z),
                          x < 7 ? at((int32_t)b, x + 1, y, z) : at(nb[1], 0, y, z),
                          y > 0 ? at((int32_t)b, x, y - 1, z) : at(nb[2], x, 7, z),
                          y < 7 ? at((int32_t)b, x, y + 1, z) : at(nb[3], x, 0, z),
                          z > 0 ? at((int32_t)b, x, y, z - 1) : at(nb[4], x, y, 7),
                          z < 7 ? at((int32_t)b, x, y, z + 1) : at(nb[5], x, y, 0)};
    uint32_t empty = 0;
    for (int k = 0; k < 6; k++) empty |= side[k] ? 0u : 1u << k;
    v = classifyByte(empty);
  }
  out[(size_t)b * 512 + i] = v;
}

// CollisionBrick.recomputeMetadata (src/collider.zig:90-133) per brick: solid
// count, voxel bounds and the local AABB of its non-empty voxels (minima and
// maxima, exact in any order).  Block per brick, thread per voxel.
__global__ __launch_bounds__(512) void k_collider_meta(const int32_t* pos, const uint8_t* vox,
                                                       float vs, uint32_t* cnt, int32_t* bounds,
                                                       float* box) {
  __shared__ int32_t ib[6][512];
  __shared__ float fb[6][512];
  __shared__ uint32_t n;
  const uint32_t b = blockIdx.x, i = threadIdx.x;
  if (i == 0) n = 0;
  int32_t v[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
  float f[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  __syncthreads();
  if (vox[(size_t)b * 512 + i] & 7u) {
    atomicAdd(&n, 1u);
    const int c[3] = {(int)(i & 7), (int)((i >> 3) & 7), (int)(i >> 6)};
    for (int k = 0; k < 3; k++) {
      const int32_t vx = pos[3 * b + k] * 8 + c[k];
      v[k] = vx;
      v[3 + k] = vx + 1;
      const float lp = ((float)pos[3 * b + k] * 8.0f + (float)c[k]) * vs;
      f[k] = lp;
      f[3 + k] = lp + vs;
    }
  }
  for (int k = 0; k < 6; k++) {
    ib[k][i] = v[k];
    fb[k][i] = f[k];
  }
  __syncthreads();
  for (uint32_t d = 256; d > 0; d >>= 1) {
    if (i < d)
      for (int k = 0; k < 3; k++) {
        ib[k][i] = min(ib[k][i], ib[k][i + d]);
        ib[3 + k][i] = max(ib[3 + k][i], ib[3 + k][i + d]);
        fb[k][i] = fminf(fb[k][i], fb[k][i + d]);
        fb[3 + k][i] = fmaxf(fb[3 + k][i], fb[3 + k][i + d]);
      }
    __syncthreads();
  }
  if (i < 6) {
    bounds[(size_t)b * 6 + i] = ib[i][0];
    box[(size_t)b * 6 + i] = fb[i][0];
  }
  if (i == 0) cnt[b] = n;
}

// computeMassProperties (src/body.zig:462-519) with getCenterOfMass
// (src/collider.zig:376-403): the voxels are summed one by one, bricks in the
// order given (ascending (z, y, x) position), so the float sums follow that
// order exactly; one lane accumulates while the block stages each brick's
// voxel bytes in LDS.  out: inv_mass, inverse inertia (9), centre of mass (3).
// computeMassProperties (src/body.zig:462-519), one workgroup per job (a
// collider at a density), all jobs of a batch in one launch.  Every float sum
// runs in the canonical order (bricks as given, voxels by index inside a
// brick), one accumulator per lane: the 64 lanes compute the per-voxel terms of
// a brick into LDS, then lane k adds term k of each solid voxel in order.  The
// terms are the same expressions as a sequential loop's, so the sums are too.
__global__ __launch_bounds__(64) void k_collider_mass(const MassJob* jobs, const int32_t* pos_all,
                                                      const uint8_t* vox_all, float* out_all) {
  __shared__ float term[6][512];
  __shared__ uint64_t occ[8];  // solid voxels of the staged brick, bit i % 64 of word i / 64
  __shared__ float s_com[3];
  const MassJob job = jobs[blockIdx.x];
  const int32_t* pos = pos_all + (size_t)job.brick_off * 3;
  const uint8_t* vox = vox_all + (size_t)job.brick_off * 512;
  float* out = out_all + (size_t)blockIdx.x * kMassOut;
  const uint32_t nb = job.nb, solid = job.solid;
  const float vs = job.vs, density = job.density;
  const uint32_t lane = threadIdx.x;
  const float vv = vs * vs * vs;
  const float vm = vv * density;
  const float total = vm * (float)solid;
  const float half = vs * 0.5f;
  auto voxelPos = [&](uint32_t b, uint32_t i) {
    const int32_t vx = pos[3 * b] * 8 + (int32_t)(i & 7), vy = pos[3 * b + 1] * 8 + (int32_t)((i >> 3) & 7),
                  vz = pos[3 * b + 2] * 8 + (int32_t)(i >> 6);
    V3 l{(float)vx * vs + half, (float)vy * vs + half, (float)vz * vs + half};
    return add(l, V3{vs * 0.5f, vs * 0.5f, vs * 0.5f});
  };
  // stage brick b: the occupancy mask, and pass 0 (positions) or 1 (inertia terms)
  auto stage = [&](uint32_t b, int pass, V3 com) {
    __syncthreads();
    const uint8_t* v = vox + (size_t)b * 512;
    for (uint32_t w = 0; w < 8; w++) {
      const uint32_t i = w * 64 + lane;
      const bool on = (v[i] & 7u) != 0;
      const uint64_t m = __ballot(on);
      if (lane == 0) occ[w] = m;
      if (!on) continue;
      const V3 p = voxelPos(b, i);
      if (pass == 0) {
        term[0][i] = p.x;
        term[1][i] = p.y;
        term[2][i] = p.z;
      } else {
        const V3 r = sub(p, com);
        const float r2 = dot(r, r);
        term[0][i] = vm * (r2 - r.x * r.x);
        term[1][i] = vm * (r2 - r.y * r.y);
        term[2][i] = vm * (r2 - r.z * r.z);
        term[3][i] = vm * r.x * r.y;
        term[4][i] = vm * r.x * r.z;
        term[5][i] = vm * r.y * r.z;
      }
    }
    __syncthreads();
  };
  // pass 0: sum of positions (lanes 0..2) and the count
  float acc = 0.0f;
  uint32_t count = 0;
  for (uint32_t b = 0; b < nb; b++) {
    stage(b, 0, V3{0.f, 0.f, 0.f});
    for (uint32_t w = 0; w < 8; w++) {
      uint64_t m = occ[w];
      count += (uint32_t)__popcll(m);
      if (lane < 3)
        while (m) {
          const uint32_t i = w * 64 + (uint32_t)__builtin_ctzll(m);
          m &= m - 1;
          acc = acc + term[lane][i];
        }
    }
  }
  if (lane < 3) s_com[lane] = acc * (1.0f / (float)count);  // Vec3.scale by 1/count
  __syncthreads();
  const V3 com{s_com[0], s_com[1], s_com[2]};
  // pass 1: inertia, I00 I11 I22 added and I01 I02 I12 subtracted (lanes 0..5)
  acc = 0.0f;
  for (uint32_t b = 0; b < nb; b++) {
    stage(b, 1, com);
    for (uint32_t w = 0; w < 8; w++) {
      uint64_t m = occ[w];
      if (lane < 6)
        while (m) {
          const uint32_t i = w * 64 + (uint32_t)__builtin_ctzll(m);
          m &= m - 1;
          acc = lane < 3 ? acc + term[lane][i] : acc - term[lane][i];
        }
    }
  }
  __syncthreads();
  if (lane < 6) term[lane][0] = acc;
  __syncthreads();
  if (lane == 0) {
    const float d0 = term[0][0], d1 = term[1][0], d2 = term[2][0];
    const float o01 = term[3][0], o02 = term[4][0], o12 = term[5][0];
    M3 m{{d0, o01, o02, o01, d1, o12, o02, o12, d2}};
    const M3 inv = minverse(m);
    out[0] = 1.0f / total;
    for (int k = 0; k < 9; k++) out[1 + k] = inv.m[k];
    out[10] = com.x;
    out[11] = com.y;
    out[12] = com.z;
  }
}

// End of update(): substeps_this_frame and updateBodyCounts (src/system.zig:
// 527-533).  accumulated_time keeps the value k_update_begin recorded right
// after `accumulated_time += dt`, before any step (src/system.zig:512-513).
__global__ void k_update_end(DevCfg cfg, DevBodies B, vxp_stats* st, uint32_t n_worlds,
                             uint32_t substeps_this_frame) {
  uint32_t w = blockIdx.x;
  if (w >= n_worlds) return;
  __shared__ uint32_t sl, ac;
  if (threadIdx.x == 0) sl = ac = 0;
  __syncthreads();
  uint32_t g0 = w * cfg.max_bodies, a = 0, s_ = 0;
  for (uint32_t s = threadIdx.x; s < cfg.max_bodies; s += blockDim.x) {
    uint32_t g = g0 + s;
    if (!B.active[g] || B.type[g] != VXP_BODY_DYNAMIC) continue;
    if (B.sleeping[g]) s_++; else a++;
  }
  if (a) atomicAdd(&ac, a);
  if (s_) atomicAdd(&sl, s_);
  __syncthreads();
  if (threadIdx.x == 0) {
    st[w].sleeping_bodies = sl;
    st[w].active_bodies = ac;
    st[w].substeps_this_frame = substeps_this_frame;
  }
}

// PhysicsStats.reset (src/system.zig:129-154) at the start of update()
__global__ void k_update_begin(vxp_stats* st, uint32_t n_worlds, float acc) {
  uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_worlds) return;
  vxp_stats& s = st[w];
  uint32_t tb = s.total_bodies, db = s.dynamic_bodies, sb = s.static_bodies;
  vxp_stats z = {};
  z.total_bodies = tb;
  z.dynamic_bodies = db;
  z.static_bodies = sb;
  z.accumulated_time = acc;
  s = z;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
// Threads of the block-per-world scans over body slots (step begin and end,
// sweep axis): up to four slots per thread, power of two in 256..1024.
static inline uint32_t slotThreads(uint32_t max_bodies) {
  return max_bodies > 2048 ? 1024u : max_bodies > 1024 ? 512u : 256u;
}

void launch_update_begin(hipStream_t s, vxp_stats* st, uint32_t nw, float acc) {
  k_update_begin<<<cdiv(nw, 256), 256, 0, s>>>(st, nw, acc);
}
void launch_step_begin(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       vxp_stats* st, uint32_t nw, uint32_t* step_ctr) {
  k_step_begin<<<nw, slotThreads(cfg.max_bodies), 0, s>>>(cfg, B, ws, st, nw, step_ctr);
}
// ---------------------------------------------------------------------------
// Broad-phase trees (pair order VXP_PAIR_ORDER_TREE)
// ---------------------------------------------------------------------------
// BroadPhase.updateBody for the moved bodies of one world, in slot order
// (src/system.zig:570-583): the block compacts the moved slots 256 at a time
// and thread 0 re-inserts them; the dynamic tree is staged in LDS when it fits.
__global__ __launch_bounds__(256) void k_tree_update(DevCfg cfg, DevBodies B, const WorldState* ws,
                                                     TreeSet T, const uint8_t* moved,
                                                     size_t lds_bytes) {
  extern __shared__ TreeNode tn[];
  __shared__ uint32_t list[256], wcnt[4], n_list;
  __shared__ TreeHead head;
  const uint32_t w = blockIdx.x, tid = threadIdx.x;
  if (ws[w].skip) return;
  Tree g = T.tree(w, 1);
  const bool L = (size_t)T.nc * sizeof(TreeNode) <= lds_bytes;
  if (tid == 0) head = *g.h;
  __syncthreads();
  const uint32_t len = head.len;
  if (L)
    for (uint32_t i = tid; i < len; i += blockDim.x) tn[i] = g.n[i];
  Tree t{L ? tn : g.n, &head, T.nc};
  const uint32_t lane = tid & (kWave - 1), wave = tid / kWave;
  for (uint32_t b0 = 0; b0 < cfg.max_bodies; b0 += 256) {
    const uint32_t slot = b0 + tid;
    const size_t gs = (size_t)w * cfg.max_bodies + slot;
    const bool m = slot < cfg.max_bodies && moved[gs];
    const uint64_t bal = __ballot(m);
    if (lane == 0) wcnt[wave] = __popcll(bal);
    __syncthreads();
    if (m) {
      uint32_t at = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
      for (uint32_t k = 0; k < wave; k++) at += wcnt[k];
      list[at] = slot;
    }
    if (tid == 0) n_list = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    __syncthreads();
    if (tid == 0)
      for (uint32_t k = 0; k < n_list; k++) {
        const size_t gb = (size_t)w * cfg.max_bodies + list[k];
        const float4 lo = B.fat_lo[gb], hi = B.fat_hi[gb];
        const float flo[3] = {lo.x, lo.y, lo.z}, fhi[3] = {hi.x, hi.y, hi.z};
        treeMoveLeaf(t, T.leaf[gb], flo, fhi);
      }
    __syncthreads();
  }
  if (L)  // re-insertions free a node before they allocate one, but copy what head.len says
    for (uint32_t i = tid; i < head.len; i += blockDim.x) g.n[i] = tn[i];
  if (tid == 0) *g.h = head;
}

// findOverlappingPairs + queryAgainstTree (src/broad_phase.zig:488-526,606-619)
// for one dynamic-tree node: position (w, q, i), q = 0 self pairs (partners
// of higher id), q = 1 against the static tree.  Offsets come from an
// exclusive scan of the counts in (world, q, node) order, which is the
// reference's emission order.
template <bool kWrite>
__global__ __launch_bounds__(256) void k_tree_pairs(WorldState* ws, TreeSet T, uint32_t nw,
                                                    uint32_t* counts, const uint32_t* offs,
                                                    uint64_t* out, uint32_t* pos,
                                                    uint32_t id_bits) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per_w = 2 * (size_t)T.nc;
  if (gid >= nw * per_w) return;
  const uint32_t w = (uint32_t)(gid / per_w), q = (uint32_t)((gid % per_w) / T.nc),
                 i = (uint32_t)(gid % T.nc);
  const Tree d = T.tree(w, 1);
  uint32_t n = 0;
  if (i < d.h->len) {
    const TreeNode me = d.n[i];
    if (me.height >= 0 && treeLeaf(me)) {
      const Tree o = T.tree(w, q == 0 ? 1 : 0);
      const size_t base = kWrite ? offs[gid] : 0;
      const bool ok = treeQuery(o.n, o.h->root, me.lo, me.hi, [&](uint32_t other) {
        if (q == 1 || other > me.user) {
          if (kWrite) {
            const uint64_t a = min(me.user, other), b = max(me.user, other);
            out[base + n] = ((uint64_t)w << (2 * id_bits)) | (a << id_bits) | b;
            pos[base + n] = (uint32_t)(base + n);
          }
          n++;
        }
      });
      if (!ok) atomicOr(&ws[w].error, kErrTreeStack);
    }
  }
  if (!kWrite) counts[gid] = n;
}

__global__ void k_pair_awake(DevCfg cfg, DevBodies B, const uint64_t* keys, const uint32_t* perm,
                             uint32_t n, uint32_t* awake) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > n) return;
  if (p == n) {
    awake[n] = 0;
    return;
  }
  const uint64_t k = keys[p];
  const uint32_t w = keyWorld(k);
  const int32_t sa = B.idmap[(size_t)w * cfg.id_cap + keyA(k)];
  const int32_t sb = B.idmap[(size_t)w * cfg.id_cap + keyB(k)];
  const bool asl = sa >= 0 && B.sleeping[(size_t)w * cfg.max_bodies + sa];
  const bool bsl = sb >= 0 && B.sleeping[(size_t)w * cfg.max_bodies + sb];
  awake[perm ? perm[p] : p] = !(asl && bsl);
}

__global__ void k_pair_drop(const uint64_t* keys, const uint32_t* perm, uint32_t n,
                            const uint32_t* awake, const uint32_t* rank, uint32_t cap,
                            uint8_t* drop) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  // first pair of this world: its segment starts at the same index in either order
  const uint32_t w = keyWorld(keys[p]);
  uint32_t lo = 0, hi = p;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keyWorld(keys[mid]) < w) lo = mid + 1; else hi = mid;
  }
  const uint32_t t = perm ? perm[p] : p;
  drop[p] = awake[t] && rank[t] - rank[lo] >= cap;
}

// Per world, the choices of the thread-pool narrow phase (src/system.zig:618-638,
// src/narrow_phase.zig:1546): thr_cut = one past the slot of the 256th active
// body (the narrow phase sees the slots below it; all of them when fewer are
// active), thr_par = the parallel branch: at least 8 pairs reach the narrow
// phase (not both asleep, within the cap) and the pool has 2+ threads.  One
// wave per world; awake/rank: the not-both-asleep flags and their exclusive
// scan (n + 1 entries) in the narrow phase's order.
__global__ void k_thread_mode(DevCfg cfg, DevBodies B, WorldState* ws, const uint64_t* keys,
                              uint32_t n, const uint32_t* rank, uint32_t cap) {
  const uint32_t w = blockIdx.x, lane = threadIdx.x, MB = cfg.max_bodies;
  uint32_t run = 0, cut = MB;
  for (uint32_t base = 0; base < MB; base += kWave) {
    const uint32_t sl = base + lane;
    const bool a = sl < MB && B.active[(size_t)w * MB + sl];
    const uint64_t m = __ballot(a);
    const uint32_t c = (uint32_t)__popcll(m);
    if (run + c >= 256u) {
      const uint32_t below = (uint32_t)__popcll(m & laneMaskLt());
      const uint64_t hit = __ballot(a && below == 255u - run);
      cut = base + (uint32_t)(__ffsll((long long)hit) - 1) + 1u;
      break;
    }
    run += c;
  }
  uint32_t np = 0;
  if (n) {
    // the world's pairs: the same index range in either pair order
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (keyWorld(keys[mid]) < w) lo = mid + 1; else hi = mid;
    }
    uint32_t e = lo, top = n;
    while (e < top) {
      const uint32_t mid = (e + top) >> 1;
      if (keyWorld(keys[mid]) <= w) e = mid + 1; else top = mid;
    }
    np = rank[e] - rank[lo];
    if (cap) np = min(np, cap);
  }
  if (lane == 0) {
    ws[w].thr_cut = cut;
    ws[w].thr_par = np >= 8u && cfg.thread_multi;
  }
}

__global__ void k_gather_flags(const uint32_t* in, const uint32_t* idx, uint32_t n, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
  if (i == n) out[n] = 0;
}

__global__ void k_invert(const uint32_t* perm, uint32_t n, uint32_t* inv) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) inv[perm[p]] = p;
}

// Writes *a and *b (when given), then seq, into host memory that the host
// polls; the system-scope release orders the words before the number.
__global__ void k_publish(uint32_t* host, const uint32_t* a, const uint32_t* b, uint32_t seq) {
  if (threadIdx.x != 0) return;
  host[1] = a ? *a : 0u;
  host[2] = b ? *b : 0u;
  __hip_atomic_store(&host[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
void launch_publish(hipStream_t s, uint32_t* host, const uint32_t* a, const uint32_t* b,
                    uint32_t seq) {
  k_publish<<<1, 64, 0, s>>>(host, a, b, seq);
}
// host[1..4] = *w[0..3] (0 for a null word), then host[0] = seq.
__global__ void k_publish4(uint32_t* host, const uint32_t* w0, const uint32_t* w1,
                           const uint32_t* w2, const uint32_t* w3, uint32_t seq) {
  if (threadIdx.x != 0) return;
  host[1] = w0 ? *w0 : 0u;
  host[2] = w1 ? *w1 : 0u;
  host[3] = w2 ? *w2 : 0u;
  host[4] = w3 ? *w3 : 0u;
  __hip_atomic_store(&host[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
void launch_publish4(hipStream_t s, uint32_t* host, const uint32_t* w0, const uint32_t* w1,
                     const uint32_t* w2, const uint32_t* w3, uint32_t seq) {
  k_publish4<<<1, 64, 0, s>>>(host, w0, w1, w2, w3, seq);
}
// Device-side sizes of a recorded step: out[0] = the probe's work items
// (wi_off[n - 1] + counts[n - 1], 0 without pairs), read after the chunk scan;
// out[1] = the candidates (wi_cbase[out[0]]), read after the candidate scan.
__global__ void k_rec_total_wi(const uint32_t* wi_off, const uint32_t* counts, const uint32_t* npairs,
                               uint32_t* out) {
  if (threadIdx.x != 0) return;
  const uint32_t n = *npairs;
  out[0] = n ? wi_off[n - 1] + counts[n - 1] : 0u;
}
// out[2] = the manifolds a step carries over unchanged: every manifold of the
// worlds whose step is skipped (all bodies asleep), read after the step begins.
__global__ __launch_bounds__(1024) void k_rec_carried(const WorldState* ws, uint32_t nw,
                                                      uint32_t* out) {
  __shared__ uint32_t s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  uint32_t c = 0;
  for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x)
    if (ws[w].skip) c += ws[w].prev_count;
  c = waveSum(c);
  if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(&s_sum, c);
  __syncthreads();
  if (threadIdx.x == 0) out[2] = s_sum;
}
void launch_rec_carried(hipStream_t s, const WorldState* ws, uint32_t nw, uint32_t* out) {
  k_rec_carried<<<1, 1024, 0, s>>>(ws, nw, out);
}
// The pair-key sort of a recorded step runs both ways, each on a gated count:
// out[3] = pairs for the device-wide sort (a world holds more than `limit`
// pairs), out[4] = pairs for the per-world sort (otherwise), the other 0.
__global__ void k_rec_gate(const uint32_t* npairs, const uint32_t* maxw, uint32_t limit,
                           uint32_t* out) {
  if (threadIdx.x != 0) return;
  const bool big = *maxw > limit;
  out[3] = big ? *npairs : 0u;
  out[4] = big ? 0u : *npairs;
}
void launch_rec_gate(hipStream_t s, const uint32_t* npairs, const uint32_t* maxw, uint32_t limit,
                     uint32_t* out) {
  k_rec_gate<<<1, 64, 0, s>>>(npairs, maxw, limit, out);
}
__global__ void k_rec_total_cand(const uint32_t* wi_cbase, uint32_t* out) {
  if (threadIdx.x == 0) out[1] = wi_cbase[out[0]];
}
void launch_rec_total_wi(hipStream_t s, const uint32_t* wi_off, const uint32_t* counts,
                         const uint32_t* npairs, uint32_t* out) {
  k_rec_total_wi<<<1, 64, 0, s>>>(wi_off, counts, npairs, out);
}
void launch_rec_total_cand(hipStream_t s, const uint32_t* wi_cbase, uint32_t* out) {
  k_rec_total_cand<<<1, 64, 0, s>>>(wi_cbase, out);
}

void launch_tree_update(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const WorldState* ws,
                        const TreeSet& T, const uint8_t* moved, uint32_t nw, size_t lds_limit) {
  const size_t need = (size_t)T.nc * sizeof(TreeNode);
  const size_t lds = need <= lds_limit ? need : 0;
  k_tree_update<<<nw, 256, lds, s>>>(cfg, B, ws, T, moved, lds);
}
void launch_tree_pairs(hipStream_t s, WorldState* ws, const TreeSet& T, uint32_t nw,
                       uint32_t* counts, const uint32_t* offs, uint64_t* out, uint32_t* pos,
                       uint32_t id_bits) {
  const size_t n = (size_t)nw * 2 * T.nc;
  const uint32_t blocks = (uint32_t)((n + 255) / 256);
  if (out)
    k_tree_pairs<true><<<blocks, 256, 0, s>>>(ws, T, nw, counts, offs, out, pos, id_bits);
  else
    k_tree_pairs<false><<<blocks, 256, 0, s>>>(ws, T, nw, counts, offs, out, pos, id_bits);
}
void launch_pair_awake(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const uint64_t* keys,
                       const uint32_t* perm, uint32_t n, uint32_t* awake) {
  k_pair_awake<<<cdiv(n + 1, 256), 256, 0, s>>>(cfg, B, keys, perm, n, awake);
}
void launch_pair_drop(hipStream_t s, const uint64_t* keys, const uint32_t* perm, uint32_t n,
                      const uint32_t* awake, const uint32_t* rank, uint32_t cap, uint8_t* drop) {
  if (n) k_pair_drop<<<cdiv(n, 256), 256, 0, s>>>(keys, perm, n, awake, rank, cap, drop);
}
void launch_thread_mode(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                        const uint64_t* keys, uint32_t n, const uint32_t* rank, uint32_t cap,
                        uint32_t nw) {
  k_thread_mode<<<nw, kWave, 0, s>>>(cfg, B, ws, keys, n, rank, cap);
}
void launch_gather_flags(hipStream_t s, const uint32_t* in, const uint32_t* idx, uint32_t n,
                         uint32_t* out) {
  k_gather_flags<<<cdiv(n + 1, 256), 256, 0, s>>>(in, idx, n, out);
}
void launch_invert(hipStream_t s, const uint32_t* perm, uint32_t n, uint32_t* inv) {
  if (n) k_invert<<<cdiv(n, 256), 256, 0, s>>>(perm, n, inv);
}

void launch_fat_aabb(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const DevColliders& C,
                     const WorldState* ws, uint32_t n_slots, uint8_t* moved) {
  k_fat_aabb<<<cdiv(n_slots, 256), 256, 0, s>>>(cfg, B, C, ws, n_slots, moved);
}
void launch_sweep_keys(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       uint64_t* keys, uint32_t* vals, uint32_t n_slots, uint32_t nw) {
  k_world_axis<<<nw, slotThreads(cfg.max_bodies), 0, s>>>(cfg, B, ws, nw);
  k_sweep_keys<<<cdiv(n_slots, 256), 256, 0, s>>>(cfg, B, ws, keys, vals, n_slots);
}
bool launch_sweep_sort(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       uint32_t* order, uint32_t nw) {
  const uint32_t MB = cfg.max_bodies;
  if (MB > 8192) return false;
  k_world_axis<<<nw, slotThreads(cfg.max_bodies), 0, s>>>(cfg, B, ws, nw);
  if (MB <= 1024) k_sweep_sort<256, 4><<<nw, 256, 0, s>>>(cfg, B, ws, order);
  else if (MB <= 2048) k_sweep_sort<256, 8><<<nw, 256, 0, s>>>(cfg, B, ws, order);
  else if (MB <= 4096) k_sweep_sort<1024, 4><<<nw, 1024, 0, s>>>(cfg, B, ws, order);
  else if (MB <= 5120) k_sweep_sort<1024, 5><<<nw, 1024, 0, s>>>(cfg, B, ws, order);
  else k_sweep_sort<1024, 8><<<nw, 1024, 0, s>>>(cfg, B, ws, order);
  return true;
}
void launch_sweep_count(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const WorldState* ws,
                        const uint32_t* order, uint32_t n_slots, float4* slo, float4* shi,
                        uint32_t* send, uint32_t* wcnt) {
  k_sweep_gather<<<cdiv(n_slots, 256), 256, 0, s>>>(cfg, B, order, n_slots, slo, shi);
  k_sweep_count<<<cdiv(n_slots, 256), 256, 0, s>>>(cfg, ws, slo, shi, n_slots, send, wcnt);
}
uint32_t sweep_waves() { return kSweepWaves; }

bool launch_lbvh_sort(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                      uint32_t* order, uint32_t nw) {
  const uint32_t MB = cfg.max_bodies;
  if (MB > 8192) return false;
  k_world_axis<<<nw, slotThreads(cfg.max_bodies), 0, s>>>(cfg, B, ws, nw);
  if (MB <= 1024) k_lbvh_sort<256, 4><<<nw, 256, 0, s>>>(cfg, B, ws, order);
  else if (MB <= 2048) k_lbvh_sort<256, 8><<<nw, 256, 0, s>>>(cfg, B, ws, order);
  else if (MB <= 4096) k_lbvh_sort<1024, 4><<<nw, 1024, 0, s>>>(cfg, B, ws, order);
  else if (MB <= 5120) k_lbvh_sort<1024, 5><<<nw, 1024, 0, s>>>(cfg, B, ws, order);
  else k_lbvh_sort<1024, 8><<<nw, 1024, 0, s>>>(cfg, B, ws, order);
  return true;
}
void launch_lbvh_keys(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                      uint64_t* keys, uint32_t* vals, uint32_t n_slots, uint32_t nw) {
  k_world_axis<<<nw, slotThreads(cfg.max_bodies), 0, s>>>(cfg, B, ws, nw);
  k_lbvh_keys<<<cdiv(n_slots, 256), 256, 0, s>>>(cfg, B, ws, keys, vals, n_slots);
}
void launch_lbvh_build(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const WorldState* ws,
                       const uint32_t* order, uint32_t n_slots, const LbvhBufs& L) {
  const uint32_t blocks = cdiv(n_slots, 256);
  k_lbvh_gather<<<blocks, 256, 0, s>>>(cfg, B, ws, order, n_slots, L.llo, L.lhi, L.code);
  k_lbvh_build<<<blocks, 256, 0, s>>>(cfg, ws, n_slots, L.code, L.child, L.range, L.iparent,
                                      L.lparent, L.flag);
  if (cfg.max_bodies <= 8192)
    k_lbvh_refit_wg<<<n_slots / cfg.max_bodies, slotThreads(cfg.max_bodies),
                      (size_t)cfg.max_bodies * 4, s>>>(cfg, ws, L.llo, L.lhi, L.child, L.range,
                                                        L.iparent, L.lparent, L.ilo, L.ihi);
  else
    k_lbvh_refit<<<blocks, 256, 0, s>>>(cfg, ws, n_slots, L.llo, L.lhi, L.child, L.range,
                                        L.iparent, L.lparent, L.flag, L.ilo, L.ihi);
}
void launch_lbvh_pairs(hipStream_t s, const DevCfg& cfg, WorldState* ws, uint32_t n_slots,
                       const LbvhBufs& L, uint32_t* cnt, const uint32_t* off, uint64_t* out,
                       uint32_t id_bits, bool narrow) {
  const uint32_t blocks = cdiv(n_slots, 256);
  if (out)
    k_lbvh_pairs<false><<<blocks, 256, 0, s>>>(cfg, ws, n_slots, L.llo, L.lhi, L.ilo, L.ihi, cnt,
                                               off, out, id_bits, narrow);
  else
    k_lbvh_pairs<true><<<blocks, 256, 0, s>>>(cfg, ws, n_slots, L.llo, L.lhi, L.ilo, L.ihi, cnt,
                                              off, out, id_bits, narrow);
}

// Pairs each world emitted (from the scan of per-leaf counts, LBVH) and the
// largest world's count: one block.
__global__ void k_world_pairs(WorldState* ws, uint32_t nw, const uint32_t* leaf_off, uint32_t MB,
                              uint32_t* max_out) {
  __shared__ uint32_t s_m;
  if (threadIdx.x == 0) s_m = 0;
  __syncthreads();
  uint32_t m = 0;
  for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) {
    if (leaf_off) ws[w].np_emit = leaf_off[(size_t)(w + 1) * MB] - leaf_off[(size_t)w * MB];
    m = max(m, ws[w].np_emit);
  }
  atomicMax(&s_m, m);
  __syncthreads();
  if (threadIdx.x == 0) *max_out = s_m;
}

// The pair keys of each world sorted in one block (the broad phase emits each
// world's packed keys (world, a, b) contiguously, worlds in ascending order):
// the world's segment found by binary search on the world field, its
// 2 * id_bits key bits radix-sorted in LDS, written as canonical keys at the
// same positions.  Requires every world's segment to fit kBS * kIPT.
template <uint32_t kBS, uint32_t kIPT, class R>
__global__ __launch_bounds__(kBS) void k_segment_sort(const R* raw, uint32_t n, uint32_t id_bits,
                                                      uint64_t* keys, const uint32_t* n_dev) {
  n = devCount(n, n_dev);
  __shared__ uint32_t s_k[kBS * kIPT];
  __shared__ uint32_t s_wcnt[2 * (kBS / 64) * kRadixBins], s_run[2 * kRadixBins];
  const uint32_t w = blockIdx.x, sh = 2 * id_bits;
  auto lower = [&](uint32_t x) {  // first position whose world is >= x
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((uint32_t)((uint64_t)raw[mid] >> sh) < x) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  const uint32_t b = lower(w), e = lower(w + 1);
  const uint32_t cnt = min(e - b, kBS * kIPT);  // the host sized the block for the largest world
  if (!cnt) return;
  for (uint32_t d = threadIdx.x; d < 2 * (kBS / 64) * kRadixBins; d += kBS) s_wcnt[d] = 0;
  const uint32_t m = sh >= 32 ? 0xFFFFFFFFu : (1u << sh) - 1u;
  uint32_t k[kIPT], v[kIPT];
#pragma unroll
  for (uint32_t i = 0; i < kIPT; i++) {
    const uint32_t j = i * kBS + threadIdx.x;
    k[i] = j < cnt ? (uint32_t)raw[b + j] & m : 0u;
    v[i] = 0;
  }
  blockRadixSort<kBS, kIPT, false>(k, v, cnt, sh, s_k, nullptr, s_wcnt, s_run);
  const uint32_t im = (1u << id_bits) - 1u;
#pragma unroll
  for (uint32_t i = 0; i < kIPT; i++) {
    const uint32_t j = i * kBS + threadIdx.x;
    if (j < cnt) keys[b + j] = makeKey(w, k[i] >> id_bits, k[i] & im);
  }
}

// Packed (world, a, b) keys -> the canonical (world << 44 | a << 22 | b).
__global__ void k_expand_keys(uint64_t* keys, uint32_t n, uint32_t id_bits, const uint32_t* n_dev) {
  n = devCount(n, n_dev);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t k = keys[i], m = (1ull << id_bits) - 1ull;
    keys[i] = makeKey((uint32_t)(k >> (2 * id_bits)), (uint32_t)((k >> id_bits) & m), (uint32_t)(k & m));
  }
}
// Blocks of a launch over n items (per items per block): all of them, or, for
// a launch whose count is on the device (sized for a capacity n, as in a
// recorded step), at most `most` blocks that loop over the items, so a large
// capacity costs no idle blocks.
static inline uint32_t gridFor(uint32_t n, uint32_t per, const uint32_t* n_dev, uint32_t most) {
  const uint32_t all = cdiv(n, per);
  return n_dev ? std::min(all, most) : all;
}
constexpr uint32_t kLoopBlocks = 16384;  // 64 blocks per CU

void launch_expand_keys(hipStream_t s, uint64_t* keys, uint32_t n, uint32_t id_bits,
                        const uint32_t* n_dev) {
  if (n) k_expand_keys<<<gridFor(n, 256, n_dev, kLoopBlocks), 256, 0, s>>>(keys, n, id_bits, n_dev);
}
// The same from packed keys of 32 bits into a separate array.
__global__ void k_expand_keys32(const uint32_t* in, uint64_t* keys, uint32_t n, uint32_t id_bits,
                                const uint32_t* n_dev) {
  n = devCount(n, n_dev);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t k = in[i], m = (1u << id_bits) - 1u;
    keys[i] = makeKey(k >> (2 * id_bits), (k >> id_bits) & m, k & m);
  }
}
void launch_expand_keys32(hipStream_t s, const uint32_t* in, uint64_t* keys, uint32_t n,
                          uint32_t id_bits, const uint32_t* n_dev) {
  if (n)
    k_expand_keys32<<<gridFor(n, 256, n_dev, kLoopBlocks), 256, 0, s>>>(in, keys, n, id_bits, n_dev);
}
void launch_world_pairs(hipStream_t s, WorldState* ws, uint32_t nw, const uint32_t* leaf_off,
                        uint32_t MB, uint32_t* max_out) {
  k_world_pairs<<<1, 1024, 0, s>>>(ws, nw, leaf_off, MB, max_out);
}
bool launch_segment_sort(hipStream_t s, const void* raw, bool narrow, uint32_t n, uint32_t nw,
                         uint32_t id_bits, uint32_t max_world, uint64_t* keys,
                         const uint32_t* n_dev) {
  if (2 * id_bits > 32 || max_world > 16384) return false;
  if (!n) return true;
#define VXP_SEGSORT(BS, IPT)                                                                   \
  do {                                                                                         \
    if (narrow)                                                                                \
      k_segment_sort<BS, IPT, uint32_t><<<nw, BS, 0, s>>>((const uint32_t*)raw, n, id_bits, keys, n_dev); \
    else                                                                                       \
      k_segment_sort<BS, IPT, uint64_t><<<nw, BS, 0, s>>>((const uint64_t*)raw, n, id_bits, keys, n_dev); \
  } while (0)
  if (max_world <= 1024) VXP_SEGSORT(256, 4);
  else if (max_world <= 2048) VXP_SEGSORT(256, 8);
  else if (max_world <= 4096) VXP_SEGSORT(512, 8);
  else if (max_world <= 8192) VXP_SEGSORT(1024, 8);
  else VXP_SEGSORT(1024, 16);
#undef VXP_SEGSORT
  return true;
}
void launch_sweep_pairs(hipStream_t s, WorldState* ws, const float4* slo, const float4* shi,
                        uint32_t n_slots, const uint32_t* send, const uint32_t* woff_i,
                        uint32_t* wcount, const uint32_t* woff_w, uint64_t* out, bool count,
                        uint32_t id_bits, bool narrow, bool per_world, uint32_t cap) {
  constexpr uint32_t blocks = kSweepWaves / 4;
  if (count)
    k_sweep_pairs<true><<<blocks, 256, 0, s>>>(ws, slo, shi, n_slots, send, woff_i, wcount, woff_w,
                                               out, id_bits, narrow, per_world, cap);
  else
    k_sweep_pairs<false><<<blocks, 256, 0, s>>>(ws, slo, shi, n_slots, send, woff_i, wcount,
                                                woff_w, out, id_bits, narrow, false, cap);
}
void launch_pair_process(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                         const DevColliders& C, WorldState* ws, vxp_stats* st,
                         const uint64_t* keys, uint32_t n, PairRec* pr, const DevManifold* prev,
                         const uint8_t* drop, uint32_t* cflag, uint32_t* produce,
                         uint32_t* p_vchk, uint32_t* wlist, uint32_t* step_ctr, uint32_t* counts,
                         bool chunk_bound, const uint32_t* n_dev) {
  if (n)
    // 512 threads per block: each block appends its work pairs and adds its
    // chunk bound with one device atomic per counter, and these single-address
    // atomics serialize (128-thread blocks: 365 us per frame late in the
    // 512-world batch, 512: 154 us)
    k_pair_process<<<gridFor(n, 512, n_dev, kLoopBlocks), 512, 0, s>>>(
        cfg, B, C, ws, st, keys, n, pr, prev, drop, cflag, produce, p_vchk, wlist, step_ctr, counts,
        chunk_bound, n_dev);
}
void launch_cache_replay(hipStream_t s, const DevCfg& cfg, WorldState* ws, vxp_stats* st,
                         PairRec* pr, CacheKeyRec* keys, const uint32_t* cflag,
                         const uint32_t* cidx, const uint32_t* order, uint32_t* clist,
                         uint32_t n_pairs, uint32_t nw, uint32_t* counts, const uint32_t* n_dev) {
  if (n_pairs)
    k_cache_scatter<<<gridFor(n_pairs, 256, n_dev, kLoopBlocks), 256, 0, s>>>(cflag, cidx, order,
                                                                             n_pairs, clist, n_dev);
  k_cache_replay<<<nw, kCacheThreads, cfg.cache_cap * sizeof(CacheKeyRec), s>>>(
      cfg, ws, st, pr, keys, cidx, clist, counts, nw);
}
void launch_wi_fill(hipStream_t s, const uint32_t* counts, const uint32_t* wi_off, uint32_t n_pairs,
                    uint32_t* wi_pair, uint32_t* wi_ccnt, const uint32_t* n_dev) {
  if (n_pairs)
    k_wi_fill<<<gridFor(n_pairs, 256, n_dev, kLoopBlocks), 256, 0, s>>>(
        counts, wi_off, n_pairs, wi_pair, wi_ccnt, n_dev);
}
// At most four rounds of 28 waves per CU, a whole number of XCD rounds of
// runs of `run` blocks when it can be.
static uint32_t probeBlocks(uint32_t items, uint32_t waves, uint32_t run) {
  uint32_t most = 4 * 256 * 28 / waves;
  if (most >= kXcds * run) most -= most % (kXcds * run);
  return std::min(cdiv(items, waves), most);
}
void launch_probe_count(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                        const DevColliders& C, const PairRec* pr, const uint32_t* wi_pair,
                        const uint32_t* wi_off, uint32_t total_wi, uint32_t* wi_ccnt,
                        uint32_t* p_vchk, uint8_t* lane_n, uint32_t* lane_list,
                        uint32_t* list_n, ProbeShape shape, const uint32_t* n_dev) {
  if (!total_wi) return;
  const uint32_t threads = shape.threads, run = shape.xcd_run ? shape.xcd_run : 1u;
  k_probe_count<<<probeBlocks(total_wi, threads / kWave, run), threads,
                  threads / kWave * kProbeLdsWords * 4, s>>>(
      cfg, B, C, pr, wi_pair, wi_off, total_wi, wi_ccnt, p_vchk, lane_n, lane_list, list_n, run,
      n_dev);
}
void launch_probe_write(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                        const DevColliders& C, const PairRec* pr, const uint32_t* wi_pair,
                        const uint32_t* wi_off, const uint32_t* lane_list, uint32_t n_items,
                        const uint32_t* wi_cbase, const uint8_t* lane_n, Cand* cand,
                        const uint32_t* n_dev) {
  if (!n_items) return;
  k_probe_write<<<gridFor(n_items, 256, n_dev, kLoopBlocks), 256, 0, s>>>(
      cfg, B, C, pr, wi_pair, wi_off, lane_list, n_items, wi_cbase, lane_n, cand, n_dev);
}
void launch_reduce(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const DevColliders& C,
                   WorldState* ws,
                   vxp_stats* st, PairRec* pr, const uint64_t* keys, uint32_t n,
                   const uint32_t* wlist, uint32_t* step_ctr,
                   const uint32_t* wi_off, const uint32_t* wi_cbase, const uint32_t* wi_ccnt,
                   const uint32_t* p_vchk, const Cand* cand, const DevManifold* prev,
                   const CacheKeyRec* ckeys, CacheVal* cvals, CacheVal* cstage,
                   DevManifold* staging, uint32_t* produce, RStat* rstat, uint32_t* coop) {
  if (!n) return;
  k_reduce_settle<<<gridFor(n, 256, step_ctr, kLoopBlocks), 256, 0, s>>>(
      cfg, pr, wlist, step_ctr, wi_off, wi_cbase, p_vchk, cvals, cstage,
      produce, rstat, coop, n);
  const uint32_t blocks = std::min<uint32_t>(2048, cdiv(n, 4));
  k_reduce<<<blocks, 256, 0, s>>>(cfg, B, C, ws, pr, keys, coop, step_ctr, wi_off, wi_cbase,
                                  wi_ccnt, p_vchk, cand, prev, cvals, cstage, staging, produce,
                                  rstat);
  k_reduce_tail<<<gridFor(n, 256, step_ctr, kLoopBlocks), 256, 0, s>>>(
      cfg, ws, st, pr, n, wlist, step_ctr, ckeys, cvals, cstage, rstat);
}
void launch_world_segments(hipStream_t s, WorldState* ws, uint32_t nw, uint32_t* total) {
  k_world_segments<<<1, 1024, 0, s>>>(ws, nw, total);
}
void launch_scatter(hipStream_t s, const WorldState* ws, const PairRec* pr, uint32_t n,
                    const uint32_t* produce, const uint32_t* pidx, const DevManifold* staging,
                    DevManifold* out, const uint32_t* n_dev) {
  if (n)
    k_scatter<<<gridFor(n, 256, n_dev, kLoopBlocks), 256, 0, s>>>(ws, pr, n, produce, pidx, staging,
                                                                  out, n_dev);
}
void launch_copy_skipped(hipStream_t s, const WorldState* ws, const DevManifold* prev,
                         DevManifold* out, uint32_t nw) {
  k_copy_skipped<<<nw, 256, 0, s>>>(ws, prev, out);
}
void launch_islands(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                    vxp_stats* st, const DevManifold* ml, const DevManifold* prev,
                    const IslandArrays& I, uint32_t nw, uint32_t max_manifolds, size_t lds_limit) {
  const size_t lds = std::min(lds_limit, ((size_t)3 * cfg.id_cap + 2 * (size_t)max_manifolds) * 4);
  k_islands<<<nw, 1024, lds, s>>>(cfg, B, ws, st, ml, prev, I, lds);
}
void launch_ref_coloring(hipStream_t s, const DevCfg& cfg, const WorldState* ws,
                         const DevManifold* ml, uint64_t* masks, uint8_t* colors, uint32_t nw) {
  k_ref_coloring<<<nw, 64, 0, s>>>(cfg, ws, ml, masks, colors);
}
void launch_body_state(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                       const WorldState* ws, const SolverScratch& S, uint32_t n_slots) {
  if (n_slots) k_body_state<<<cdiv(n_slots, 256), 256, 0, s>>>(cfg, B, ws, S, n_slots);
}
void launch_solve_prep(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       vxp_stats* st, DevManifold* ml, const SolverScratch& S, uint32_t* prio,
                       uint32_t* step_ctr, uint32_t nw, uint32_t max_manifolds,
                       size_t lds_limit, const uint32_t* man_total) {
  size_t lds = std::min(lds_limit, std::max((size_t)max_manifolds * kColourLdsPerManifold,
                                            (size_t)cfg.max_bodies * 8));
  k_solve_prep<<<nw, kSolveThreads, lds, s>>>(cfg, B, ws, st, ml, S, prio, step_ctr, lds, man_total);
}
// Longest units first: the queue is reordered by descending power-of-two
// bucket of manifold count (order inside a bucket is free), so that the
// persistent grid starts the units that bound the launch before the small ones.
// Units are independent: any order gives the same results.
__global__ __launch_bounds__(1024) void k_unit_order(SolverScratch S, const uint32_t* step_ctr,
                                                     uint32_t* tmp) {
  __shared__ uint32_t hist[33], cur[33];
  const uint32_t tid = threadIdx.x, n = step_ctr[kCtrSolveUnits];
  if (tid < 33) hist[tid] = 0;
  __syncthreads();
  auto bucket = [&](uint32_t q) {  // 0 for the largest units
    const uint32_t mn = S.units[q].y;
    return (uint32_t)__builtin_clz(mn | 1u);  // 32 - bit length: larger counts first
  };
  for (uint32_t i = tid; i < n; i += blockDim.x) {
    const uint32_t q = S.uqueue[i];
    tmp[i] = q;
    atomicAdd(&hist[bucket(q)], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t a = 0;
    for (uint32_t b = 0; b < 33; b++) {
      cur[b] = a;
      a += hist[b];
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < n; i += blockDim.x) {
    const uint32_t q = tmp[i];
    S.uqueue[atomicAdd(&cur[bucket(q)], 1u)] = q;
  }
}
void launch_unit_order(hipStream_t s, const SolverScratch& S, const uint32_t* step_ctr,
                       uint32_t* tmp) {
  k_unit_order<<<1, 1024, 0, s>>>(S, step_ctr, tmp);
}

void launch_solve_units(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                        DevManifold* ml, const SolverScratch& S, uint32_t* step_ctr, uint32_t nw,
                        uint32_t n_cu, size_t lds_limit, uint32_t threads) {
  size_t vel_bytes = (size_t)cfg.max_bodies * 2 * sizeof(float4);
  uint32_t grid = std::min<uint32_t>(nw * kUnitsMax, 2 * std::max(n_cu, 1u));
  const bool lds = vel_bytes <= lds_limit;
  if (threads == 256) {
    if (lds)
      k_solve_units<true, 256><<<grid, 256, vel_bytes, s>>>(cfg, B, ws, ml, S, step_ctr);
    else
      k_solve_units<false, 256><<<grid, 256, 0, s>>>(cfg, B, ws, ml, S, step_ctr);
  } else {
    if (lds)
      k_solve_units<true, 512><<<grid, 512, vel_bytes, s>>>(cfg, B, ws, ml, S, step_ctr);
    else
      k_solve_units<false, 512><<<grid, 512, 0, s>>>(cfg, B, ws, ml, S, step_ctr);
  }
}
void launch_solve_tail(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       vxp_stats* st, const SolverScratch& S, const uint32_t* step_ctr,
                       uint32_t nw, uint32_t n_cu) {
  uint32_t grid = std::min<uint32_t>(nw * kUnitsMax, 4 * std::max(n_cu, 1u));
  for (uint32_t it = 1; it < cfg.pos_iters; it++)
    k_solve_pos<<<grid, 256, 0, s>>>(cfg, B, ws, S, step_ctr, it);
  k_solve_final<<<nw, kSolveThreads, 0, s>>>(cfg, B, ws, st, S);
}
void launch_brick_masks(hipStream_t s, const uint8_t* vox, uint32_t nb, uint64_t* mask) {
  if (nb) k_brick_masks<<<nb, 64, 0, s>>>(vox, mask);
}
void launch_collider_classify(hipStream_t s, const uint64_t* solid, const int32_t* nbr,
                              uint32_t nb, uint8_t* out) {
  if (nb) k_collider_classify<<<nb, 512, 0, s>>>(solid, nbr, out);
}
void launch_collider_meta(hipStream_t s, const int32_t* pos, const uint8_t* vox, uint32_t nb,
                          float vs, uint32_t* cnt, int32_t* bounds, float* box) {
  if (nb) k_collider_meta<<<nb, 512, 0, s>>>(pos, vox, vs, cnt, bounds, box);
}
void launch_collider_mass(hipStream_t s, const MassJob* jobs, uint32_t n_jobs, const int32_t* pos,
                          const uint8_t* vox, float* out) {
  if (n_jobs) k_collider_mass<<<n_jobs, 64, 0, s>>>(jobs, pos, vox, out);
}
void launch_update_end(hipStream_t s, const DevCfg& cfg, const DevBodies& B, vxp_stats* st,
                       uint32_t nw, uint32_t substeps) {
  k_update_end<<<nw, slotThreads(cfg.max_bodies), 0, s>>>(cfg, B, st, nw, substeps);
}

void launch_copy_many(hipStream_t s, const CopyDesc* d, uint32_t n) {
  if (n) k_copy_many<<<dim3(32, n), 256, 0, s>>>(d, n);
}
void launch_rekey_manifolds(hipStream_t s, DevManifold* ml, const uint32_t* ranges, uint32_t n) {
  if (n) k_rekey_manifolds<<<dim3(4, n), 256, 0, s>>>(ml, ranges, n);
}
void launch_clear_slots(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const uint32_t* slots,
                        uint32_t n) {
  if (n) k_clear_slots<<<dim3(4, n), 256, 0, s>>>(cfg, B, slots, n);
}

}  // namespace vxd
