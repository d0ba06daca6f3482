// vxp_dev.hpp — device data model of the step pipeline (structure-of-arrays
// body state, collider brick pool + dense brick directories, pair / manifold /
// candidate records) shared by the kernels and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vxp.h"
#include "vxp_math.hpp"
#include "vxp_tree.hpp"

namespace vxd {

constexpr int kWave = 64;
constexpr uint32_t kNone = 0xFFFFFFFFu;
// Device counters cleared at the start of every step (step_ctr[]).
constexpr uint32_t kCtrReduceWork = 0;  // pairs on the reduction work list
constexpr uint32_t kUnitsPerWorld = 64;  // solve units per world, at most
constexpr uint32_t kCtrSolveUnits = 1;  // solve units queued by the solver's prep
constexpr uint32_t kCtrSolveUnitPull = 2;  // solve units taken by workgroups
constexpr uint32_t kCtrChunkBound = 3;  // probe work items if every cache-eligible pair missed
constexpr uint32_t kCtrProbeItems = 4;  // probe lanes with candidates (the write pass's list)
constexpr uint32_t kCtrReduceCoop = 5;  // work-list pairs the reduction's waves take
constexpr uint32_t kStepCounters = 6;

// Collider header.  The brick directory is a dense grid over brick positions
// floor(min_bound/8) .. floor(max_bound/8) (inclusive: the overlap region loop
// of src/narrow_phase.zig:759-764 reaches max_bound); each cell holds the pool
// index of the brick that VoxelCollider.getBrick returns for that position
// (hashBrickPos lookup, aliasing included), or -1.
struct DevCollider {
  float vs;
  int32_t minb[3], maxb[3];
  float lo[3], hi[3];  // local AABB
  int32_t org[3], dim[3];
  uint32_t dir_off;
  uint32_t solid;
};

// One computeMassProperties job: a collider's bricks at brick_off..brick_off+nb
// of the staged brick arrays, its solid count and voxel size, and a density.
// Results go to 16 floats per job (inv_mass, 3x3 inverse inertia, centre of
// mass, unused).
struct MassJob {
  uint32_t brick_off, nb, solid;
  float vs, density;
  uint32_t pad[3];
};
constexpr uint32_t kMassOut = 16;

// Configuration values the kernels read (vxp_config subset, by value).
struct DevCfg {
  float gravity[3];
  float dt;            // fixed step
  float h;             // dt / substeps
  float bias_rate, mass_coeff, impulse_coeff;
  float slop, pcr, mpc, wss;
  uint32_t substeps, vel_iters, pos_iters;
  uint32_t max_bodies;   // slots per world
  uint32_t id_cap;       // id-indexed arrays per world
  uint32_t cache_cap;
  uint8_t enable_sleeping, warm_starting, enable_friction, enable_restitution;
  uint8_t disable_rotation, use_sat, use_cache, enable_islands;
  uint8_t store_impulses;
  uint8_t threaded;      // PhysicsConfig.enable_threading: the thread-pool narrow phase
  uint8_t thread_multi;  // ... with 2+ threads (thread_count 0 or >= 2)
  uint8_t pad;
  uint32_t pair_color_max;  // colours of at most this many manifolds run on lane pairs
  uint32_t n_cu;            // compute units (solve units: about two per CU over the batch)
};

// Body state, global slot g = world * max_bodies + slot.
struct DevBodies {
  float4 *pos, *rot, *lv, *av;
  float* inv_mass;
  float* iil;  // 9 per body: local inverse inertia
  float *fric, *rest, *ldamp, *adamp, *sleep_timer;
  uint8_t *type, *sleeping, *active;
  int32_t* collider;
  uint32_t* id;
  float4 *fat_lo, *fat_hi;
  int32_t* idmap;  // [world * id_cap + id] -> slot or -1
};

constexpr uint32_t kMaskWords = 32;  // brick mask words: 4 planes of 8 z slices
struct DevColliders {
  const DevCollider* hdr;
  const int32_t* dir;
  const uint8_t* vox;     // 512 bytes per brick
  const uint64_t* mask;   // kMaskWords per brick: non-empty, corner, edge, inside planes
                          // of 8 words; bit x + 8y of word z
};

// Per-world step bookkeeping.
struct WorldState {
  uint32_t skip;           // all dynamic bodies asleep: step() returns early
  uint32_t awake;
  uint32_t pair_lo, pair_hi;
  uint32_t man_start, man_count;      // current manifold segment
  uint32_t prev_start, prev_count;    // previous frame's segment
  uint32_t produced;
  uint32_t frame;                     // BrickPairCache.current_frame
  uint32_t error;
  uint32_t axis;                      // sort-and-sweep axis (longest extent)
  // LBVH broad phase: active slots (tree leaves) and the Morton grid of the
  // doubled fat-box centres, q = (c - c_lo) * c_scale clamped to [0, 1023]
  uint32_t n_leaves;
  float c_lo[3], c_scale[3];
  // thread-pool narrow phase (enable_threading): slots below thr_cut hold the
  // first 256 active bodies; thr_par: the parallel branch runs this step
  uint32_t thr_cut, thr_par;
  uint32_t np_emit;  // pairs the broad phase emits for this world this step
  // solver (written by its prep each step)
  uint32_t n_units, nsolve;           // solve units, solved manifolds
  uint32_t max_unit_m;                // manifolds of the largest unit
  uint32_t pos_run, pos_done;         // position passes run; early exit taken
  uint32_t maxpen;                    // float bits of the maximum penetration
  uint32_t pos_min[3];                // float bits of each pass's minimum separation
  // IslandManager (src/islands.zig:196-270)
  uint32_t isl_need;                  // needs_rebuild
  vxp_island_stats isl;               // stats after the last sleep update
};

// The broad-phase trees of every world: tree (w, which) is nodes[(2w + which)
// * nc ..], head heads[2w + which]; which 0 = static, 1 = dynamic.  leaf[g] is
// the node of body slot g in its tree.
struct TreeSet {
  TreeNode* nodes;
  TreeHead* heads;
  uint32_t* leaf;
  uint32_t nc;
  __host__ __device__ Tree tree(uint32_t w, uint32_t which) const {
    return Tree{nodes + ((size_t)w * 2 + which) * nc, heads + (size_t)w * 2 + which, nc};
  }
};

// WorldState.error bits
constexpr uint32_t kErrColours = 1;     // solver colouring past its colour limit
constexpr uint32_t kErrTreeStack = 2;   // a tree query deeper than its stack
constexpr uint32_t kErrLbvh = 4;        // an LBVH walk longer than its tree (malformed tree)

// Id-indexed island arrays, [world * id_cap + body id]: union-find parents
// (labels after a step), endpoint flags, component sizes, and the sleep timer
// and flag of each island, kept under its label.
struct IslandArrays {
  uint32_t *parent, *endpoint, *size;
  float* timer;
  uint32_t* sleep;
};

// Pair flags
enum : uint32_t {
  PF_NONE = 0,     // no manifold this step
  PF_CARRY = 1,    // both asleep: carry previous manifold
  PF_DETECT = 2,   // voxel detection, no cache
  PF_CACHE = 3,    // cache-eligible, before replay
  PF_HIT = 4,      // cache hit
  PF_MISS = 5,     // cache miss: detect then fill
};

struct PairRec {
  uint32_t world, sa, sb;  // slots within the world
  uint32_t flag;
  int32_t mn[3];
  uint32_t nx, ny, nz;
  uint32_t nchunks;
  uint32_t aux;       // PF_CARRY: prev manifold index; HIT/MISS: cache entry
  uint32_t ida, idb;
  uint32_t vchk;
  uint64_t th;        // cache transform hash
};

// Manifold record (device): pair key + slots + the exported vxp_manifold.
struct DevManifold {
  uint64_t key;  // (world << 44) | (id_a << 22) | id_b
  uint32_t sa, sb;
  vxp_manifold m;
};

struct Cand {  // one contact candidate: what the top-4 replay reads, and where it came from
  float pen;
  uint32_t fid_lo, fid_hi;
  uint32_t code;  // lane (A voxel in the work item) << 5 | probe index (B neighbour)
};

struct CacheKeyRec {
  uint32_t a, b;
  uint64_t th;
  uint64_t last_used;
  uint32_t valid;
  uint32_t owner_frame;
  uint32_t owner_pair;
  uint32_t pad;
};
struct CacheVal {
  vxp_contact c[4];
  uint32_t count, has_collision, vchk, pad;
};

// Narrow-phase counters of one pair, summed per world after the reduction.
struct RStat {
  uint32_t produced, vchk, colliding, contacts, cands, pad;
};

// One device-to-device byte range of a batched copy.
struct CopyDesc {
  const void* src;
  void* dst;
  uint64_t bytes;
};

inline __host__ __device__ uint64_t makeKey(uint32_t w, uint32_t a, uint32_t b) {
  return ((uint64_t)w << 44) | ((uint64_t)a << 22) | (uint64_t)b;
}
inline __host__ __device__ uint32_t keyWorld(uint64_t k) { return (uint32_t)(k >> 44); }
inline __host__ __device__ uint32_t keyA(uint64_t k) { return (uint32_t)((k >> 22) & 0x3FFFFF); }
inline __host__ __device__ uint32_t keyB(uint64_t k) { return (uint32_t)(k & 0x3FFFFF); }
inline __host__ __device__ uint64_t exportKey(uint64_t k) {
  return ((uint64_t)keyA(k) << 32) | keyB(k);
}

// Solver scratch: per constraint point (4 per manifold), structure of arrays.
struct SolverScratch {
  float4 *ws_pos, *ws_rot, *ws_lv, *ws_av;  // per body slot
  float* ws_ii;                              // 9 per body slot
  uint32_t *adj_cnt, *adj_off, *adj;         // CSR of awake-dynamic incidences
  int32_t* color;                            // per manifold
  uint32_t* order;                           // per manifold: grouped by colour
  float4* rec;                               // 9 per manifold (colour order)
  float4* pts;                               // 4 x 8 per manifold (colour order)
  uint64_t* tstamp;                          // optional phase timestamps (16 per world)
  // solve units: components of the graph of mobile bodies, cut into units
  uint32_t *upar, *ucnt, *uoff;              // per body slot: parent, manifolds, offsets
  uint32_t* ubody;                           // per body slot: mobile bodies by unit
  uint32_t* uorder;                          // per manifold: solved manifolds by unit
  uint4* units;                              // per world x 64: (man off, n, body off, n)
  uint32_t* uqueue;                          // world * 64 + unit, in queue order
  uint32_t *pcnt, *fq;                       // per manifold: uncoloured lower neighbours, queue
  uint4* cend;                               // per manifold: endpoints' incidence ranges
};

}  // namespace vxd
