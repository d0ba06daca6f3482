/*
 * vxp.h — C ABI of the MI355X step pipeline for zig-voxel-physics.
 *
 * Every entry point below replaces one entry point of the reference's
 * `VoxelPhysicsSystem` / `VoxelCollider` API (Zig).  Signatures use plain
 * pointers and sizes only; no device pointers or torch types cross this
 * boundary.  Every call returns an int status (VXP_OK or a negative code);
 * nothing throws across the ABI.
 *
 * A `vxp_system` holds one or more independent worlds (a batch).  The
 * reference API maps onto world 0; the `_w` variants take a world index and
 * drive the batched path (many worlds stepped per launch).
 */
#ifndef VXP_H
#define VXP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (src/system.zig:487 MaxBodiesReached; OOM) ---------- */
#define VXP_OK 0
#define VXP_ERR_INVALID_HANDLE (-1)
#define VXP_ERR_CAPACITY (-2)      /* error.MaxBodiesReached */
#define VXP_ERR_OOM (-3)           /* allocator OOM / HBM budget */
#define VXP_ERR_DEVICE (-4)        /* HIP runtime or kernel fault */
#define VXP_ERR_ARGUMENT (-5)

/* ---- body types (src/body.zig:29-48, enum(u2)) -------------------------- */
#define VXP_BODY_STATIC 0
#define VXP_BODY_KINEMATIC 1
#define VXP_BODY_DYNAMIC 2

/* ---- voxel byte: u3 type | u5 normal index << 3 (src/collision_voxel.zig:133-191) */
#define VXP_VOXEL_EMPTY 0
#define VXP_VOXEL_CORNER 1
#define VXP_VOXEL_EDGE 2
#define VXP_VOXEL_FACE 3
#define VXP_VOXEL_INSIDE 4
#define VXP_VOXEL(type, normal) ((uint8_t)(((type) & 7u) | (((normal) & 31u) << 3)))

/* Solve order inside one step.  The reference visits manifolds in Zig
 * AutoHashMap order (src/system.zig:687-702), which cannot be reproduced
 * outside Zig's stdlib; both orders below are deterministic restatements. */
#define VXP_SOLVE_ORDER_COLOR 0     /* per-colour Gauss-Seidel (GPU order)  */
#define VXP_SOLVE_ORDER_CANONICAL 1 /* ascending pair key                   */

/* Order in which the narrow phase takes the broad-phase pairs: the pair cap
 * and the brick-pair cache follow it (detection results do not depend on it). */
#define VXP_PAIR_ORDER_KEY 0   /* ascending pair key */
#define VXP_PAIR_ORDER_TREE 1  /* the reference's AABB-tree order (src/broad_phase.zig:488-526):
                                  with pair_cap = 512, the cap-faithful mode */

/*
 * Configuration: PhysicsConfig (src/system.zig:47-88) + SolverConfig
 * (src/solver.zig:75-119) + NarrowPhaseConfig (src/narrow_phase.zig:1325-1343)
 * field for field, plus the documented deviation knobs at the end.
 * Fill with vxp_config_default() and override.
 */
typedef struct vxp_config {
  /* PhysicsConfig */
  float gravity[3];            /* (0,-9.81,0) */
  float fixed_timestep;        /* 1/60 */
  uint32_t max_substeps;       /* 8: frame catch-up cap */
  uint32_t max_bodies;         /* 1024 per world */
  float default_friction;      /* 0.6 */
  float default_restitution;   /* 0.2 */
  float default_density;       /* 1.0 (never read by the reference) */
  uint8_t enable_sleeping;     /* 1 */
  uint8_t enable_warm_starting;/* 1 */
  uint8_t enable_islands;      /* 1: island labels + island stats each step */
  uint8_t enable_coloring;     /* 0: reference GraphColoring artifact (built
                                  only with enable_threading in the reference) */
  /* SolverConfig */
  uint32_t substeps;           /* 4 */
  uint32_t velocity_iterations;/* 1 */
  uint32_t position_iterations;/* 2 */
  float hertz;                 /* 30 */
  float damping_ratio;         /* 1 */
  float position_correction_rate; /* 0.2 */
  float max_position_correction;  /* 0.04 */
  float slop;                  /* 0.005 */
  float warm_start_scale;      /* 0.8 */
  uint8_t disable_rotation;    /* 0 */
  uint8_t warm_starting;       /* 1 */
  uint8_t enable_friction;     /* 1 */
  uint8_t enable_restitution;  /* 1 */
  /* NarrowPhaseConfig */
  uint8_t use_brick_sat;       /* 1 */
  uint8_t use_brick_pair_cache;/* 1 */
  /* PhysicsConfig.enable_threading (0): the thread-pool narrow phase of
     src/system.zig:618-638 (first 256 active bodies only, read-only cache
     for 8+ pairs; see thread_count) */
  uint8_t enable_threading;
  uint8_t pad0;
  uint32_t brick_pair_cache_size; /* 1024 */
  /* deviation knobs */
  uint32_t pair_cap;           /* 0 = unlimited; 512 = the reference's
                                  filtered-pair array (src/system.zig:594) */
  uint32_t solve_order;        /* VXP_SOLVE_ORDER_* */
  uint32_t cpu_threads;        /* CPU restatement only: narrow-phase workers */
  uint32_t pair_order;         /* VXP_PAIR_ORDER_* */
  uint32_t thread_count;       /* PhysicsConfig.thread_count (0 = the host's CPU
                                  count); with enable_threading only 1 versus 2+
                                  changes results */
} vxp_config;

/* Per-step statistics: PhysicsStats (src/system.zig:91-155) counters, for
 * the last executed step of the last update() call, plus build counters. */
typedef struct vxp_stats {
  uint32_t total_bodies, dynamic_bodies, static_bodies, sleeping_bodies, active_bodies;
  uint32_t broad_phase_pairs, narrow_phase_tested, narrow_phase_colliding;
  uint32_t contacts_generated, voxels_checked, pairs_skipped_sleeping;
  uint32_t cache_hits, cache_misses, manifolds_skipped_sleeping;
  uint32_t island_count, largest_island_size, substeps_this_frame;
  uint32_t manifolds, solver_manifolds, colors, position_iterations_run;
  uint32_t steps_executed;
  uint32_t awake_bodies_stepped;  /* sum over executed steps of awake dynamic bodies */
  uint32_t solver_points;         /* contact points in solved manifolds (last step) */
  float max_penetration;
  float accumulated_time;
  uint32_t contact_candidates;    /* generateContact candidates before the top-4 (last step) */
  uint32_t active_islands, sleeping_islands;  /* IslandManager stats as the island build left them */
} vxp_stats;

/* IslandManager.IslandStats (src/islands.zig:222-243) of one world. */
typedef struct vxp_island_stats {
  uint32_t total_islands, active_islands, sleeping_islands, largest_island_size;
  uint32_t total_bodies, total_contacts, rebuilds_skipped, rebuilds_performed;
} vxp_island_stats;

/* ContactPoint (src/narrow_phase.zig:336-393) */
typedef struct vxp_contact {
  float point[3];
  float normal[3];
  float penetration;
  float local_anchor_a[3];
  float local_anchor_b[3];
  float normal_impulse;
  float tangent_impulse[2];
  float normal_mass;
  float tangent_mass[2];
  uint32_t pad;
  uint64_t feature_id;
} vxp_contact;

/* ContactManifold (src/narrow_phase.zig:396-502) */
typedef struct vxp_manifold {
  uint32_t body_a_id, body_b_id;
  float friction, restitution;
  uint32_t contact_count;
  uint32_t color;              /* solver colour this step (0xFFFFFFFF: not solved) */
  vxp_contact contacts[4];
} vxp_manifold;

/* BodyHandle (src/system.zig:158-167) */
typedef struct vxp_handle {
  uint32_t index, generation;
} vxp_handle;

/* Body state, as the reference's VoxelBody hot fields (src/body.zig:177-203). */
typedef struct vxp_body_state {
  float position[3];
  float rotation[4];           /* x, y, z, w */
  float linear_velocity[3];
  float angular_velocity[3];
  float sleep_timer;
  uint32_t id;
  uint8_t body_type;
  uint8_t is_sleeping;
  uint8_t active;
  uint8_t pad;
} vxp_body_state;

typedef struct vxp_system vxp_system;

/* Defaults of PhysicsConfig{} / SolverConfig{} / NarrowPhaseConfig{}. */
int vxp_config_default(vxp_config* cfg);

/* VoxelPhysicsSystem.init / deinit (src/system.zig:213-291).
 * n_worlds >= 1 independent worlds share the device and every launch. */
int vxp_system_create(const vxp_config* cfg, uint32_t n_worlds, vxp_system** out);
int vxp_system_destroy(vxp_system* sys);

/* VoxelCollider.init + setCollisionVoxel... + recomputeMetadata
 * (src/collider.zig:168-179,267-273,283-333).  `brick_pos` holds 3*n_bricks
 * brick coordinates, `voxels` 512*n_bricks voxel bytes indexed x+8y+64z.
 * The collider is copied to the device; the caller may free its arrays.
 * A collider may be shared by any number of bodies of any world. */
int vxp_collider_create(vxp_system* sys, float voxel_size, uint32_t n_bricks,
                        const int32_t* brick_pos, const uint8_t* voxels, uint32_t* out_collider);
/* The same from solid occupancy: every solid voxel is classified with
 * classifyFromNeighbors over its six face neighbours (src/collision_voxel.zig:
 * 194-235; a neighbour in a brick that is not given is empty) on the device,
 * then stored as by vxp_collider_create.  `solid` holds 8 words per brick,
 * bit x + 8y of word z; brick positions must be distinct and within +-2^20. */
int vxp_collider_create_solid(vxp_system* sys, float voxel_size, uint32_t n_bricks,
                              const int32_t* brick_pos, const uint64_t* solid,
                              uint32_t* out_collider);
/* VoxelCollider metadata after recomputeMetadata (src/collider.zig:283-333):
 * bounds = min_bound xyz, max_bound xyz (voxels, max exclusive); aabb =
 * local AABB min xyz, max xyz; solid = solid_count. */
int vxp_collider_info(vxp_system* sys, uint32_t collider, int32_t bounds[6], float aabb[6],
                      uint32_t* solid);
/* computeMassProperties (src/body.zig:462-519) of a collider at a density:
 * out = inv_mass, inverse inertia (row-major 3x3), getCenterOfMass xyz
 * (src/collider.zig:376-403). */
int vxp_collider_mass_properties(vxp_system* sys, uint32_t collider, float density,
                                 float out[13]);
/* Allocation-free variant of vxp_collider_mass_properties, with the same
 * results bit for bit.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops
 * that admit worlds between ticks, harnesses that forbid hipMalloc).  It calls
 * no hipMalloc/hipFree: the collider's bricks, the job and the results are
 * staged in one 4 MiB device array per GPU, held until the process ends.
 * Colliders of more than vxp_noalloc_max_bricks() stored bricks are rejected
 * with VXP_ERR_CAPACITY before anything is launched.  Calls share that array:
 * calls that overlap on one GPU (two threads, two systems) race. */
int vxp_collider_mass_properties_noalloc(vxp_system* sys, uint32_t collider, float density,
                                         float out[13]);
/* The largest collider (stored bricks) vxp_collider_mass_properties_noalloc accepts. */
uint32_t vxp_noalloc_max_bricks(void);
/* The collider's stored bricks (after hash merging, src/collider.zig:232-239),
 * ascending (z, y, x) position: 3 ints and 512 voxel bytes each; `count`
 * receives the brick count (call with capacity 0 to size the arrays). */
int vxp_collider_bricks(vxp_system* sys, uint32_t collider, int32_t* brick_pos, uint8_t* voxels,
                        uint32_t capacity, uint32_t* count);

/* createDynamicBody / createStaticBody / createKinematicBody
 * (src/system.zig:294-322,354-377,380-402) in world 0 and in world `w`. */
int vxp_body_create(vxp_system* sys, uint32_t body_type, const float pos[3], const float rot[4],
                    uint32_t collider, float density, vxp_handle* out);
int vxp_body_create_w(vxp_system* sys, uint32_t world, uint32_t body_type, const float pos[3],
                      const float rot[4], uint32_t collider, float density, vxp_handle* out);
/* createDynamicBodyWithMass (src/system.zig:323-351): a dynamic body with the given
 * inverse mass and local inverse inertia (row-major 3x3) instead of the
 * collider's computeMassProperties. */
int vxp_body_create_with_mass_w(vxp_system* sys, uint32_t world, const float pos[3],
                                const float rot[4], uint32_t collider, float inv_mass,
                                const float inv_inertia[9], vxp_handle* out);
/* destroyBody (src/system.zig:405-433) */
int vxp_body_destroy_w(vxp_system* sys, uint32_t world, vxp_handle h);

/* update(dt) (src/system.zig:508-534) on every world; step(dt) (src/system.zig:537-555). */
int vxp_system_update(vxp_system* sys, float dt);
int vxp_system_step(vxp_system* sys, float dt);
/* Batched: `frames` consecutive update(dt) calls on every world, one device
 * launch sequence per phase for all worlds together. */
int vxp_system_update_frames(vxp_system* sys, float dt, uint32_t frames);

/* setGravity / setSolverConfig (src/system.zig:1034-1052): gravity and the
 * SolverConfig fields of `cfg` (substeps, iterations, soft-constraint and
 * position-correction parameters, warm starting, friction, restitution,
 * rotation) apply from the next step; every other field must equal the
 * system's (VXP_ERR_ARGUMENT otherwise). */
int vxp_system_set_config(vxp_system* sys, const vxp_config* cfg);

/* getBody / accessors (src/system.zig:436-445,885-1008): copy state out / in.
 * `out` receives max_bodies entries (slot order) of world `w`. */
int vxp_bodies_get_w(vxp_system* sys, uint32_t world, vxp_body_state* out, uint32_t capacity,
                     uint32_t* count);
/* Body mutators between frames, one per reference method (VoxelPhysicsSystem
 * forwards to VoxelBody: src/system.zig:886-961, src/body.zig:266-414), with
 * the same type guards, arithmetic and wake-ups:
 *   VXP_OP_APPLY_FORCE            applyForce(v)            lv += v * inv_mass
 *   VXP_OP_APPLY_FORCE_AT_POINT   applyForceAtPoint(v, p)  + applyTorque((p - pos) x v)
 *   VXP_OP_APPLY_IMPULSE          applyImpulse(v)
 *   VXP_OP_APPLY_IMPULSE_AT_POINT applyImpulseAtPoint(v, p)
 *   VXP_OP_APPLY_TORQUE           applyTorque(v)           av += I_world^-1 v
 *   VXP_OP_APPLY_ANGULAR_IMPULSE  applyAngularImpulse(v)
 *   VXP_OP_SET_LINEAR_VELOCITY    setLinearVelocity(v)     (dynamic bodies)
 *   VXP_OP_SET_ANGULAR_VELOCITY   setAngularVelocity(v)    (dynamic bodies)
 *   VXP_OP_SET_POSITION           setPosition(v)
 *   VXP_OP_SET_ROTATION           setRotation(v[0..3] = x, y, z, w)
 *   VXP_OP_WAKE_UP                wakeUp()
 *   VXP_OP_SLEEP                  sleep()                  zeroes both velocities
 * `v` (3 floats, 4 for a rotation) is unused by WAKE_UP and SLEEP, `point` is
 * used by the *_AT_POINT ops only. */
#define VXP_OP_APPLY_FORCE 0
#define VXP_OP_APPLY_FORCE_AT_POINT 1
#define VXP_OP_APPLY_IMPULSE 2
#define VXP_OP_APPLY_IMPULSE_AT_POINT 3
#define VXP_OP_APPLY_TORQUE 4
#define VXP_OP_APPLY_ANGULAR_IMPULSE 5
#define VXP_OP_SET_LINEAR_VELOCITY 6
#define VXP_OP_SET_ANGULAR_VELOCITY 7
#define VXP_OP_SET_POSITION 8
#define VXP_OP_SET_ROTATION 9
#define VXP_OP_WAKE_UP 10
#define VXP_OP_SLEEP 11
int vxp_body_op_w(vxp_system* sys, uint32_t world, vxp_handle h, uint32_t op, const float* v,
                  const float point[3]);
/* getWorldAABB (src/system.zig:1003-1008, src/body.zig:430-435): the
 * collider's local AABB through the body's transform; out = min xyz, max xyz. */
int vxp_body_world_aabb_w(vxp_system* sys, uint32_t world, vxp_handle h, float out[6]);
/* findBodyById (src/system.zig:462-473): VXP_ERR_INVALID_HANDLE if no active
 * body of world `w` has this id. */
int vxp_body_find_w(vxp_system* sys, uint32_t world, uint32_t id, vxp_handle* out);
/* setPosition/setRotation/setLinearVelocity/setAngularVelocity + wakeUp on one body. */
int vxp_body_set_state_w(vxp_system* sys, uint32_t world, vxp_handle h, const float pos[3],
                         const float rot[4], const float lin_vel[3], const float ang_vel[3]);

/* getStats (src/system.zig:1024-1026) of world `w`. */
int vxp_stats_get_w(vxp_system* sys, uint32_t world, vxp_stats* out);

/* getStats of every world at once: `out` receives n_worlds entries. */
int vxp_stats_get_all(vxp_system* sys, vxp_stats* out, uint32_t capacity);

/* getContactManager (src/system.zig:1029-1031): the current manifolds of
 * world `w` in ascending pair-key order. */
int vxp_manifolds_get_w(vxp_system* sys, uint32_t world, vxp_manifold* out, uint32_t capacity,
                        uint32_t* count);
/* BroadPhase.findOverlappingPairs pairs of the last step (src/broad_phase.zig:606-619)
 * in the narrow phase's order: ascending, or the trees' emission order under
 * VXP_PAIR_ORDER_TREE; each entry (a<<32)|b with a < b. */
int vxp_pairs_get_w(vxp_system* sys, uint32_t world, uint64_t* out, uint32_t capacity,
                    uint32_t* count);
/* Island labels of the last step (src/islands.zig:318-441): for each
 * manifold endpoint body id (ascending), the minimum body id of its island. */
int vxp_islands_get_w(vxp_system* sys, uint32_t world, uint32_t* body_ids, uint32_t* labels,
                      uint32_t capacity, uint32_t* count);
/* getIslandStats (src/system.zig:868-873): the island manager's statistics
 * after the last step's sleep update.  VXP_ERR_ARGUMENT when islands are
 * disabled (the reference returns null). */
int vxp_island_stats_get_w(vxp_system* sys, uint32_t world, vxp_island_stats* out);
/* isBodyInSleepingIsland (src/system.zig:845-853, src/islands.zig:554-562):
 * *out = 1 if the body's island (as last built) is asleep, else 0 (a body in
 * no island, islands disabled); an invalid handle sets 0 and returns
 * VXP_ERR_INVALID_HANDLE. */
int vxp_body_in_sleeping_island_w(vxp_system* sys, uint32_t world, vxp_handle h, int* out);
/* wakeBodyIsland (src/system.zig:855-866, src/islands.zig:497-516): clears the
 * sleep flag and timer of the body's island, forces an island rebuild on the
 * next step, and wakes the body.  An invalid handle changes nothing and
 * returns VXP_ERR_INVALID_HANDLE. */
int vxp_body_wake_island_w(vxp_system* sys, uint32_t world, vxp_handle h);
/* One node of a broad-phase tree (TreeNode, src/broad_phase.zig:40-62). */
typedef struct vxp_tree_node {
  float min[3], max[3];
  uint32_t body_id;                    /* leaves */
  uint32_t parent, child1, child2;     /* 0xFFFFFFFF: none; free nodes: parent = next free */
  int32_t height;                      /* -1: on the free list */
  uint32_t pad;
} vxp_tree_node;
/* The static (which = 0) or dynamic (which = 1) AABBTree of world `w`
 * (src/broad_phase.zig:64-529) as the last step left it: nodes.items in index
 * order and the root.  Needs pair_order = VXP_PAIR_ORDER_TREE (VXP_ERR_ARGUMENT
 * otherwise). */
int vxp_tree_get_w(vxp_system* sys, uint32_t world, uint32_t which, vxp_tree_node* out,
                   uint32_t capacity, uint32_t* count, uint32_t* root);
/* GraphColoring.build colours (src/islands.zig:668-736) of the current
 * manifolds in pair-key order (needs enable_coloring). */
int vxp_coloring_get_w(vxp_system* sys, uint32_t world, uint8_t* colors, uint32_t capacity,
                       uint32_t* count);

/* Kernel timing (HIP events on the pipeline's stream), for measurement.
 * Kernel ids: */
#define VXP_K_STEP_BEGIN 0
#define VXP_K_FAT_AABB 1
#define VXP_K_PAIRS 2
#define VXP_K_PAIR_PROCESS 3
#define VXP_K_CACHE_REPLAY 4
#define VXP_K_PROBE 5
#define VXP_K_REDUCE 6
#define VXP_K_COMPACT 7
#define VXP_K_ISLANDS 8
#define VXP_K_SOLVE 9        /* solver prep: body state, colouring, solve units */
#define VXP_K_SOLVE_UNITS 10 /* the solve units: substeps, first position pass */
#define VXP_K_SOLVE_TAIL 11  /* further position passes, writeback, sleep */
#define VXP_K_COUNT 12
/* Timing modes: every phase with HIP events on the system's stream, plus the
 * solver's device timestamps (and the island labels kept on that stream);
 * or only VXP_K_SOLVE_UNITS, two events per step, which leaves the step's
 * schedule as it is unprofiled. */
#define VXP_PROFILE_OFF 0
#define VXP_PROFILE_ALL 1
#define VXP_PROFILE_SOLVE_UNITS 2
/* Set the timing mode (VXP_PROFILE_*; VXP_ERR_ARGUMENT otherwise); a mode
 * other than OFF clears the accumulators. */
int vxp_profile_enable(vxp_system* sys, int on);
/* Accumulated device time (ms) and launch count of kernel `id`. */
int vxp_profile_get(vxp_system* sys, uint32_t id, double* total_ms, uint64_t* launches);
/* Solver phase timestamps of the last profiled step of world `w` (device
 * wall clock ticks, 100 MHz): prep start, colouring start, colouring end,
 * units formed, writeback start, writeback end, sleep update done, colouring
 * rounds start (ts[0..7]); the world's largest solve unit: start, records
 * built, substep 0 prepared, substep 0 velocity iterations done, substep 0
 * relaxed, substeps done, first position pass done, end (ts[8..15]). */
int vxp_profile_solver_phases(vxp_system* sys, uint32_t world, uint64_t ts[16]);

/* ---- serving: continuous batching over resident worlds -------------------
 * No counterpart in the reference, which steps one world per update() call;
 * this is the request loop for callers that advance many worlds on demand.
 * A server keeps worlds resident in device memory (a pool) and steps up to
 * `batch_slots` of them together, one update(fixed_timestep) frame per tick.
 * Requests "advance world W by k frames" queue in submission order; every tick
 * admits queued requests into free slots (one request per world at a time,
 * FIFO per world), advances every occupied slot by one frame with one launch
 * per pipeline phase, and retires the requests whose frames are done.  Worlds
 * are admitted by an HBM budget: the pool holds as many worlds as the budget
 * covers and vxp_server_world_create fails with VXP_ERR_OOM beyond that.
 * Times are seconds of CLOCK_MONOTONIC. */
typedef struct vxp_server vxp_server;

typedef struct vxp_request_done {
  uint64_t tag;
  uint32_t world, frames;
  double submit_s, admit_s, done_s;
  uint64_t body_substeps;  /* awake dynamic bodies x substeps over its frames */
} vxp_request_done;

typedef struct vxp_server_info {
  uint32_t worlds, capacity, slots, occupied, queued, pad;
  uint64_t ticks, completed, body_substeps;
  uint64_t resident_bytes, budget_bytes;
} vxp_server_info;

int vxp_server_create(const vxp_config* cfg, uint32_t max_worlds, uint32_t batch_slots,
                      uint64_t hbm_budget_bytes, vxp_server** out);
int vxp_server_destroy(vxp_server* srv);
/* VoxelCollider.init + setCollisionVoxel... (src/collider.zig:168-179,267-273), shared by all worlds. */
int vxp_server_collider_create(vxp_server* srv, float voxel_size, uint32_t n_bricks,
                               const int32_t* brick_pos, const uint8_t* voxels,
                               uint32_t* out_collider);
/* A new empty resident world (VoxelPhysicsSystem.init, src/system.zig:213-259). */
int vxp_server_world_create(vxp_server* srv, uint32_t* out_world);
/* create*Body (src/system.zig:294-402) in a resident world that is not in a slot. */
int vxp_server_body_create(vxp_server* srv, uint32_t world, uint32_t body_type, const float pos[3],
                           const float rot[4], uint32_t collider, float density, vxp_handle* out);
/* Queue "advance `world` by `frames` frames"; submit_s < 0 stamps the call time. */
int vxp_server_submit(vxp_server* srv, uint32_t world, uint32_t frames, uint64_t tag,
                      double submit_s);
/* One tick: admit, step every occupied slot one frame, retire.  *n_done
 * receives the number of requests retired by this tick. */
int vxp_server_tick(vxp_server* srv, uint32_t* n_done);
/* Drain the retired requests (oldest first). */
int vxp_server_poll(vxp_server* srv, vxp_request_done* out, uint32_t capacity, uint32_t* count);
/* State of a resident world that is not in a slot (as vxp_bodies_get_w,
 * vxp_stats_get_w, vxp_manifolds_get_w). */
int vxp_server_bodies_get(vxp_server* srv, uint32_t world, vxp_body_state* out, uint32_t capacity,
                          uint32_t* count);
int vxp_server_stats_get(vxp_server* srv, uint32_t world, vxp_stats* out);
int vxp_server_manifolds_get(vxp_server* srv, uint32_t world, vxp_manifold* out,
                             uint32_t capacity, uint32_t* count);
int vxp_server_info_get(vxp_server* srv, vxp_server_info* out);

/* Launch-shape tuning.  Every choice computes bit-identical results; only the
 * speed differs.  An online tuner times the candidates of one knob at a time
 * on live frames (3 frames per candidate, timed over the phase the knob
 * shapes: solve, broad phase or voxel probe), switches to the fastest when it
 * beats the value in use by 10 %, and goes on to the next knob; a round over
 * every knob that is not forced starts at creation and again 120 frames after
 * the last one ended (the best shapes change as piles form).
 * pair_color_max: colours of at most this many manifolds are solved
 * on lane pairs (one body per lane) instead of one lane per manifold.
 * unit_threads: threads per solve-unit workgroup, 256 or 512 (default: 256
 * for systems of 64 worlds or more, where many units share the device; 512
 * otherwise, where one large unit bounds the solve).  broad_phase: how the
 * pair set of the untraced pair order is found (VXP_PAIR_ORDER_TREE keeps the
 * reference's trees): a sort-and-sweep along each world's best axis, or a
 * Morton-sorted LBVH per world; both find the pairs of
 * BroadPhase.findOverlappingPairs (src/broad_phase.zig:488-526). */
#define VXP_BROAD_AUTO 0u
#define VXP_BROAD_SWEEP 1u
#define VXP_BROAD_LBVH 2u
typedef struct vxp_tuning {
  uint32_t pair_color_max;
  uint32_t windows;       /* tuning windows completed */
  uint32_t fixed;         /* set by vxp_tuning_set: the tuner is off */
  uint32_t unit_threads;  /* get: in use; set: 256 or 512, 0 = default */
  uint32_t broad_phase;   /* get: in use; set: VXP_BROAD_*, AUTO = tuned */
  uint32_t probe_threads; /* voxel probe workgroup: 64, 128 or 256 threads; 0 = tuned */
  uint32_t probe_xcd_run; /* probe workgroups kept on one XCD per run (power of two
                             <= 4096; consecutive work items of a pair share its
                             L2); 0 = tuned */
} vxp_tuning;
int vxp_tuning_get(vxp_system* sys, vxp_tuning* out);
/* Fix the tuning (the tuner stops); pair_color_max = 0xFFFFFFFF restarts it
 * instead, over the knobs left at 0.  The other knobs are forced either way
 * when not 0 (VXP_ERR_ARGUMENT for a value outside those listed above). */
int vxp_tuning_set(vxp_system* sys, const vxp_tuning* tuning);

/* ---- recorded step (HIP graphs) ---------------------------------------------
 * No counterpart in the reference, whose update() steps on the CPU; this is
 * for batched systems and the serving loop, where the step's ~45 launches and
 * its per-step size reads leave the GPU idle between kernels.  With recording
 * on, a step of the default pipeline (sort-and-sweep pairs in key order, no
 * pair cap, no thread-pool narrow phase, timing off, no tuning window in
 * progress) runs as three recorded HIP graphs whose launches are sized for
 * capacities and read their counts on the device; the host reads the counts
 * twice per step (instead of four times) and a step past a capacity is rolled
 * back and replayed with larger graphs.  Other steps run as before.  Results
 * are bit-identical either way. */
typedef struct vxp_graph_info {
  uint64_t recorded_steps; /* steps run from recorded graphs */
  uint64_t redos;          /* steps rolled back and replayed past a capacity */
  uint64_t recordings;     /* graph sets recorded (shapes, buffers or capacities changed) */
  double recording_ms;     /* host time spent recording them */
  uint32_t cap_pairs, cap_world_pairs, cap_work_items, cap_candidates;
} vxp_graph_info;
int vxp_graph_enable(vxp_system* sys, int on);
int vxp_graph_info_get(vxp_system* sys, vxp_graph_info* out);
/* The server's batch steps recorded (as vxp_graph_enable on its batch system). */
int vxp_server_graph_enable(vxp_server* srv, int on);
int vxp_server_graph_info_get(vxp_server* srv, vxp_graph_info* out);

/* Last error message of this thread (never NULL). */
const char* vxp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VXP_H */
