// vxp_kernels.hpp — host-callable launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "vxp_dev.hpp"

namespace vxd {

void launch_update_begin(hipStream_t s, vxp_stats* st, uint32_t nw, float acc);
void launch_step_begin(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       vxp_stats* st, uint32_t nw, uint32_t* step_ctr);
void launch_fat_aabb(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const DevColliders& C,
                     const WorldState* ws, uint32_t n_slots, uint8_t* moved);
void launch_sweep_keys(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       uint64_t* keys, uint32_t* vals, uint32_t n_slots, uint32_t nw);
// Per-world block sort of the sweep order (k_world_axis included); false when
// max_bodies is above the largest block sort and the device-wide sort is needed.
bool launch_sweep_sort(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       uint32_t* order, uint32_t nw);
void launch_sweep_count(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const WorldState* ws,
                        const uint32_t* order, uint32_t n_slots, float4* slo, float4* shi,
                        uint32_t* send, uint32_t* wcnt);
uint32_t sweep_waves();
// The largest count of pairs one world emitted this step into *max_out (from
// the per-leaf offsets when given, LBVH; else the counts the sweep's count
// pass added when per_world: several worlds; one world's count is the total).
void launch_world_pairs(hipStream_t s, WorldState* ws, uint32_t nw, const uint32_t* leaf_off,
                        uint32_t MB, uint32_t* max_out);
// Each world's packed pair keys (raw: 32-bit when narrow) sorted in one
// workgroup into canonical keys; false (nothing launched) when a world holds
// more than 16384 pairs or the ids need more than 32 key bits.
//
// Launchers that take `const uint32_t* n_dev` size their grid for n items and
// act on min(*n_dev, n) when n_dev is given (a step recorded for replay, whose
// counts exist only on the device); nullptr: exactly n.
bool launch_segment_sort(hipStream_t s, const void* raw, bool narrow, uint32_t n, uint32_t nw,
                         uint32_t id_bits, uint32_t max_world, uint64_t* keys,
                         const uint32_t* n_dev = nullptr);
void launch_sweep_pairs(hipStream_t s, WorldState* ws, const float4* slo, const float4* shi,
                        uint32_t n_slots, const uint32_t* send, const uint32_t* woff_i,
                        uint32_t* wcount, const uint32_t* woff_w, uint64_t* out, bool count,
                        uint32_t id_bits, bool narrow,
                        bool per_world, uint32_t cap = 0xFFFFFFFFu);  // cap: output slots
// LBVH broad phase (per-world trees at each world's slot offsets).
struct LbvhBufs {
  float4 *llo, *lhi, *ilo, *ihi;  // leaves in Morton order; internal nodes
  uint32_t* code;                 // leaf Morton codes
  uint2 *child, *range;           // internal nodes: children, leaf range
  uint32_t *iparent, *lparent, *flag;
};
// Per-world block sort of the Morton order (k_world_axis included); false
// when max_bodies is above the largest block sort.
bool launch_lbvh_sort(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                      uint32_t* order, uint32_t nw);
// (world << 32 | code) keys for the device-wide sort (k_world_axis included).
void launch_lbvh_keys(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                      uint64_t* keys, uint32_t* vals, uint32_t n_slots, uint32_t nw);
// Leaves, internal nodes and their boxes from the Morton order.
void launch_lbvh_build(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const WorldState* ws,
                       const uint32_t* order, uint32_t n_slots, const LbvhBufs& L);
// Count (cnt[0..n_slots], cnt[n_slots] = 0) or write (from off) each leaf's pairs.
void launch_lbvh_pairs(hipStream_t s, const DevCfg& cfg, WorldState* ws, uint32_t n_slots,
                       const LbvhBufs& L, uint32_t* cnt, const uint32_t* off, uint64_t* out,
                       uint32_t id_bits, bool narrow);
void launch_expand_keys(hipStream_t s, uint64_t* keys, uint32_t n, uint32_t id_bits,
                        const uint32_t* n_dev = nullptr);
void launch_expand_keys32(hipStream_t s, const uint32_t* in, uint64_t* keys, uint32_t n,
                          uint32_t id_bits, const uint32_t* n_dev = nullptr);
void launch_pair_process(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                         const DevColliders& C, WorldState* ws, vxp_stats* st,
                         const uint64_t* keys, uint32_t n, PairRec* pr, const DevManifold* prev,
                         const uint8_t* drop, uint32_t* cflag, uint32_t* produce,
                         uint32_t* p_vchk, uint32_t* wlist, uint32_t* step_ctr, uint32_t* counts,
                         bool chunk_bound, const uint32_t* n_dev = nullptr);
// counts[p] (the probe's work items of pair p) is set for the misses.
void launch_cache_replay(hipStream_t s, const DevCfg& cfg, WorldState* ws, vxp_stats* st,
                         PairRec* pr, CacheKeyRec* keys, const uint32_t* cflag,
                         const uint32_t* cidx, const uint32_t* order, uint32_t* clist,
                         uint32_t n_pairs, uint32_t nw, uint32_t* counts,
                         const uint32_t* n_dev = nullptr);
// Broad-phase trees (pair order VXP_PAIR_ORDER_TREE): per world, re-insert the
// leaves of the dynamic tree whose fat box moved, in slot order.
// host[1..2] = *a, *b (0 when null), then host[0] = seq (host-visible memory).
void launch_publish(hipStream_t s, uint32_t* host, const uint32_t* a, const uint32_t* b,
                    uint32_t seq);
void launch_publish4(hipStream_t s, uint32_t* host, const uint32_t* w0, const uint32_t* w1,
                     const uint32_t* w2, const uint32_t* w3, uint32_t seq);
// Recorded step sizes: out[0] = probe work items (after the chunk scan),
// out[1] = contact candidates (after the candidate scan).
void launch_rec_total_wi(hipStream_t s, const uint32_t* wi_off, const uint32_t* counts,
                         const uint32_t* npairs, uint32_t* out);
void launch_rec_total_cand(hipStream_t s, const uint32_t* wi_cbase, uint32_t* out);
// out[3] = npairs if a world holds more than `limit` pairs (else 0), out[4] the reverse.
void launch_rec_gate(hipStream_t s, const uint32_t* npairs, const uint32_t* maxw, uint32_t limit,
                     uint32_t* out);
// out[2] = manifolds carried over by the skipped worlds (after the step begins).
void launch_rec_carried(hipStream_t s, const WorldState* ws, uint32_t nw, uint32_t* out);
void launch_tree_update(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const WorldState* ws,
                        const TreeSet& T, const uint8_t* moved, uint32_t nw, size_t lds_limit);
// Pairs in tree order: count (out == nullptr) or write packed keys
// (world, a, b) and their positions; counts/offsets per (world, query, node).
void launch_tree_pairs(hipStream_t s, WorldState* ws, const TreeSet& T, uint32_t nw,
                       uint32_t* counts, const uint32_t* offs, uint64_t* out, uint32_t* pos,
                       uint32_t id_bits);
// The pair cap in the narrow phase's order: awake[order position] of each
// pair (not both bodies asleep), then drop[p] for awake pairs at or past the
// cap within their world (rank = scanned awake flags).
void launch_pair_awake(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const uint64_t* keys,
                       const uint32_t* perm, uint32_t n, uint32_t* awake);
void launch_pair_drop(hipStream_t s, const uint64_t* keys, const uint32_t* perm, uint32_t n,
                      const uint32_t* awake, const uint32_t* rank, uint32_t cap, uint8_t* drop);
// out[i] = in[idx[i]] for i < n, out[n] = 0; inv[perm[p]] = p
// enable_threading: per world, the slots the thread-pool narrow phase sees
// and whether its parallel branch runs (rank: exclusive scan of the
// not-both-asleep flags, n + 1 entries; cap 0 = none).
void launch_thread_mode(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                        const uint64_t* keys, uint32_t n, const uint32_t* rank, uint32_t cap,
                        uint32_t nw);
void launch_gather_flags(hipStream_t s, const uint32_t* in, const uint32_t* idx, uint32_t n,
                         uint32_t* out);
void launch_invert(hipStream_t s, const uint32_t* perm, uint32_t n, uint32_t* inv);
// wi_pair[w] = the pair of work item w (after the scan of the chunk counts);
// wi_ccnt[w] = 0 for the probe's count pass to add to
void launch_wi_fill(hipStream_t s, const uint32_t* counts, const uint32_t* wi_off, uint32_t n_pairs,
                    uint32_t* wi_pair, uint32_t* wi_ccnt, const uint32_t* n_dev = nullptr);
// Launch shape of the probe: threads per workgroup (64, 128 or 256; each wave
// takes a contiguous range of work items) and the run of consecutive
// workgroups kept on one XCD.
struct ProbeShape {
  uint32_t threads, xcd_run;
};
// The probe's count pass over total_wi work items (lane_n, wi_ccnt, and each
// pair's voxels_checked added to p_vchk;
// the lanes with candidates appended to lane_list as work item * 64 + lane,
// their number at *list_n, which starts at 0; total_wi below 2^26) and its
// write pass over the n_items entries of lane_list (*n_dev when given, at
// most n_items).
void launch_probe_count(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                        const DevColliders& C, const PairRec* pr, const uint32_t* wi_pair,
                        const uint32_t* wi_off, uint32_t total_wi, uint32_t* wi_ccnt,
                        uint32_t* p_vchk, uint8_t* lane_n, uint32_t* lane_list,
                        uint32_t* list_n, ProbeShape shape, const uint32_t* n_dev = nullptr);
void launch_probe_write(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                        const DevColliders& C, const PairRec* pr, const uint32_t* wi_pair,
                        const uint32_t* wi_off, const uint32_t* lane_list, uint32_t n_items,
                        const uint32_t* wi_cbase, const uint8_t* lane_n, Cand* cand,
                        const uint32_t* n_dev = nullptr);
void launch_reduce(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const DevColliders& C,
                   WorldState* ws,
                   vxp_stats* st, PairRec* pr, const uint64_t* keys, uint32_t n,
                   const uint32_t* wlist, uint32_t* step_ctr,
                   const uint32_t* wi_off, const uint32_t* wi_cbase, const uint32_t* wi_ccnt,
                   const uint32_t* p_vchk, const Cand* cand, const DevManifold* prev,
                   const CacheKeyRec* ckeys, CacheVal* cvals, CacheVal* cstage,
                   DevManifold* staging, uint32_t* produce, RStat* rstat, uint32_t* coop);
void launch_world_segments(hipStream_t s, WorldState* ws, uint32_t nw, uint32_t* total);
void launch_scatter(hipStream_t s, const WorldState* ws, const PairRec* pr, uint32_t n,
                    const uint32_t* produce, const uint32_t* pidx, const DevManifold* staging,
                    DevManifold* out, const uint32_t* n_dev = nullptr);
void launch_copy_skipped(hipStream_t s, const WorldState* ws, const DevManifold* prev,
                         DevManifold* out, uint32_t nw);
void launch_islands(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                    vxp_stats* st, const DevManifold* ml, const DevManifold* prev,
                    const IslandArrays& I, uint32_t nw, uint32_t max_manifolds, size_t lds_limit);
void launch_ref_coloring(hipStream_t s, const DevCfg& cfg, const WorldState* ws,
                         const DevManifold* ml, uint64_t* masks, uint8_t* colors, uint32_t nw);
// The solver in three steps: prep (block / world), units (persistent grid over
// the solve units, body velocities in LDS when they fit in lds_limit), tail
// (position passes 2.., then writeback / sleep / stats per world).
void launch_body_state(hipStream_t s, const DevCfg& cfg, const DevBodies& B,
                       const WorldState* ws, const SolverScratch& S, uint32_t n_slots);
void launch_solve_prep(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       vxp_stats* st, DevManifold* ml, const SolverScratch& S, uint32_t* prio,
                       uint32_t* step_ctr, uint32_t nw, uint32_t max_manifolds,
                       size_t lds_limit, const uint32_t* man_total);  // man_total: all manifolds
// Reorders the solve-unit queue, largest units first (tmp: queue-sized scratch).
void launch_unit_order(hipStream_t s, const SolverScratch& S, const uint32_t* step_ctr,
                       uint32_t* tmp);
void launch_solve_units(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                        DevManifold* ml, const SolverScratch& S, uint32_t* step_ctr, uint32_t nw,
                        uint32_t n_cu, size_t lds_limit, uint32_t threads);
void launch_solve_tail(hipStream_t s, const DevCfg& cfg, const DevBodies& B, WorldState* ws,
                       vxp_stats* st, const SolverScratch& S, const uint32_t* step_ctr,
                       uint32_t nw, uint32_t n_cu);
// per-brick voxel masks of the brick pool (non-empty, corner, edge)
void launch_brick_masks(hipStream_t s, const uint8_t* vox, uint32_t nb, uint64_t* mask);
// collider build (per voxel on the device; the brick bookkeeping is the host's)
void launch_collider_classify(hipStream_t s, const uint64_t* solid, const int32_t* nbr,
                              uint32_t nb, uint8_t* out);
void launch_collider_meta(hipStream_t s, const int32_t* pos, const uint8_t* vox, uint32_t nb,
                          float vs, uint32_t* cnt, int32_t* bounds, float* box);
// computeMassProperties of n_jobs (collider, density) jobs, one workgroup each:
// bricks staged in pos (3 ints) / vox (512 bytes), kMassOut floats out per job
void launch_collider_mass(hipStream_t s, const MassJob* jobs, uint32_t n_jobs, const int32_t* pos,
                          const uint8_t* vox, float* out);
void launch_update_end(hipStream_t s, const DevCfg& cfg, const DevBodies& B, vxp_stats* st,
                       uint32_t nw, uint32_t substeps);

// sort / scan primitives (vxp_prims.hip)
size_t sort_pairs_u64_u32_bytes(uint32_t n, uint32_t end_bit);
hipError_t sort_pairs_u64_u32(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout,
                              const uint32_t* vin, uint32_t* vout, uint32_t n, uint32_t end_bit,
                              hipStream_t s);
size_t sort_keys_u64_bytes(uint32_t n, uint32_t end_bit);
hipError_t sort_keys_u64(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout, uint32_t n,
                         uint32_t end_bit, hipStream_t s);
size_t sort_keys_u32_bytes(uint32_t n, uint32_t end_bit);
hipError_t sort_keys_u32(void* tmp, size_t bytes, const uint32_t* kin, uint32_t* kout, uint32_t n,
                         uint32_t end_bit, hipStream_t s);
size_t exclusive_scan_u32_bytes(uint32_t n);
// Device-count forms (sized for `cap` items, acting on min(*n_dev + n_add, cap)).
hipError_t exclusive_scan_u32_n(void* tmp, const uint32_t* in, uint32_t* out, uint32_t cap,
                                const uint32_t* n_dev, uint32_t n_add, hipStream_t s);
hipError_t sort_keys_u64_n(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout,
                           uint32_t cap, const uint32_t* n_dev, uint32_t end_bit, hipStream_t s);
hipError_t sort_keys_u32_n(void* tmp, size_t bytes, const uint32_t* kin, uint32_t* kout,
                           uint32_t cap, const uint32_t* n_dev, uint32_t end_bit, hipStream_t s);
hipError_t exclusive_scan_u32(void* tmp, size_t bytes, const uint32_t* in, uint32_t* out,
                              uint32_t n, hipStream_t s);

void launch_copy_many(hipStream_t s, const CopyDesc* d, uint32_t n);
void launch_rekey_manifolds(hipStream_t s, DevManifold* ml, const uint32_t* ranges, uint32_t n);
void launch_clear_slots(hipStream_t s, const DevCfg& cfg, const DevBodies& B, const uint32_t* slots,
                        uint32_t n);

}  // namespace vxd
