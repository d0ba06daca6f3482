// vxp_host.hpp — host runtime internals shared by the C ABI (vxp_api.cpp)
// and the serving loop (vxp_serve.cpp): device buffers, host colliders, the
// system object with its host mirror of the body SoA, and the step driver.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>
#include <initializer_list>

#include "../../include/vxp.h"
#include "vxp_dev.hpp"
#include "vxp_kernels.hpp"

using namespace vxm;
using namespace vxd;

namespace vxh {
// Last error message of the calling thread.
std::string& lastError();
}  // namespace vxh
#define g_err (::vxh::lastError())

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      g_err = std::string(#expr) + ": " + hipGetErrorString(e_);                         \
      return VXP_ERR_DEVICE;                                                             \
    }                                                                                    \
  } while (0)

namespace vxh {

// Growable device buffer (contents are not preserved on growth unless asked).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need, bool keep = false, hipStream_t s = nullptr) {
    if (need <= bytes && p) return hipSuccess;
    size_t nb = std::max(need, bytes * 2);
    if (nb < 256) nb = 256;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, nb);
    if (e != hipSuccess) return e;
    if (keep && p && bytes) {
      e = hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
    }
    if (p) (void)hipFree(p);
    p = q;
    bytes = nb;
    return hipSuccess;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct MassProps {
  float inv_mass;
  M3 inv_inertia;
};

// ---- collider (src/collider.zig) -------------------------------------------
struct HostCollider {
  float vs = 0.1f;
  std::vector<int32_t> pos;   // merged bricks, ascending (z, y, x)
  std::vector<uint8_t> vox;   // 512 per merged brick
  std::unordered_map<uint64_t, int32_t> by_hash;
  int32_t minb[3] = {0, 0, 0}, maxb[3] = {0, 0, 0};
  float lo[3], hi[3];
  uint32_t solid = 0;
  int32_t org[3], dim[3];
  std::vector<int32_t> dir;
  // computeMassProperties results by density bits (a pure function of both)
  std::unordered_map<uint32_t, MassProps> mass_cache;
};


// defined in vxp_api.cpp
int stageCollider(vxp_system* s, const HostCollider& c);
int colliderMetadata(vxp_system* s, HostCollider& c);
int massProperties(vxp_system* s, const HostCollider& c, float density, MassProps* out, V3* com);
int resolveMasses(vxp_system* s);
// defined in vxp_noalloc.hip
int noallocInit(int device);

}  // namespace vxh

using namespace vxh;

struct vxp_system {
  vxp_config cfg;
  uint32_t nw = 1, MB = 0, id_cap = 0;
  hipStream_t stream = nullptr;
  // Second stream for the island labels, which run beside the solver's prep
  // and units: ev_fork orders them after the manifolds, ev_join orders the
  // solver's writeback (and everything after it) after them.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool side_pending = false;  // ev_join recorded, the main stream not yet joined
  float acc = 0.0f;

  std::vector<HostCollider> colliders;
  bool coll_dirty = true;
  DevBuf d_hdr, d_dir, d_vox, d_mask;

  // host mirror of the body SoA
  std::vector<float4> pos, rot, lv, av, fat_lo, fat_hi;
  std::vector<float> inv_mass, iil, fric, rest, ldamp, adamp, sleep_timer;
  std::vector<uint8_t> type, sleeping, active;
  std::vector<int32_t> collider;
  std::vector<uint32_t> id;
  std::vector<int32_t> idmap;
  std::vector<uint32_t> generation, next_id, body_count;
  std::vector<std::vector<uint32_t>> free_list;
  std::vector<uint32_t> n_total, n_dyn, n_static;
  bool host_dirty = true, dev_newer = false;

  DevBuf b_pos, b_rot, b_lv, b_av, b_fat_lo, b_fat_hi, b_inv_mass, b_iil, b_fric, b_rest,
      b_ldamp, b_adamp, b_sleep_timer, b_type, b_sleeping, b_active, b_collider, b_id, b_idmap;

  // per world / per step device state
  DevBuf d_ws, d_stats, d_cache_keys, d_cache_vals;
  DevBuf d_skeys, d_svals, d_skeys2, d_svals2, d_sort_tmp;
  DevBuf d_slo, d_shi, d_send, d_wcnt, d_woff, d_swc, d_swo, d_pairs_raw, d_pairs, d_pr, d_counts, d_wi_off, d_wi_cbase, d_wi_ccnt, d_pvchk, d_lane_n,
      d_cand,
      d_produce, d_pidx, d_staging, d_cstage, d_total, d_rstat, d_clist, d_cflag, d_cidx, d_wlist, d_step_ctr;
  DevBuf d_lb_ilo, d_lb_ihi, d_lb_code, d_lb_child, d_lb_range, d_lb_iparent, d_lb_lparent,
      d_lb_flag;  // LBVH broad phase
  DevBuf d_maxw;  // the most pairs one world emitted this step
  DevBuf d_ml[2];
  int cur = 0;
  DevBuf d_isl_parent, d_isl_end, d_isl_size, d_ref_masks, d_ref_colors;
  DevBuf d_isl_timer, d_isl_sleep;  // island sleep state (persistent)
  uint32_t isl_cap = 0;             // id_cap the island arrays are laid out with
  // broad-phase trees (pair order VXP_PAIR_ORDER_TREE): host mirror, device
  // copy, and the per-step buffers of the tree order and the pair cap
  std::vector<TreeNode> t_nodes;
  std::vector<TreeHead> t_heads;
  std::vector<uint32_t> t_leaf;
  uint32_t t_nc = 0;
  bool t_host_dirty = false, t_dev_newer = false;
  DevBuf d_t_nodes, d_t_heads, d_t_leaf, d_t_moved, d_t_cnt, d_t_off, d_t_pos, d_perm, d_inv,
      d_pairs_tree, d_awake, d_arank, d_drop, d_cflag_r, d_cidx_r;
  DevBuf d_coop;     // work-list pairs left to the reduction's waves (any order)
  DevBuf d_wi_pair;  // the pair of each probe work item
  DevBuf d_wi_list;  // probe lanes with candidates (work item * 64 + lane), in no particular order
  DevBuf s_pos, s_rot, s_lv, s_av, s_ii, s_adj_cnt, s_adj_off, s_adj, s_color, s_order, s_rec,
      s_pts, s_prio, s_tstamp, s_upar, s_ucnt, s_uoff, s_ubody, s_uorder, s_units, s_uqueue, s_uqueue2,
      s_pcnt, s_fq, s_cend;
  uint32_t n_pairs = 0, n_manifolds = 0;
  uint32_t last_wi = 0, last_cand = 0;  // probe work items / candidates of the last step
  // Recorded step (vxp_graph_enable): the step as three HIP graphs per
  // manifold-buffer parity, sized for capacities (pairs, pairs of one world,
  // probe work items, contact candidates) with the counts read on the device.
  // Segment 1 ends with the counts that the capacities bound, read by the host
  // before segment 2 (a step past a capacity is redone from the world-state
  // and stats snapshot that segment 1 takes first); segment 2 ends with the
  // candidate count (past its capacity: segment 3 is recorded again, larger).
  struct RecGraph {
    hipGraphExec_t g[3] = {nullptr, nullptr, nullptr};
    uint64_t sig = 0;  // the buffers, shapes and capacities it was recorded with
  };
  bool rec_on = false;
  RecGraph rec[2];
  uint32_t rc_pairs = 0, rc_wp = 0, rc_wi = 0, rc_cand = 0, rc_carry = 0;
  uint32_t rc_mm = 0;         // most manifolds of one world the recorded LDS requests assume
  uint32_t last_max_man = 0;  // most manifolds of one world in the last step
  uint64_t rec_steps = 0, rec_redos = 0, rec_records = 0;
  double rec_ms = 0.0;  // host time spent recording
  CopyDesc rec_desc[4] = {};  // the snapshot copies as last uploaded to d_rdesc
  bool rec_nman_pending = false;  // n_manifolds awaits the last recorded step's count
  DevBuf d_rsnap, d_rdesc, d_rsize;
  uint32_t* h_read = nullptr;  // pinned host words for the per-step size reads: [seq, a, b]
  uint32_t* d_read = nullptr;  // their device address
  uint32_t read_seq = 0;
  // Online tuning of launch shapes that do not change results: during a window
  // of frames the candidates take turns (one per frame, timed with events
  // around the solver); the fastest on average is kept until the next window.
  // Online tuner of the launch shapes (every choice gives the same bits).  A
  // window times the candidates of one knob in turn, round-robin over 3
  // passes, on the phase the knob shapes, and keeps the fastest; a round
  // tunes every knob that is not forced, one window after the other.
  enum Knob { kPairColorMax, kUnitThreads, kBroadPhase, kProbeThreads, kProbeXcdRun, kKnobs };
  enum Phase { kPhSolve, kPhPairs, kPhProbe };
  struct Tuner {
    std::vector<uint32_t> cand[kKnobs] = {
        {0, 64, 128, 256, 512},  // lane-pair colour limit
        {256, 512},              // threads per solve-unit workgroup
        {1, 2},                  // VXP_BROAD_SWEEP, VXP_BROAD_LBVH
        {64, 128, 256},          // threads per probe workgroup
        {8, 64, 512}};           // probe workgroups per XCD run
    // the probe's one-wave workgroups: 6-8 % faster count passes than 256
    // threads on both the 4096-body scene and the 512-world batch
    uint32_t pick[kKnobs] = {128, 0, 1, 64, 64};   // pick[kUnitThreads] 0: by world count
    uint32_t forced[kKnobs] = {0, 0, 0, 0, 0};     // 0: tuned
    uint32_t use[kKnobs] = {128, 512, 1, 64, 64};  // this step's values
    std::vector<double> ms;
    std::vector<uint32_t> n;
    uint32_t frame = 0, window_left = 0, next_window = 0, windows = 0, knob = 0;
    bool fixed = false, pending = false, marked_a = false, marked_b = false;
    uint32_t pending_idx = 0;
    hipEvent_t a = nullptr, b = nullptr;
  } tune;
  size_t lds_limit = 0;  // dynamic LDS available to the solver block
  uint32_t n_cu = 256;   // compute units of the device
  int device = 0;        // the device current at creation (the streams' device)
  // collider build scratch (classification, metadata, mass properties)
  DevBuf d_cb_pos, d_cb_vox, d_cb_cnt, d_cb_bounds, d_cb_box, d_cb_out, d_cb_solid, d_cb_nbr,
      d_cb_jobs;
  // dynamic bodies created since the last batch of mass-property jobs: their
  // inverse mass and inertia are filled in by resolveMasses (one launch for
  // every (collider, density) not yet computed)
  struct PendingMass {
    size_t g;
    uint32_t collider;
    float density;
  };
  std::vector<PendingMass> mass_pending;
  // kernel timing (HIP events on `stream`)
  int prof = VXP_PROFILE_OFF;
  double prof_ms[VXP_K_COUNT] = {};
  uint64_t prof_n[VXP_K_COUNT] = {};
  std::vector<hipEvent_t> ev_pool;
  struct Pending {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
};


namespace vxh {
size_t slots(const vxp_system* s);
int uploadColliders(vxp_system* s);
int ensureDevice(vxp_system* s);
int ensureIslands(vxp_system* s);
int ensureHost(vxp_system* s);
DevBodies devBodies(vxp_system* s);
DevColliders devColliders(vxp_system* s);
DevCfg devCfg(vxp_system* s);
int stepAll(vxp_system* s);
int checkErrors(vxp_system* s);
int streamWait(vxp_system* s);
int readSizes(vxp_system* s, std::initializer_list<const uint32_t*> src, uint32_t* out);
int updateOnce(vxp_system* s, float dt);
int settleManifolds(vxp_system* s);  // n_manifolds after a recorded step
}  // namespace vxh

#define RC(x)            \
  do {                   \
    int rc_ = (x);       \
    if (rc_) return rc_; \
  } while (0)
