// vxp_api.cpp — host runtime and C ABI (include/vxp.h) of the MI355X step
// pipeline.  The host keeps the reference's bookkeeping (body slots with
// generations and a free list, monotonically increasing body ids, colliders
// built from setCollisionVoxel writes, the fixed-step accumulator); all
// per-step work runs in the gfx950 kernels of vxp_kernels.hip on one HIP
// stream.  Body state lives on the device between steps; getters copy it out
// and setters copy it in before the next step.
#include <chrono>

#include "vxp_host.hpp"

namespace vxh {
std::string& lastError() {
  static thread_local std::string e;
  return e;
}

// The merged bricks of a collider, uploaded for the collider-build kernels.
int stageCollider(vxp_system* s, const HostCollider& c) {
  const size_t nb = c.pos.size() / 3;
  HIP_TRY(s->d_cb_pos.ensure(std::max<size_t>(nb * 12, 16)));
  HIP_TRY(s->d_cb_vox.ensure(std::max<size_t>(nb * 512, 512)));
  if (nb) {
    HIP_TRY(hipMemcpyAsync(s->d_cb_pos.p, c.pos.data(), nb * 12, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_cb_vox.p, c.vox.data(), nb * 512, hipMemcpyHostToDevice, s->stream));
  }
  return VXP_OK;
}

// VoxelCollider.recomputeMetadata (src/collider.zig:90-133, 283-333): the
// per-voxel work (solid counts, voxel bounds, brick AABBs) runs on the device
// per brick; the per-brick results are combined here (sums, minima, maxima:
// exact in any order) and the dense brick directory is built over the bounds.
int colliderMetadata(vxp_system* s, HostCollider& c) {
  c.minb[0] = c.minb[1] = c.minb[2] = INT32_MAX;
  c.maxb[0] = c.maxb[1] = c.maxb[2] = INT32_MIN;
  c.lo[0] = c.lo[1] = c.lo[2] = INFINITY;
  c.hi[0] = c.hi[1] = c.hi[2] = -INFINITY;
  c.solid = 0;
  const size_t nb = c.pos.size() / 3;
  if (nb) {
    RC(stageCollider(s, c));
    HIP_TRY(s->d_cb_cnt.ensure(nb * 4));
    HIP_TRY(s->d_cb_bounds.ensure(nb * 24));
    HIP_TRY(s->d_cb_box.ensure(nb * 24));
    launch_collider_meta(s->stream, s->d_cb_pos.as<int32_t>(), s->d_cb_vox.as<uint8_t>(),
                         (uint32_t)nb, c.vs, s->d_cb_cnt.as<uint32_t>(),
                         s->d_cb_bounds.as<int32_t>(), s->d_cb_box.as<float>());
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> cnt(nb);
    std::vector<int32_t> bounds(nb * 6);
    std::vector<float> box(nb * 6);
    HIP_TRY(hipMemcpyAsync(cnt.data(), s->d_cb_cnt.p, nb * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(bounds.data(), s->d_cb_bounds.p, nb * 24, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(box.data(), s->d_cb_box.p, nb * 24, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t b = 0; b < nb; b++) {
      if (!cnt[b]) continue;
      c.solid += cnt[b];
      for (int k = 0; k < 3; k++) {
        c.minb[k] = std::min(c.minb[k], bounds[6 * b + k]);
        c.maxb[k] = std::max(c.maxb[k], bounds[6 * b + 3 + k]);
        c.lo[k] = fminf(c.lo[k], box[6 * b + k]);
        c.hi[k] = fmaxf(c.hi[k], box[6 * b + 3 + k]);
      }
    }
  }
  if (c.solid == 0) {
    for (int k = 0; k < 3; k++) c.minb[k] = c.maxb[k] = 0;
  }
  // dense directory over brick positions floor(min/8) .. floor(max/8)
  size_t cells = 1;
  for (int k = 0; k < 3; k++) {
    c.org[k] = floorDiv8(c.minb[k]);
    c.dim[k] = floorDiv8(c.maxb[k]) - c.org[k] + 1;
    cells *= (size_t)c.dim[k];
  }
  c.dir.assign(cells, -1);
  for (int32_t z = 0; z < c.dim[2]; z++)
    for (int32_t y = 0; y < c.dim[1]; y++)
      for (int32_t x = 0; x < c.dim[0]; x++) {
        auto it = c.by_hash.find(hashBrickPos(c.org[0] + x, c.org[1] + y, c.org[2] + z));
        if (it != c.by_hash.end())
          c.dir[((size_t)z * c.dim[1] + y) * c.dim[0] + x] = it->second;
      }
  return VXP_OK;
}

// computeMassProperties (src/body.zig:462-519) of a batch of (collider,
// density) jobs on the device: one k_collider_mass workgroup per job, every
// job's bricks staged once, one launch and one readback for the whole batch.
// out[j] receives kMassOut floats (inv_mass, inverse inertia, centre of mass).
// Colliders without solid voxels are not launched: zero mass properties.
static int massBatch(vxp_system* s, const std::vector<std::pair<const HostCollider*, float>>& jobs,
                     std::vector<float>& out) {
  out.assign(jobs.size() * kMassOut, 0.0f);
  std::vector<MassJob> dj;
  std::vector<uint32_t> slot;  // out index of each launched job
  std::unordered_map<const HostCollider*, uint32_t> off;
  std::vector<int32_t> pos;
  std::vector<uint8_t> vox;
  for (size_t j = 0; j < jobs.size(); j++) {
    const HostCollider* c = jobs[j].first;
    if (c->solid == 0) continue;
    auto it = off.find(c);
    if (it == off.end()) {
      it = off.emplace(c, (uint32_t)(pos.size() / 3)).first;
      pos.insert(pos.end(), c->pos.begin(), c->pos.end());
      vox.insert(vox.end(), c->vox.begin(), c->vox.end());
    }
    MassJob m{};
    m.brick_off = it->second;
    m.nb = (uint32_t)(c->pos.size() / 3);
    m.solid = c->solid;
    m.vs = c->vs;
    m.density = jobs[j].second;
    dj.push_back(m);
    slot.push_back((uint32_t)j);
  }
  if (dj.empty()) return VXP_OK;
  const size_t nb = pos.size() / 3;
  HIP_TRY(s->d_cb_pos.ensure(std::max<size_t>(nb * 12, 16)));
  HIP_TRY(s->d_cb_vox.ensure(std::max<size_t>(nb * 512, 512)));
  HIP_TRY(s->d_cb_jobs.ensure(dj.size() * sizeof(MassJob)));
  HIP_TRY(s->d_cb_out.ensure(dj.size() * kMassOut * 4));
  HIP_TRY(hipMemcpyAsync(s->d_cb_pos.p, pos.data(), nb * 12, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->d_cb_vox.p, vox.data(), nb * 512, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->d_cb_jobs.p, dj.data(), dj.size() * sizeof(MassJob), hipMemcpyHostToDevice,
                         s->stream));
  launch_collider_mass(s->stream, s->d_cb_jobs.as<MassJob>(), (uint32_t)dj.size(),
                       s->d_cb_pos.as<int32_t>(), s->d_cb_vox.as<uint8_t>(), s->d_cb_out.as<float>());
  HIP_TRY(hipGetLastError());
  std::vector<float> r(dj.size() * kMassOut);
  HIP_TRY(hipMemcpyAsync(r.data(), s->d_cb_out.p, r.size() * 4, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (size_t k = 0; k < dj.size(); k++)
    std::memcpy(&out[(size_t)slot[k] * kMassOut], &r[k * kMassOut], kMassOut * 4);
  return VXP_OK;
}

static MassProps massFromOut(const float* r) {
  MassProps m;
  m.inv_mass = r[0];
  for (int i = 0; i < 9; i++) m.inv_inertia.m[i] = r[1 + i];
  return m;
}

// One collider at one density.  com may be null.
int massProperties(vxp_system* s, const HostCollider& c, float density, MassProps* out, V3* com) {
  std::vector<float> r;
  RC(massBatch(s, {{&c, density}}, r));
  *out = massFromOut(r.data());
  if (com) *com = V3{r[10], r[11], r[12]};
  return VXP_OK;
}

// The mass properties of every dynamic body created since the last call
// (createDynamicBody, src/system.zig:294-322): the (collider, density) pairs
// not yet known are computed in one batch, then the bodies' inverse mass and
// local inverse inertia are filled in on the host mirror.
int resolveMasses(vxp_system* s) {
  if (s->mass_pending.empty()) return VXP_OK;
  std::vector<std::pair<const HostCollider*, float>> jobs;
  std::vector<std::pair<uint32_t, uint32_t>> keys;  // (collider, density bits) of each job
  std::unordered_set<uint64_t> seen;
  for (const auto& p : s->mass_pending) {
    uint32_t key;
    std::memcpy(&key, &p.density, 4);
    const HostCollider& hc = s->colliders[p.collider];
    if (hc.mass_cache.count(key) || !seen.insert((uint64_t)p.collider << 32 | key).second) continue;
    keys.push_back({p.collider, key});
    jobs.push_back({&hc, p.density});
  }
  std::vector<float> r;
  RC(massBatch(s, jobs, r));
  for (size_t j = 0; j < jobs.size(); j++)
    s->colliders[keys[j].first].mass_cache.emplace(keys[j].second, massFromOut(&r[j * kMassOut]));
  for (const auto& p : s->mass_pending) {
    uint32_t key;
    std::memcpy(&key, &p.density, 4);
    const MassProps& mp = s->colliders[p.collider].mass_cache.at(key);
    s->inv_mass[p.g] = mp.inv_mass;
    for (int i = 0; i < 9; i++) s->iil[p.g * 9 + i] = mp.inv_inertia.m[i];
  }
  s->mass_pending.clear();
  s->host_dirty = true;
  return VXP_OK;
}

}  // namespace vxh

namespace vxh {

size_t slots(const vxp_system* s) { return (size_t)s->nw * s->MB; }

int uploadColliders(vxp_system* s) {
  if (!s->coll_dirty) return VXP_OK;
  std::vector<DevCollider> hdr;
  std::vector<int32_t> dir;
  std::vector<uint8_t> vox;
  uint32_t brick_base = 0;
  for (const HostCollider& c : s->colliders) {
    DevCollider h;
    h.vs = c.vs;
    for (int k = 0; k < 3; k++) {
      h.minb[k] = c.minb[k];
      h.maxb[k] = c.maxb[k];
      h.lo[k] = c.lo[k];
      h.hi[k] = c.hi[k];
      h.org[k] = c.org[k];
      h.dim[k] = c.dim[k];
    }
    h.dir_off = (uint32_t)dir.size();
    h.solid = c.solid;
    for (int32_t d : c.dir) dir.push_back(d < 0 ? -1 : d + (int32_t)brick_base);
    vox.insert(vox.end(), c.vox.begin(), c.vox.end());
    brick_base += (uint32_t)(c.vox.size() / 512);
    hdr.push_back(h);
  }
  if (hdr.empty()) hdr.push_back(DevCollider{});
  if (dir.empty()) dir.push_back(-1);
  if (vox.empty()) vox.resize(512, 0);
  HIP_TRY(s->d_hdr.ensure(hdr.size() * sizeof(DevCollider)));
  HIP_TRY(s->d_dir.ensure(dir.size() * 4));
  HIP_TRY(s->d_vox.ensure(vox.size()));
  HIP_TRY(hipMemcpy(s->d_hdr.p, hdr.data(), hdr.size() * sizeof(DevCollider), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_dir.p, dir.data(), dir.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_vox.p, vox.data(), vox.size(), hipMemcpyHostToDevice));
  const uint32_t nb = (uint32_t)(vox.size() / 512);
  HIP_TRY(s->d_mask.ensure((size_t)nb * kMaskWords * 8));
  launch_brick_masks(s->stream, s->d_vox.as<uint8_t>(), nb, s->d_mask.as<uint64_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->coll_dirty = false;
  return VXP_OK;
}

template <class T>
int up(DevBuf& b, const std::vector<T>& v) {
  HIP_TRY(b.ensure(std::max<size_t>(v.size() * sizeof(T), 16)));
  if (!v.empty()) HIP_TRY(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return VXP_OK;
}
template <class T>
int down(const DevBuf& b, std::vector<T>& v) {
  if (!v.empty()) HIP_TRY(hipMemcpy(v.data(), b.p, v.size() * sizeof(T), hipMemcpyDeviceToHost));
  return VXP_OK;
}

int ensureDevice(vxp_system* s) {
  int rc = uploadColliders(s);
  if (rc) return rc;
  if ((rc = resolveMasses(s))) return rc;
  if (s->t_host_dirty) {
    if ((rc = up(s->d_t_nodes, s->t_nodes)) || (rc = up(s->d_t_heads, s->t_heads)) ||
        (rc = up(s->d_t_leaf, s->t_leaf)))
      return rc;
    s->t_host_dirty = false;
  }
  if (!s->host_dirty) return VXP_OK;
  if ((rc = up(s->b_pos, s->pos)) || (rc = up(s->b_rot, s->rot)) || (rc = up(s->b_lv, s->lv)) ||
      (rc = up(s->b_av, s->av)) || (rc = up(s->b_fat_lo, s->fat_lo)) ||
      (rc = up(s->b_fat_hi, s->fat_hi)) || (rc = up(s->b_inv_mass, s->inv_mass)) ||
      (rc = up(s->b_iil, s->iil)) || (rc = up(s->b_fric, s->fric)) || (rc = up(s->b_rest, s->rest)) ||
      (rc = up(s->b_ldamp, s->ldamp)) || (rc = up(s->b_adamp, s->adamp)) ||
      (rc = up(s->b_sleep_timer, s->sleep_timer)) || (rc = up(s->b_type, s->type)) ||
      (rc = up(s->b_sleeping, s->sleeping)) || (rc = up(s->b_active, s->active)) ||
      (rc = up(s->b_collider, s->collider)) || (rc = up(s->b_id, s->id)) ||
      (rc = up(s->b_idmap, s->idmap)))
    return rc;
  s->host_dirty = false;
  return VXP_OK;
}

int ensureHost(vxp_system* s) {
  if (!s->dev_newer) return VXP_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  int rc;
  if ((rc = down(s->b_pos, s->pos)) || (rc = down(s->b_rot, s->rot)) || (rc = down(s->b_lv, s->lv)) ||
      (rc = down(s->b_av, s->av)) || (rc = down(s->b_fat_lo, s->fat_lo)) ||
      (rc = down(s->b_fat_hi, s->fat_hi)) || (rc = down(s->b_sleep_timer, s->sleep_timer)) ||
      (rc = down(s->b_sleeping, s->sleeping)))
    return rc;
  s->dev_newer = false;
  return VXP_OK;
}

// ---- broad-phase trees (pair order VXP_PAIR_ORDER_TREE) ----------------------
// The host mirror is changed by body creation and destruction, the device copy
// by the steps; each side fetches the other's changes before it edits.
static bool treeMode(const vxp_system* s) { return s->cfg.pair_order == VXP_PAIR_ORDER_TREE; }

int treesToHost(vxp_system* s) {
  if (!s->t_dev_newer) return VXP_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(s->t_nodes.data(), s->d_t_nodes.p, s->t_nodes.size() * sizeof(TreeNode),
                    hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(s->t_heads.data(), s->d_t_heads.p, s->t_heads.size() * sizeof(TreeHead),
                    hipMemcpyDeviceToHost));
  s->t_dev_newer = false;
  return VXP_OK;
}

Tree hostTree(vxp_system* s, uint32_t w, uint32_t which) {
  return Tree{s->t_nodes.data() + ((size_t)w * 2 + which) * s->t_nc,
              s->t_heads.data() + (size_t)w * 2 + which, s->t_nc};
}

TreeSet treeSet(vxp_system* s) {
  return TreeSet{s->d_t_nodes.as<TreeNode>(), s->d_t_heads.as<TreeHead>(), s->d_t_leaf.as<uint32_t>(),
                 s->t_nc};
}

DevBodies devBodies(vxp_system* s) {
  DevBodies B;
  B.pos = s->b_pos.as<float4>();
  B.rot = s->b_rot.as<float4>();
  B.lv = s->b_lv.as<float4>();
  B.av = s->b_av.as<float4>();
  B.inv_mass = s->b_inv_mass.as<float>();
  B.iil = s->b_iil.as<float>();
  B.fric = s->b_fric.as<float>();
  B.rest = s->b_rest.as<float>();
  B.ldamp = s->b_ldamp.as<float>();
  B.adamp = s->b_adamp.as<float>();
  B.sleep_timer = s->b_sleep_timer.as<float>();
  B.type = s->b_type.as<uint8_t>();
  B.sleeping = s->b_sleeping.as<uint8_t>();
  B.active = s->b_active.as<uint8_t>();
  B.collider = s->b_collider.as<int32_t>();
  B.id = s->b_id.as<uint32_t>();
  B.fat_lo = s->b_fat_lo.as<float4>();
  B.fat_hi = s->b_fat_hi.as<float4>();
  B.idmap = s->b_idmap.as<int32_t>();
  return B;
}

DevColliders devColliders(vxp_system* s) {
  return DevColliders{s->d_hdr.as<DevCollider>(), s->d_dir.as<int32_t>(), s->d_vox.as<uint8_t>(),
                      s->d_mask.as<uint64_t>()};
}

DevCfg devCfg(vxp_system* s) {
  const vxp_config& c = s->cfg;
  DevCfg d{};
  for (int k = 0; k < 3; k++) d.gravity[k] = c.gravity[k];
  d.dt = c.fixed_timestep;
  d.h = c.fixed_timestep / (float)c.substeps;
  // computeSoftParams (src/solver.zig:57-72)
  if (c.hertz <= 0.0f) {
    d.bias_rate = 0.0f;
    d.mass_coeff = 1.0f;
    d.impulse_coeff = 1.0f;
  } else {
    float omega = 2.0f * kPi * c.hertz;
    float a1 = 2.0f * c.damping_ratio + omega * d.h;
    float a2 = d.h * omega * a1;
    float a3 = 1.0f / (1.0f + a2);
    d.bias_rate = omega / a1;
    d.mass_coeff = a2 * a3;
    d.impulse_coeff = a3;
  }
  d.slop = c.slop;
  d.pcr = c.position_correction_rate;
  d.mpc = c.max_position_correction;
  d.wss = c.warm_start_scale;
  d.substeps = c.substeps;
  d.vel_iters = c.velocity_iterations;
  d.pos_iters = c.position_iterations;
  d.max_bodies = s->MB;
  d.id_cap = s->id_cap;
  d.cache_cap = c.brick_pair_cache_size ? c.brick_pair_cache_size : 1024;
  d.enable_sleeping = c.enable_sleeping;
  d.threaded = c.enable_threading ? 1 : 0;
  d.thread_multi = c.thread_count != 1 ? 1 : 0;
  d.warm_starting = c.warm_starting;
  d.enable_friction = c.enable_friction;
  d.enable_restitution = c.enable_restitution;
  d.disable_rotation = c.disable_rotation;
  d.use_sat = c.use_brick_sat;
  d.use_cache = c.use_brick_pair_cache;
  d.enable_islands = c.enable_islands;
  d.store_impulses = c.enable_warm_starting;
  d.n_cu = s->n_cu;
  return d;
}

// Grow the id-indexed arrays (idmap and island scratch) to hold id `need`.
void growIds(vxp_system* s, uint32_t need) {
  if (need < s->id_cap) return;
  uint32_t nc = std::max(need + 1, s->id_cap * 2);
  std::vector<int32_t> m((size_t)s->nw * nc, -1);
  for (uint32_t w = 0; w < s->nw; w++)
    for (uint32_t i = 0; i < s->id_cap; i++) m[(size_t)w * nc + i] = s->idmap[(size_t)w * s->id_cap + i];
  s->idmap.swap(m);
  s->id_cap = nc;
  s->host_dirty = true;
}

// The island arrays at the current id_cap, zeroed when first allocated; when
// ids have grown since, each world's rows move to the wider layout.
int ensureIslands(vxp_system* s) {
  if (s->isl_cap == s->id_cap && s->d_isl_timer.p) return VXP_OK;
  const size_t n = (size_t)s->nw * s->id_cap * 4;
  DevBuf* bufs[4] = {&s->d_isl_parent, &s->d_isl_end, &s->d_isl_timer, &s->d_isl_sleep};
  for (DevBuf* b : bufs) {
    DevBuf nb;
    HIP_TRY(nb.ensure(n));
    HIP_TRY(hipMemsetAsync(nb.p, 0, n, s->stream));
    if (b->p && s->isl_cap)
      HIP_TRY(hipMemcpy2DAsync(nb.p, (size_t)s->id_cap * 4, b->p, (size_t)s->isl_cap * 4,
                               (size_t)s->isl_cap * 4, s->nw, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    b->release();
    *b = nb;
  }
  HIP_TRY(s->d_isl_size.ensure(n));
  s->isl_cap = s->id_cap;
  return VXP_OK;
}

// Threads per solve-unit workgroup: forced, or 256 when many worlds' units
// share the device (two workgroups fit a CU) and 512 when a few large units
// bound the solve (more lane pairs per colour).
// This step's launch shapes (chosen by tuneStep).
uint32_t unitThreads(const vxp_system* s) { return s->tune.use[vxp_system::kUnitThreads]; }
uint32_t broadPhase(const vxp_system* s) { return s->tune.use[vxp_system::kBroadPhase]; }

IslandArrays islandArrays(vxp_system* s) {
  return IslandArrays{s->d_isl_parent.as<uint32_t>(), s->d_isl_end.as<uint32_t>(),
                      s->d_isl_size.as<uint32_t>(), s->d_isl_timer.as<float>(),
                      s->d_isl_sleep.as<uint32_t>()};
}

hipEvent_t takeEvent(vxp_system* s) {
  if (!s->ev_pool.empty()) {
    hipEvent_t e = s->ev_pool.back();
    s->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
// Brackets a group of launches of kernel `id` with events (when profiling).
// Phase `id` is timed in mode VXP_PROFILE_ALL, and the solve units also in
// VXP_PROFILE_SOLVE_UNITS.
bool profTimed(const vxp_system* s, int id) {
  return s->prof == VXP_PROFILE_ALL || (s->prof == VXP_PROFILE_SOLVE_UNITS && id == VXP_K_SOLVE_UNITS);
}
hipEvent_t profBegin(vxp_system* s, int id) {
  if (!profTimed(s, id)) return nullptr;
  hipEvent_t a = takeEvent(s);
  (void)hipEventRecord(a, s->stream);
  return a;
}
void profEnd(vxp_system* s, int id, hipEvent_t a) {
  if (!a) return;
  hipEvent_t b = takeEvent(s);
  (void)hipEventRecord(b, s->stream);
  s->pending.push_back({id, a, b});
}
// After a stream synchronization: fold the recorded intervals in.
void profFlush(vxp_system* s) {
  for (auto& p : s->pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      s->prof_ms[p.id] += ms;
      s->prof_n[p.id]++;
    }
    s->ev_pool.push_back(p.a);
    s->ev_pool.push_back(p.b);
  }
  s->pending.clear();
}

// Waits for the system's stream by polling: a blocking wait can sleep the host
// thread, which then wakes late and leaves the GPU idle before the next launch.
int streamWait(vxp_system* s) {
  for (;;) {
    hipError_t e = hipStreamQuery(s->stream);
    if (e == hipSuccess) return VXP_OK;
    if (e != hipErrorNotReady) {
      g_err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
      return VXP_ERR_DEVICE;
    }
    __builtin_ia32_pause();
  }
}

// Reads up to two device words: a one-thread kernel, queued behind the work
// that produces them, writes them and then a new sequence number into pinned
// host memory, and the host polls that number (the stream is checked now and
// then so that a failed launch ends the wait).
int readSizes(vxp_system* s, std::initializer_list<const uint32_t*> src, uint32_t* out) {
  const uint32_t* p[2] = {nullptr, nullptr};
  uint32_t n = 0;
  for (const uint32_t* q : src) {
    if (n == 2) {
      g_err = "readSizes: at most two words";
      return VXP_ERR_ARGUMENT;
    }
    p[n++] = q;
  }
  const uint32_t seq = ++s->read_seq;
  launch_publish(s->stream, s->d_read, p[0], p[1], seq);
  HIP_TRY(hipGetLastError());
  for (uint32_t spin = 1;; spin++) {
    if (__atomic_load_n(&s->h_read[0], __ATOMIC_ACQUIRE) == seq) break;
    if ((spin & 1023u) == 0) {
      hipError_t e = hipStreamQuery(s->stream);
      if (e != hipSuccess && e != hipErrorNotReady) {
        g_err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
        return VXP_ERR_DEVICE;
      }
      if (e == hipSuccess && __atomic_load_n(&s->h_read[0], __ATOMIC_ACQUIRE) != seq) {
        g_err = "size read: the stream drained without the published words";
        return VXP_ERR_DEVICE;
      }
    }
    __builtin_ia32_pause();
  }
  for (uint32_t k = 0; k < n; k++) out[k] = __atomic_load_n(&s->h_read[1 + k], __ATOMIC_RELAXED);
  return VXP_OK;
}

static vxp_system::Phase knobPhase(uint32_t k) {
  return k == vxp_system::kBroadPhase ? vxp_system::kPhPairs
         : k >= vxp_system::kProbeThreads ? vxp_system::kPhProbe
                                          : vxp_system::kPhSolve;
}

constexpr uint32_t kTuneRoundGap = 120;  // frames between tuning rounds

// At the start of a step: folds in the previous tuned step's time (its events
// completed before this step begins), ends a window when its samples are in
// (keeping the fastest candidate), starts the next window of the round when
// one is due, and sets this step's value of every knob.
int tuneStep(vxp_system* s) {
  auto& t = s->tune;
  constexpr uint32_t K = vxp_system::kKnobs;
  if (!t.pick[vxp_system::kUnitThreads])  // two 256-thread units fit a CU: many worlds
    t.pick[vxp_system::kUnitThreads] = s->nw >= 64 ? 256u : 512u;
  if (t.pending) {
    if (t.marked_a && t.marked_b) {
      float ms = 0.0f;
      HIP_TRY(hipEventSynchronize(t.b));
      HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
      t.ms[t.pending_idx] += ms;
      t.n[t.pending_idx]++;
    }
    t.pending = false;
    if (t.window_left == 0) {
      // window over: the fastest candidate replaces the value in use only if
      // it is at least 10 % faster (timings of a few live frames are noisy)
      const auto& cand = t.cand[t.knob];
      double best = 1e300, inc = 1e300;
      size_t bi = cand.size();
      for (size_t i = 0; i < cand.size(); i++) {
        if (!t.n[i]) continue;
        const double avg = t.ms[i] / t.n[i];
        if (avg < best) {
          best = avg;
          bi = i;
        }
        if (cand[i] == t.pick[t.knob]) inc = avg;
      }
      if (bi < cand.size() && best < 0.9 * inc) t.pick[t.knob] = cand[bi];
      t.windows++;
      // the next knob of the round at once; after the last, a new round later
      // (the best shapes change as a scene's piles form and its islands grow)
      uint32_t k = t.knob + 1;
      while (k < K && t.forced[k]) k++;
      if (k < K) {
        t.knob = k;
        t.next_window = t.frame;
      } else {
        t.knob = 0;
        t.next_window = t.frame + kTuneRoundGap;
      }
    }
  }
  for (uint32_t k = 0; k < K; k++) t.use[k] = t.forced[k] ? t.forced[k] : t.pick[k];
  t.frame++;
  if (t.fixed) return VXP_OK;
  if (t.window_left == 0 && t.frame - 1 >= t.next_window) {
    while (t.knob < K && t.forced[t.knob]) t.knob++;
    if (t.knob == K) {  // every knob forced
      t.knob = 0;
      return VXP_OK;
    }
    t.window_left = 3 * (uint32_t)t.cand[t.knob].size();
    t.ms.assign(t.cand[t.knob].size(), 0.0);
    t.n.assign(t.cand[t.knob].size(), 0u);
  }
  if (t.window_left) {
    if (!t.a) {
      HIP_TRY(hipEventCreate(&t.a));
      HIP_TRY(hipEventCreate(&t.b));
    }
    t.pending_idx = t.window_left % (uint32_t)t.cand[t.knob].size();
    t.use[t.knob] = t.cand[t.knob][t.pending_idx];
    t.pending = true;
    t.marked_a = t.marked_b = false;
    t.window_left--;
  }
  return VXP_OK;
}

// Brackets the phase that the window's knob shapes (events on the main stream).
int tuneMark(vxp_system* s, vxp_system::Phase ph, bool begin) {
  auto& t = s->tune;
  if (!t.pending || knobPhase(t.knob) != ph) return VXP_OK;
  if (begin) {
    HIP_TRY(hipEventRecord(t.a, s->stream));
    t.marked_a = true;
  } else if (t.marked_a) {
    HIP_TRY(hipEventRecord(t.b, s->stream));
    t.marked_b = true;
  }
  return VXP_OK;
}

// Orders the main stream after the side stream's last work.
int joinSide(vxp_system* s) {
  if (s->side_pending) {
    HIP_TRY(hipStreamWaitEvent(s->stream, s->ev_join, 0));
    s->side_pending = false;
  }
  return VXP_OK;
}

static int stepAllImpl(vxp_system* s);
static bool recordable(const vxp_system* s);
static int stepRecorded(vxp_system* s);
int stepAll(vxp_system* s) {
  RC(ensureDevice(s));
  RC(tuneStep(s));
  const int rc = recordable(s) ? stepRecorded(s) : stepAllImpl(s);
  const int jr = joinSide(s);  // a step that ended early leaves the join undone
  return rc != VXP_OK ? rc : jr;
}

static int stepAllImpl(vxp_system* s) {
  hipStream_t st = s->stream;
  DevCfg cfg = devCfg(s);
  cfg.pair_color_max = s->tune.use[vxp_system::kPairColorMax];
  DevBodies B = devBodies(s);
  DevColliders C = devColliders(s);
  const uint32_t nw = s->nw;
  const uint32_t ns = (uint32_t)slots(s);
  WorldState* ws = s->d_ws.as<WorldState>();
  vxp_stats* stats = s->d_stats.as<vxp_stats>();

  // ContactManager.beginFrame: the current manifolds become the previous ones
  s->cur ^= 1;
  DevManifold* prev = s->d_ml[s->cur ^ 1].as<DevManifold>();
  hipEvent_t pe = profBegin(s, VXP_K_STEP_BEGIN);
  HIP_TRY(s->d_step_ctr.ensure(kStepCounters * 4));
  launch_step_begin(st, cfg, B, ws, stats, nw, s->d_step_ctr.as<uint32_t>());
  profEnd(s, VXP_K_STEP_BEGIN, pe);
  const bool tree = treeMode(s);
  uint8_t* moved = nullptr;
  if (tree) {
    HIP_TRY(s->d_t_moved.ensure(ns));
    moved = s->d_t_moved.as<uint8_t>();
  }
  pe = profBegin(s, VXP_K_FAT_AABB);
  launch_fat_aabb(st, cfg, B, C, ws, ns, moved);
  profEnd(s, VXP_K_FAT_AABB, pe);

  uint32_t wbits = 0;
  while ((1u << wbits) < nw) wbits++;
  uint32_t id_bits = 1;  // body ids are below id_cap
  while ((1u << id_bits) < s->id_cap) id_bits++;
  uint32_t npairs = 0;
  uint32_t max_world_pairs = 0xFFFFFFFFu;  // the most pairs one world emitted (untraced orders)
  HIP_TRY(s->d_maxw.ensure(4));
  RC(tuneMark(s, vxp_system::kPhPairs, true));
  if (tree) {
    // ---- pair set in tree order: re-insert the moved leaves, then every
    // dynamic leaf queries both trees (count, scan, write) --------------------
    const TreeSet T = treeSet(s);
    const uint32_t nq = nw * 2 * s->t_nc;
    HIP_TRY(s->d_t_cnt.ensure((size_t)(nq + 1) * 4));
    HIP_TRY(s->d_t_off.ensure((size_t)(nq + 1) * 4));
    pe = profBegin(s, VXP_K_PAIRS);
    launch_tree_update(st, cfg, B, ws, T, moved, nw, s->lds_limit);
    launch_tree_pairs(st, ws, T, nw, s->d_t_cnt.as<uint32_t>(), nullptr, nullptr, nullptr, id_bits);
    HIP_TRY(hipMemsetAsync(s->d_t_cnt.as<uint32_t>() + nq, 0, 4, st));
    {
      size_t cb = exclusive_scan_u32_bytes(nq + 1);
      HIP_TRY(s->d_sort_tmp.ensure(cb));
      HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, cb, s->d_t_cnt.as<uint32_t>(),
                                 s->d_t_off.as<uint32_t>(), nq + 1, st));
    }
    RC(readSizes(s, {s->d_t_off.as<uint32_t>() + nq}, &npairs));
    HIP_TRY(s->d_pairs_raw.ensure((size_t)std::max(npairs, 1u) * 8));
    HIP_TRY(s->d_t_pos.ensure((size_t)std::max(npairs, 1u) * 4));
    if (npairs)
      launch_tree_pairs(st, ws, T, nw, nullptr, s->d_t_off.as<uint32_t>(),
                        s->d_pairs_raw.as<uint64_t>(), s->d_t_pos.as<uint32_t>(), id_bits);
    profEnd(s, VXP_K_PAIRS, pe);
    s->t_dev_newer = true;
  } else if (broadPhase(s) == VXP_BROAD_LBVH) {
    // ---- pair set: LBVH, one stackless query per leaf (count, scan, write) --
    HIP_TRY(s->d_skeys.ensure((size_t)ns * 8));
    HIP_TRY(s->d_svals.ensure((size_t)ns * 4));
    HIP_TRY(s->d_skeys2.ensure((size_t)ns * 8));
    HIP_TRY(s->d_svals2.ensure((size_t)ns * 4));
    HIP_TRY(s->d_slo.ensure((size_t)ns * 16));
    HIP_TRY(s->d_shi.ensure((size_t)ns * 16));
    HIP_TRY(s->d_lb_ilo.ensure((size_t)ns * 16));
    HIP_TRY(s->d_lb_ihi.ensure((size_t)ns * 16));
    HIP_TRY(s->d_lb_code.ensure((size_t)ns * 4));
    HIP_TRY(s->d_lb_child.ensure((size_t)ns * 8));
    HIP_TRY(s->d_lb_range.ensure((size_t)ns * 8));
    HIP_TRY(s->d_lb_iparent.ensure((size_t)ns * 4));
    HIP_TRY(s->d_lb_lparent.ensure((size_t)ns * 4));
    HIP_TRY(s->d_lb_flag.ensure((size_t)ns * 4));
    HIP_TRY(s->d_wcnt.ensure((size_t)(ns + 1) * 4));
    HIP_TRY(s->d_woff.ensure((size_t)(ns + 1) * 4));
    const LbvhBufs L{s->d_slo.as<float4>(),        s->d_shi.as<float4>(),
                     s->d_lb_ilo.as<float4>(),     s->d_lb_ihi.as<float4>(),
                     s->d_lb_code.as<uint32_t>(),  s->d_lb_child.as<uint2>(),
                     s->d_lb_range.as<uint2>(),    s->d_lb_iparent.as<uint32_t>(),
                     s->d_lb_lparent.as<uint32_t>(), s->d_lb_flag.as<uint32_t>()};
    pe = profBegin(s, VXP_K_PAIRS);
    if (!launch_lbvh_sort(st, cfg, B, ws, s->d_svals2.as<uint32_t>(), nw)) {
      launch_lbvh_keys(st, cfg, B, ws, s->d_skeys.as<uint64_t>(), s->d_svals.as<uint32_t>(), ns, nw);
      size_t tb = sort_pairs_u64_u32_bytes(ns, 32 + wbits);
      HIP_TRY(s->d_sort_tmp.ensure(tb));
      HIP_TRY(sort_pairs_u64_u32(s->d_sort_tmp.p, tb, s->d_skeys.as<uint64_t>(),
                                 s->d_skeys2.as<uint64_t>(), s->d_svals.as<uint32_t>(),
                                 s->d_svals2.as<uint32_t>(), ns, 32 + wbits, st));
    }
    launch_lbvh_build(st, cfg, B, ws, s->d_svals2.as<uint32_t>(), ns, L);
    const bool narrow = 2 * id_bits + wbits <= 32;
    launch_lbvh_pairs(st, cfg, ws, ns, L, s->d_wcnt.as<uint32_t>(), nullptr, nullptr, id_bits,
                      narrow);
    {
      size_t cb = exclusive_scan_u32_bytes(ns + 1);
      HIP_TRY(s->d_sort_tmp.ensure(cb));
      HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, cb, s->d_wcnt.as<uint32_t>(),
                                 s->d_woff.as<uint32_t>(), ns + 1, st));
    }
    launch_world_pairs(st, ws, nw, s->d_woff.as<uint32_t>(), s->MB, s->d_maxw.as<uint32_t>());
    {
      uint32_t sz[2];
      RC(readSizes(s, {s->d_woff.as<uint32_t>() + ns, s->d_maxw.as<uint32_t>()}, sz));
      npairs = sz[0];
      max_world_pairs = sz[1];
    }
    HIP_TRY(s->d_pairs_raw.ensure((size_t)std::max(npairs, 1u) * 8));
    if (npairs)
      launch_lbvh_pairs(st, cfg, ws, ns, L, nullptr, s->d_woff.as<uint32_t>(),
                        s->d_pairs_raw.as<uint64_t>(), id_bits, narrow);
    profEnd(s, VXP_K_PAIRS, pe);
  } else {
  // ---- pair set: sort-and-sweep --------------------------------------------
  HIP_TRY(s->d_skeys.ensure((size_t)ns * 8));
  HIP_TRY(s->d_svals.ensure((size_t)ns * 4));
  HIP_TRY(s->d_skeys2.ensure((size_t)ns * 8));
  HIP_TRY(s->d_svals2.ensure((size_t)ns * 4));
  HIP_TRY(s->d_slo.ensure((size_t)ns * 16));
  HIP_TRY(s->d_shi.ensure((size_t)ns * 16));
  pe = profBegin(s, VXP_K_PAIRS);
  if (!launch_sweep_sort(st, cfg, B, ws, s->d_svals2.as<uint32_t>(), nw)) {
    // worlds above the block sort's size: device-wide sort of (world << 32 | coordinate)
    // keys, which are zero above bit 32 + log2(worlds)
    launch_sweep_keys(st, cfg, B, ws, s->d_skeys.as<uint64_t>(), s->d_svals.as<uint32_t>(), ns, nw);
    size_t tb = sort_pairs_u64_u32_bytes(ns, 32 + wbits);
    HIP_TRY(s->d_sort_tmp.ensure(tb));
    HIP_TRY(sort_pairs_u64_u32(s->d_sort_tmp.p, tb, s->d_skeys.as<uint64_t>(),
                               s->d_skeys2.as<uint64_t>(), s->d_svals.as<uint32_t>(),
                               s->d_svals2.as<uint32_t>(), ns, 32 + wbits, st));
  }
  HIP_TRY(s->d_send.ensure((size_t)ns * 4));
  HIP_TRY(s->d_wcnt.ensure((size_t)(ns + 1) * 4));
  HIP_TRY(s->d_woff.ensure((size_t)(ns + 1) * 4));
  launch_sweep_count(st, cfg, B, ws, s->d_svals2.as<uint32_t>(), ns, s->d_slo.as<float4>(),
                     s->d_shi.as<float4>(), s->d_send.as<uint32_t>(), s->d_wcnt.as<uint32_t>());
  {
    size_t cb = exclusive_scan_u32_bytes(ns + 1);
    HIP_TRY(s->d_sort_tmp.ensure(cb));
    HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, cb, s->d_wcnt.as<uint32_t>(),
                               s->d_woff.as<uint32_t>(), ns + 1, st));
  }
  // two passes over the same work-item partition: count per wave, scan, write;
  // packed keys (world, a, b) are zero above bit 2 * id_bits + log2(worlds)
  // and written as 32-bit words when that fits
  const bool narrow = 2 * id_bits + wbits <= 32;
  const uint32_t swv = sweep_waves();
  HIP_TRY(s->d_swc.ensure((size_t)(swv + 1) * 4));
  HIP_TRY(s->d_swo.ensure((size_t)(swv + 1) * 4));
  // per-world pair counts (for the per-world key sort) only with several
  // worlds: one world's count is the total, and 8192 waves adding to one
  // counter would serialize
  launch_sweep_pairs(st, ws, s->d_slo.as<float4>(), s->d_shi.as<float4>(), ns,
                     s->d_send.as<uint32_t>(), s->d_woff.as<uint32_t>(), s->d_swc.as<uint32_t>(),
                     nullptr, nullptr, true, id_bits, narrow, nw > 1);
  {
    size_t cb = exclusive_scan_u32_bytes(swv + 1);
    HIP_TRY(s->d_sort_tmp.ensure(cb));
    HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, cb, s->d_swc.as<uint32_t>(),
                               s->d_swo.as<uint32_t>(), swv + 1, st));
  }
  if (nw > 1) {
    launch_world_pairs(st, ws, nw, nullptr, s->MB, s->d_maxw.as<uint32_t>());
    uint32_t sz[2];
    RC(readSizes(s, {s->d_swo.as<uint32_t>() + swv, s->d_maxw.as<uint32_t>()}, sz));
    npairs = sz[0];
    max_world_pairs = sz[1];
  } else {
    RC(readSizes(s, {s->d_swo.as<uint32_t>() + swv}, &npairs));
    max_world_pairs = npairs;
  }
  HIP_TRY(s->d_pairs_raw.ensure((size_t)std::max(npairs, 1u) * 8));
  if (npairs)
    launch_sweep_pairs(st, ws, s->d_slo.as<float4>(), s->d_shi.as<float4>(), ns,
                       s->d_send.as<uint32_t>(), s->d_woff.as<uint32_t>(), nullptr,
                       s->d_swo.as<uint32_t>(), s->d_pairs_raw.as<uint64_t>(), false, id_bits,
                       narrow, false);
  profEnd(s, VXP_K_PAIRS, pe);
  }
  s->n_pairs = npairs;
  HIP_TRY(s->d_pairs.ensure((size_t)std::max(npairs, 1u) * 8));
  const uint32_t* perm = nullptr;  // key position -> tree position
  pe = profBegin(s, VXP_K_PAIRS);
  bool seg = false;
  if (npairs) {
    // packed pair keys (world, a, b) are zero above bit 2 * id_bits + log2(worlds)
    const uint32_t kbits = 2 * id_bits + wbits;
    if (tree) {
      HIP_TRY(s->d_perm.ensure((size_t)npairs * 4));
      HIP_TRY(s->d_pairs_tree.ensure((size_t)npairs * 8));
      HIP_TRY(hipMemcpyAsync(s->d_pairs_tree.p, s->d_pairs_raw.p, (size_t)npairs * 8,
                             hipMemcpyDeviceToDevice, st));
      launch_expand_keys(st, s->d_pairs_tree.as<uint64_t>(), npairs, id_bits);
      size_t kb = sort_pairs_u64_u32_bytes(npairs, kbits);
      HIP_TRY(s->d_sort_tmp.ensure(kb));
      HIP_TRY(sort_pairs_u64_u32(s->d_sort_tmp.p, kb, s->d_pairs_raw.as<uint64_t>(),
                                 s->d_pairs.as<uint64_t>(), s->d_t_pos.as<uint32_t>(),
                                 s->d_perm.as<uint32_t>(), npairs, kbits, st));
      perm = s->d_perm.as<uint32_t>();
    } else if (launch_segment_sort(st, s->d_pairs_raw.p, kbits <= 32, npairs, nw, id_bits,
                                   max_world_pairs, s->d_pairs.as<uint64_t>())) {
      seg = true;  // every world's keys sorted in one workgroup, written canonical
    } else if (kbits <= 32) {
      // 32-bit keys: sorted from the first half of the raw buffer into its second
      uint32_t* raw = s->d_pairs_raw.as<uint32_t>();
      size_t kb = sort_keys_u32_bytes(npairs, kbits);
      HIP_TRY(s->d_sort_tmp.ensure(kb));
      HIP_TRY(sort_keys_u32(s->d_sort_tmp.p, kb, raw, raw + npairs, npairs, kbits, st));
      launch_expand_keys32(st, raw + npairs, s->d_pairs.as<uint64_t>(), npairs, id_bits);
    } else {
      size_t kb = sort_keys_u64_bytes(npairs, kbits);
      HIP_TRY(s->d_sort_tmp.ensure(kb));
      HIP_TRY(sort_keys_u64(s->d_sort_tmp.p, kb, s->d_pairs_raw.as<uint64_t>(),
                            s->d_pairs.as<uint64_t>(), npairs, kbits, st));
    }
    if (tree || (!seg && kbits > 32))
      launch_expand_keys(st, s->d_pairs.as<uint64_t>(), npairs, id_bits);
  }
  profEnd(s, VXP_K_PAIRS, pe);
  RC(tuneMark(s, vxp_system::kPhPairs, false));
  const uint64_t* keys = s->d_pairs.as<uint64_t>();

  // ---- narrow phase ----------------------------------------------------------
  uint32_t np1 = std::max(npairs, 1u);
  HIP_TRY(s->d_pr.ensure((size_t)np1 * sizeof(PairRec)));
  HIP_TRY(s->d_counts.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_wi_off.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_produce.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_pvchk.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_rstat.ensure((size_t)np1 * sizeof(RStat)));
  HIP_TRY(s->d_clist.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_wlist.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_coop.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_cflag.ensure((size_t)(npairs + 1) * 4));
  HIP_TRY(s->d_cidx.ensure((size_t)(npairs + 1) * 4));
  HIP_TRY(s->d_pidx.ensure((size_t)np1 * 4));
  HIP_TRY(s->d_staging.ensure((size_t)np1 * sizeof(DevManifold)));
  HIP_TRY(s->d_cstage.ensure((size_t)np1 * sizeof(CacheVal)));
  PairRec* pr = s->d_pr.as<PairRec>();
  pe = profBegin(s, VXP_K_PAIR_PROCESS);
  // the filtered-pair cap (src/system.zig:594-616), counted in the narrow
  // phase's order over the pairs that are not both asleep
  const uint8_t* drop = nullptr;
  const bool threaded = s->cfg.enable_threading;
  if ((s->cfg.pair_cap || threaded) && npairs) {
    HIP_TRY(s->d_awake.ensure((size_t)(npairs + 1) * 4));
    HIP_TRY(s->d_arank.ensure((size_t)(npairs + 1) * 4));
    launch_pair_awake(st, cfg, B, keys, perm, npairs, s->d_awake.as<uint32_t>());
    size_t cb = exclusive_scan_u32_bytes(npairs + 1);
    HIP_TRY(s->d_sort_tmp.ensure(cb));
    HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, cb, s->d_awake.as<uint32_t>(),
                               s->d_arank.as<uint32_t>(), npairs + 1, st));
    if (s->cfg.pair_cap) {
      HIP_TRY(s->d_drop.ensure((size_t)npairs));
      launch_pair_drop(st, keys, perm, npairs, s->d_awake.as<uint32_t>(),
                       s->d_arank.as<uint32_t>(), s->cfg.pair_cap, s->d_drop.as<uint8_t>());
      drop = s->d_drop.as<uint8_t>();
    }
  }
  if (threaded)  // the thread-pool narrow phase's body list and branch, per world
    launch_thread_mode(st, cfg, B, ws, keys, npairs, npairs ? s->d_arank.as<uint32_t>() : nullptr,
                       s->cfg.pair_cap, nw);
  launch_pair_process(st, cfg, B, C, ws, stats, keys, npairs, pr, prev, drop,
                      s->d_cflag.as<uint32_t>(), s->d_produce.as<uint32_t>(),
                      s->d_pvchk.as<uint32_t>(), s->d_wlist.as<uint32_t>(),
                      s->d_step_ctr.as<uint32_t>(), s->d_counts.as<uint32_t>(),
                      false);  // the eager step reads the work items after the scan
  profEnd(s, VXP_K_PAIR_PROCESS, pe);
  pe = profBegin(s, VXP_K_CACHE_REPLAY);
  // cache-eligible flags in the narrow phase's order, and their positions
  uint32_t* cflag_r = s->d_cflag.as<uint32_t>();
  uint32_t* cidx_r = s->d_cidx.as<uint32_t>();
  const uint32_t* order = nullptr;
  if (tree && npairs) {
    HIP_TRY(s->d_inv.ensure((size_t)npairs * 4));
    HIP_TRY(s->d_cflag_r.ensure((size_t)(npairs + 1) * 4));
    HIP_TRY(s->d_cidx_r.ensure((size_t)(npairs + 1) * 4));
    launch_invert(st, perm, npairs, s->d_inv.as<uint32_t>());
    launch_gather_flags(st, s->d_cflag.as<uint32_t>(), s->d_inv.as<uint32_t>(), npairs,
                        s->d_cflag_r.as<uint32_t>());
    cflag_r = s->d_cflag_r.as<uint32_t>();
    cidx_r = s->d_cidx_r.as<uint32_t>();
    order = s->d_inv.as<uint32_t>();
  }
  if (npairs) {
    size_t cb = exclusive_scan_u32_bytes(npairs + 1);
    HIP_TRY(s->d_sort_tmp.ensure(cb));
    HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, cb, cflag_r, cidx_r, npairs + 1, st));
  }
  launch_cache_replay(st, cfg, ws, stats, pr, s->d_cache_keys.as<CacheKeyRec>(), cflag_r, cidx_r,
                      order, s->d_clist.as<uint32_t>(), npairs, nw, s->d_counts.as<uint32_t>());
  profEnd(s, VXP_K_CACHE_REPLAY, pe);
  uint32_t total_wi = 0;
  if (npairs) {
    size_t sb = exclusive_scan_u32_bytes(npairs);
    HIP_TRY(s->d_sort_tmp.ensure(sb));
    HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, sb, s->d_counts.as<uint32_t>(), s->d_wi_off.as<uint32_t>(),
                               npairs, st));
    uint32_t last[2];
    RC(readSizes(s, {s->d_wi_off.as<uint32_t>() + npairs - 1, s->d_counts.as<uint32_t>() + npairs - 1},
                 last));
    total_wi = last[0] + last[1];
  }
  HIP_TRY(s->d_wi_cbase.ensure((size_t)(total_wi + 1) * 4));
  HIP_TRY(s->d_wi_ccnt.ensure((size_t)(total_wi + 1) * 4));
  HIP_TRY(s->d_lane_n.ensure((size_t)std::max(total_wi, 1u) * 64));
  HIP_TRY(s->d_wi_pair.ensure((size_t)std::max(total_wi, 1u) * 4));
  if (total_wi >= (1u << 26)) {  // lane-list entries are work item * 64 + lane in 32 bits
    g_err = "probe: more than 2^26 work items in one step";
    return VXP_ERR_CAPACITY;
  }
  HIP_TRY(s->d_wi_list.ensure((size_t)std::max(total_wi, 1u) * 64 * 4));
  // pass 1 counts and lists the work items with candidates, the scan places
  // every work item's candidates, pass 2 writes those of the listed ones
  pe = profBegin(s, VXP_K_PROBE);
  RC(tuneMark(s, vxp_system::kPhProbe, true));
  const ProbeShape shape{s->tune.use[vxp_system::kProbeThreads],
                         s->tune.use[vxp_system::kProbeXcdRun]};
  uint32_t ncand = 0, nlist = 0;
  if (total_wi) {
    launch_wi_fill(st, s->d_counts.as<uint32_t>(), s->d_wi_off.as<uint32_t>(), npairs,
                   s->d_wi_pair.as<uint32_t>(), s->d_wi_ccnt.as<uint32_t>());
    launch_probe_count(st, cfg, B, C, pr, s->d_wi_pair.as<uint32_t>(),
                       s->d_wi_off.as<uint32_t>(), total_wi, s->d_wi_ccnt.as<uint32_t>(),
                       s->d_pvchk.as<uint32_t>(), s->d_lane_n.as<uint8_t>(),
                       s->d_wi_list.as<uint32_t>(), s->d_step_ctr.as<uint32_t>() + kCtrProbeItems,
                       shape);
    size_t cb = exclusive_scan_u32_bytes(total_wi + 1);
    HIP_TRY(s->d_sort_tmp.ensure(cb));
    HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, cb, s->d_wi_ccnt.as<uint32_t>(),
                               s->d_wi_cbase.as<uint32_t>(), total_wi + 1, st));
    uint32_t sz[2];
    RC(readSizes(s, {s->d_wi_cbase.as<uint32_t>() + total_wi,
                     s->d_step_ctr.as<uint32_t>() + kCtrProbeItems}, sz));
    ncand = sz[0];
    nlist = sz[1];
  }
  HIP_TRY(s->d_cand.ensure((size_t)std::max(ncand, 1u) * sizeof(Cand)));
  if (ncand)
    launch_probe_write(st, cfg, B, C, pr, s->d_wi_pair.as<uint32_t>(),
                       s->d_wi_off.as<uint32_t>(), s->d_wi_list.as<uint32_t>(), nlist,
                       s->d_wi_cbase.as<uint32_t>(), s->d_lane_n.as<uint8_t>(),
                       s->d_cand.as<Cand>());
  RC(tuneMark(s, vxp_system::kPhProbe, false));
  profEnd(s, VXP_K_PROBE, pe);
  s->last_wi = total_wi;
  s->last_cand = ncand;
  pe = profBegin(s, VXP_K_REDUCE);
  launch_reduce(st, cfg, B, C, ws, stats, pr, keys, npairs, s->d_wlist.as<uint32_t>(),
                s->d_step_ctr.as<uint32_t>(), s->d_wi_off.as<uint32_t>(),
                s->d_wi_cbase.as<uint32_t>(), s->d_wi_ccnt.as<uint32_t>(), s->d_pvchk.as<uint32_t>(),
                s->d_cand.as<Cand>(), prev, s->d_cache_keys.as<CacheKeyRec>(),
                s->d_cache_vals.as<CacheVal>(), s->d_cstage.as<CacheVal>(),
                s->d_staging.as<DevManifold>(), s->d_produce.as<uint32_t>(),
                s->d_rstat.as<RStat>(), s->d_coop.as<uint32_t>());
  profEnd(s, VXP_K_REDUCE, pe);
  pe = profBegin(s, VXP_K_COMPACT);
  if (npairs) {
    size_t sb = exclusive_scan_u32_bytes(npairs);
    HIP_TRY(s->d_sort_tmp.ensure(sb));
    HIP_TRY(exclusive_scan_u32(s->d_sort_tmp.p, sb, s->d_produce.as<uint32_t>(), s->d_pidx.as<uint32_t>(),
                               npairs, st));
  }
  HIP_TRY(s->d_total.ensure(64));
  launch_world_segments(st, ws, nw, s->d_total.as<uint32_t>());
  profEnd(s, VXP_K_COMPACT, pe);
  uint32_t sizes[2] = {0, 0};
  RC(readSizes(s, {s->d_total.as<uint32_t>(), s->d_total.as<uint32_t>() + 1}, sizes));
  const uint32_t nman = sizes[0], max_man = sizes[1];
  s->n_manifolds = nman;
  s->last_max_man = max_man;
  uint32_t nm1 = std::max(nman, 1u);
  HIP_TRY(s->d_ml[s->cur].ensure((size_t)nm1 * sizeof(DevManifold)));
  DevManifold* ml = s->d_ml[s->cur].as<DevManifold>();
  pe = profBegin(s, VXP_K_COMPACT);
  launch_scatter(st, ws, pr, npairs, s->d_produce.as<uint32_t>(), s->d_pidx.as<uint32_t>(),
                 s->d_staging.as<DevManifold>(), ml);
  launch_copy_skipped(st, ws, prev, ml, nw);
  profEnd(s, VXP_K_COMPACT, pe);

  // ---- islands / colouring artefacts ------------------------------------------
  if (s->cfg.enable_coloring) {
    HIP_TRY(s->d_ref_masks.ensure((size_t)nw * s->id_cap * 32));
    HIP_TRY(s->d_ref_colors.ensure((size_t)nm1));
    launch_ref_coloring(st, cfg, ws, ml, s->d_ref_masks.as<uint64_t>(), s->d_ref_colors.as<uint8_t>(), nw);
  }

  // ---- solver ------------------------------------------------------------------
  HIP_TRY(s->s_pos.ensure((size_t)ns * 16));
  HIP_TRY(s->s_rot.ensure((size_t)ns * 16));
  HIP_TRY(s->s_lv.ensure((size_t)ns * 16));
  HIP_TRY(s->s_av.ensure((size_t)ns * 16));
  HIP_TRY(s->s_ii.ensure((size_t)ns * 36));
  HIP_TRY(s->s_adj_cnt.ensure((size_t)ns * 4));
  HIP_TRY(s->s_adj_off.ensure((size_t)ns * 4));
  HIP_TRY(s->s_adj.ensure((size_t)nm1 * 8));
  HIP_TRY(s->s_color.ensure((size_t)nm1 * 4));
  HIP_TRY(s->s_order.ensure((size_t)nm1 * 4));
  HIP_TRY(s->s_prio.ensure((size_t)nm1 * 4));
  HIP_TRY(s->s_rec.ensure((size_t)nm1 * 9 * sizeof(float4)));
  HIP_TRY(s->s_pts.ensure((size_t)nm1 * 4 * 8 * sizeof(float4)));
  HIP_TRY(s->s_tstamp.ensure((size_t)nw * 16 * 8));
  HIP_TRY(s->s_upar.ensure((size_t)ns * 4));
  HIP_TRY(s->s_ucnt.ensure((size_t)ns * 4));
  HIP_TRY(s->s_uoff.ensure((size_t)ns * 4));
  HIP_TRY(s->s_ubody.ensure((size_t)ns * 4));
  HIP_TRY(s->s_uorder.ensure((size_t)nm1 * 4));
  HIP_TRY(s->s_units.ensure((size_t)nw * kUnitsPerWorld * sizeof(uint4)));
  HIP_TRY(s->s_uqueue.ensure((size_t)nw * kUnitsPerWorld * 4));
  HIP_TRY(s->s_uqueue2.ensure((size_t)nw * kUnitsPerWorld * 4));
  HIP_TRY(s->s_pcnt.ensure((size_t)nm1 * 4));
  HIP_TRY(s->s_fq.ensure((size_t)nm1 * 4));
  HIP_TRY(s->s_cend.ensure((size_t)nm1 * 16));
  SolverScratch S;
  S.ws_pos = s->s_pos.as<float4>();
  S.ws_rot = s->s_rot.as<float4>();
  S.ws_lv = s->s_lv.as<float4>();
  S.ws_av = s->s_av.as<float4>();
  S.ws_ii = s->s_ii.as<float>();
  S.adj_cnt = s->s_adj_cnt.as<uint32_t>();
  S.adj_off = s->s_adj_off.as<uint32_t>();
  S.adj = s->s_adj.as<uint32_t>();
  S.color = s->s_color.as<int32_t>();
  S.order = s->s_order.as<uint32_t>();
  S.rec = s->s_rec.as<float4>();
  S.pts = s->s_pts.as<float4>();
  S.tstamp = s->prof == VXP_PROFILE_ALL ? s->s_tstamp.as<uint64_t>() : nullptr;
  S.upar = s->s_upar.as<uint32_t>();
  S.ucnt = s->s_ucnt.as<uint32_t>();
  S.uoff = s->s_uoff.as<uint32_t>();
  S.ubody = s->s_ubody.as<uint32_t>();
  S.uorder = s->s_uorder.as<uint32_t>();
  S.units = s->s_units.as<uint4>();
  S.uqueue = s->s_uqueue.as<uint32_t>();
  S.pcnt = s->s_pcnt.as<uint32_t>();
  S.fq = s->s_fq.as<uint32_t>();
  S.cend = s->s_cend.as<uint4>();
  // The island labels (manifolds and pre-solve body state in, island arrays
  // out) and the body working state (body state in, working state out) run on
  // the side stream beside the solver's prep, which reads neither output; the
  // solve units join them.  Kept in line when every phase is being timed.
  if (s->cfg.enable_islands) RC(ensureIslands(s));
  hipStream_t side = st;
  if (s->prof != VXP_PROFILE_ALL) {
    HIP_TRY(hipEventRecord(s->ev_fork, st));
    HIP_TRY(hipStreamWaitEvent(s->side, s->ev_fork, 0));
    side = s->side;
  }
  if (s->cfg.enable_islands) {
    pe = profBegin(s, VXP_K_ISLANDS);
    launch_islands(side, cfg, B, ws, stats, ml, prev, islandArrays(s), nw, max_man, s->lds_limit);
    profEnd(s, VXP_K_ISLANDS, pe);
  }
  launch_body_state(side, cfg, B, ws, S, ns);
  if (side != st) {
    HIP_TRY(hipEventRecord(s->ev_join, s->side));
    s->side_pending = true;
  }
  uint32_t* ctr = s->d_step_ctr.as<uint32_t>();
  RC(tuneMark(s, vxp_system::kPhSolve, true));
  pe = profBegin(s, VXP_K_SOLVE);
  launch_solve_prep(st, cfg, B, ws, stats, ml, S, s->s_prio.as<uint32_t>(), ctr, nw, max_man,
                    s->lds_limit, s->d_total.as<uint32_t>());
  profEnd(s, VXP_K_SOLVE, pe);
  RC(joinSide(s));  // the units read the working state; the writeback overwrites what the labels read
  pe = profBegin(s, VXP_K_SOLVE_UNITS);
  launch_unit_order(st, S, ctr, s->s_uqueue2.as<uint32_t>());
  // an earlier launch's failure is reported here, so that only the solve
  // units' own refusal below selects the variant without LDS staging
  HIP_TRY(hipGetLastError());
  launch_solve_units(st, cfg, B, ws, ml, S, ctr, nw, s->n_cu, s->lds_limit, unitThreads(s));
  if (hipGetLastError() != hipSuccess) {  // LDS request refused: global variant
    s->lds_limit = 0;
    launch_solve_units(st, cfg, B, ws, ml, S, ctr, nw, s->n_cu, 0, unitThreads(s));
  }
  profEnd(s, VXP_K_SOLVE_UNITS, pe);
  pe = profBegin(s, VXP_K_SOLVE_TAIL);
  launch_solve_tail(st, cfg, B, ws, stats, S, ctr, nw, s->n_cu);
  profEnd(s, VXP_K_SOLVE_TAIL, pe);
  RC(tuneMark(s, vxp_system::kPhSolve, false));
  HIP_TRY(hipGetLastError());
  s->dev_newer = true;
  return VXP_OK;
}

// ---- recorded step ------------------------------------------------------------
// vxp_graph_enable: the step of the default pipeline (sort-and-sweep pairs in
// key order, no filtered-pair cap, no thread-pool narrow phase, no per-phase
// timing, no tuning window in progress) is recorded as three HIP graphs per
// manifold-buffer parity and replayed.  Every launch is sized for a capacity
// and reads its item count on the device, so a replay needs no size read
// between launches; the host reads the counts twice per step, after segment 1
// (pairs, pairs of one world, the probe work items if every cache-eligible
// pair missed: the three counts the capacities bound) and after segment 2
// (contact candidates).  Segment 1 changes no state that a redo could not
// restore (world states and stats are copied aside first; the fat boxes are a
// function of the poses), so a step past a capacity is rolled back and
// replayed with larger graphs; past the candidate capacity, only segment 3 is
// recorded again.  The kernels and their order are the eager step's, so the
// results are the same bits.

// Every device buffer of a system (released together; their addresses are
// part of what a recorded step bakes in).
static std::vector<DevBuf*> allBufs(vxp_system* s) {
  return {&s->d_hdr, &s->d_dir, &s->d_vox, &s->d_mask, &s->b_pos, &s->b_rot, &s->b_lv, &s->b_av,
          &s->b_fat_lo, &s->b_fat_hi, &s->b_inv_mass, &s->b_iil, &s->b_fric, &s->b_rest,
          &s->b_ldamp, &s->b_adamp, &s->b_sleep_timer, &s->b_type, &s->b_sleeping,
          &s->b_active, &s->b_collider, &s->b_id, &s->b_idmap, &s->d_ws, &s->d_stats,
          &s->d_cache_keys, &s->d_cache_vals, &s->d_skeys, &s->d_svals, &s->d_skeys2,
          &s->d_svals2, &s->d_sort_tmp, &s->d_slo, &s->d_shi, &s->d_send, &s->d_wcnt, &s->d_woff,
          &s->d_swc, &s->d_swo, &s->d_pairs_raw, &s->d_pairs, &s->d_pr, &s->d_counts, &s->d_wi_off,
          &s->d_wi_cbase, &s->d_wi_ccnt, &s->d_pvchk, &s->d_lane_n, &s->d_cand, &s->d_produce,
          &s->d_pidx, &s->d_staging, &s->d_cstage, &s->d_rstat, &s->d_clist, &s->d_cflag,
          &s->d_cidx, &s->d_wlist, &s->d_step_ctr, &s->d_total, &s->d_ml[0], &s->d_ml[1],
          &s->d_isl_parent, &s->d_isl_end, &s->d_isl_size, &s->d_ref_masks, &s->d_ref_colors,
          &s->s_pos, &s->s_rot, &s->s_lv, &s->s_av, &s->s_ii, &s->s_adj_cnt, &s->s_adj_off,
          &s->s_adj, &s->s_color, &s->s_order, &s->s_rec, &s->s_pts, &s->s_prio, &s->s_tstamp,
          &s->s_upar, &s->s_ucnt, &s->s_uoff, &s->s_ubody, &s->s_uorder, &s->s_units,
          &s->s_uqueue, &s->s_uqueue2, &s->s_pcnt, &s->s_fq, &s->s_cend, &s->d_cb_pos,
          &s->d_cb_vox, &s->d_cb_cnt, &s->d_cb_bounds, &s->d_cb_box, &s->d_cb_out,
          &s->d_cb_solid, &s->d_cb_nbr, &s->d_isl_timer, &s->d_isl_sleep, &s->d_t_nodes,
          &s->d_t_heads, &s->d_t_leaf, &s->d_t_moved, &s->d_t_cnt, &s->d_t_off, &s->d_t_pos,
          &s->d_perm, &s->d_inv, &s->d_pairs_tree, &s->d_awake, &s->d_arank, &s->d_drop,
          &s->d_cflag_r, &s->d_cidx_r, &s->d_wi_pair, &s->d_wi_list, &s->d_coop, &s->d_lb_ilo, &s->d_lb_ihi, &s->d_lb_code,
          &s->d_lb_child, &s->d_lb_range, &s->d_lb_iparent, &s->d_lb_lparent, &s->d_lb_flag,
          &s->d_maxw, &s->d_rsnap, &s->d_rdesc, &s->d_rsize};
}

static bool recordable(const vxp_system* s) {
  return s->rec_on && !treeMode(s) && broadPhase(s) == VXP_BROAD_SWEEP && !s->cfg.pair_cap &&
         !s->cfg.enable_threading && s->prof == VXP_PROFILE_OFF && !s->tune.pending &&
         s->lds_limit > 0 && s->n_cu > 0 && s->read_seq > 0;  // after an eager step
}

// Publication slots in the pinned words (h_read[0..2] serve readSizes).
constexpr uint32_t kPub1 = 16, kPub2 = 32, kPub3 = 48;

// What a recorded step bakes in: buffer addresses, shapes, knobs, capacities.
static uint64_t recSignature(vxp_system* s, int par) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  for (DevBuf* b : allBufs(s)) mix((uint64_t)(uintptr_t)b->p);
  mix(s->nw);
  mix(s->MB);
  mix(s->id_cap);
  mix((uint64_t)par);
  mix(s->lds_limit);
  mix(s->n_cu);
  for (uint32_t k = 0; k < vxp_system::kKnobs; k++) mix(s->tune.use[k]);
  mix(s->rc_pairs);
  mix(s->rc_wp);
  mix(s->rc_wi);
  mix(s->rc_cand);
  mix(s->rc_carry);
  mix(s->rc_mm);
  const unsigned char* c = (const unsigned char*)&s->cfg;
  for (size_t i = 0; i < sizeof(vxp_config); i++) mix(c[i]);
  return h;
}

static uint32_t idBits(const vxp_system* s) {
  uint32_t b = 1;
  while ((1u << b) < s->id_cap) b++;
  return b;
}
static uint32_t worldBits(const vxp_system* s) {
  uint32_t b = 0;
  while ((1u << b) < s->nw) b++;
  return b;
}

// Every buffer a recorded step touches, at its capacity (no allocation may
// happen while a step is being recorded).
static int recEnsure(vxp_system* s) {
  const size_t ns = slots(s), nw = s->nw;
  const uint32_t swv = sweep_waves(), wbits = worldBits(s), kbits = 2 * idBits(s) + wbits;
  const size_t P = s->rc_pairs, WI = s->rc_wi, CA = s->rc_cand;
  const size_t M = P + s->rc_carry;  // manifolds: produced (<= pairs) + carried over
  HIP_TRY(s->d_step_ctr.ensure(kStepCounters * 4));
  HIP_TRY(s->d_maxw.ensure(4));
  HIP_TRY(s->d_skeys.ensure(ns * 8));
  HIP_TRY(s->d_svals.ensure(ns * 4));
  HIP_TRY(s->d_skeys2.ensure(ns * 8));
  HIP_TRY(s->d_svals2.ensure(ns * 4));
  HIP_TRY(s->d_slo.ensure(ns * 16));
  HIP_TRY(s->d_shi.ensure(ns * 16));
  HIP_TRY(s->d_send.ensure(ns * 4));
  HIP_TRY(s->d_wcnt.ensure((ns + 1) * 4));
  HIP_TRY(s->d_woff.ensure((ns + 1) * 4));
  HIP_TRY(s->d_swc.ensure((size_t)(swv + 1) * 4));
  HIP_TRY(s->d_swo.ensure((size_t)(swv + 1) * 4));
  HIP_TRY(s->d_pairs_raw.ensure(P * 8));
  HIP_TRY(s->d_pairs.ensure(P * 8));
  HIP_TRY(s->d_pr.ensure(P * sizeof(PairRec)));
  HIP_TRY(s->d_counts.ensure(P * 4));
  HIP_TRY(s->d_wi_off.ensure(P * 4));
  HIP_TRY(s->d_produce.ensure(P * 4));
  HIP_TRY(s->d_pvchk.ensure(P * 4));
  HIP_TRY(s->d_rstat.ensure(P * sizeof(RStat)));
  HIP_TRY(s->d_clist.ensure(P * 4));
  HIP_TRY(s->d_wlist.ensure(P * 4));
  HIP_TRY(s->d_coop.ensure(P * 4));
  HIP_TRY(s->d_cflag.ensure((P + 1) * 4));
  HIP_TRY(s->d_cidx.ensure((P + 1) * 4));
  HIP_TRY(s->d_pidx.ensure(P * 4));
  HIP_TRY(s->d_staging.ensure(P * sizeof(DevManifold)));
  HIP_TRY(s->d_cstage.ensure(P * sizeof(CacheVal)));
  HIP_TRY(s->d_wi_cbase.ensure((WI + 1) * 4));
  HIP_TRY(s->d_wi_ccnt.ensure((WI + 1) * 4));
  HIP_TRY(s->d_lane_n.ensure(WI * 64));
  HIP_TRY(s->d_wi_pair.ensure(WI * 4));
  if (WI >= (1u << 26)) {  // lane-list entries are work item * 64 + lane in 32 bits
    g_err = "recorded step: more than 2^26 probe work items";
    return VXP_ERR_CAPACITY;
  }
  HIP_TRY(s->d_wi_list.ensure(WI * 64 * 4));
  HIP_TRY(s->d_cand.ensure(CA * sizeof(Cand)));
  HIP_TRY(s->d_total.ensure(64));
  HIP_TRY(s->d_ml[0].ensure(M * sizeof(DevManifold), true, s->stream));
  HIP_TRY(s->d_ml[1].ensure(M * sizeof(DevManifold), true, s->stream));
  if (s->cfg.enable_coloring) {
    HIP_TRY(s->d_ref_masks.ensure(nw * s->id_cap * 32));
    HIP_TRY(s->d_ref_colors.ensure(M));
  }
  if (s->cfg.enable_islands) RC(ensureIslands(s));
  HIP_TRY(s->s_pos.ensure(ns * 16));
  HIP_TRY(s->s_rot.ensure(ns * 16));
  HIP_TRY(s->s_lv.ensure(ns * 16));
  HIP_TRY(s->s_av.ensure(ns * 16));
  HIP_TRY(s->s_ii.ensure(ns * 36));
  HIP_TRY(s->s_adj_cnt.ensure(ns * 4));
  HIP_TRY(s->s_adj_off.ensure(ns * 4));
  HIP_TRY(s->s_adj.ensure(M * 8));
  HIP_TRY(s->s_color.ensure(M * 4));
  HIP_TRY(s->s_order.ensure(M * 4));
  HIP_TRY(s->s_prio.ensure(M * 4));
  HIP_TRY(s->s_rec.ensure(M * 9 * sizeof(float4)));
  HIP_TRY(s->s_pts.ensure(M * 4 * 8 * sizeof(float4)));
  HIP_TRY(s->s_tstamp.ensure(nw * 16 * 8));
  HIP_TRY(s->s_upar.ensure(ns * 4));
  HIP_TRY(s->s_ucnt.ensure(ns * 4));
  HIP_TRY(s->s_uoff.ensure(ns * 4));
  HIP_TRY(s->s_ubody.ensure(ns * 4));
  HIP_TRY(s->s_uorder.ensure(M * 4));
  HIP_TRY(s->s_units.ensure(nw * kUnitsPerWorld * sizeof(uint4)));
  HIP_TRY(s->s_uqueue.ensure(nw * kUnitsPerWorld * 4));
  HIP_TRY(s->s_uqueue2.ensure(nw * kUnitsPerWorld * 4));
  HIP_TRY(s->s_pcnt.ensure(M * 4));
  HIP_TRY(s->s_fq.ensure(M * 4));
  HIP_TRY(s->s_cend.ensure(M * 16));
  size_t tmp = std::max({exclusive_scan_u32_bytes((uint32_t)ns + 1),
                         exclusive_scan_u32_bytes(swv + 1),
                         exclusive_scan_u32_bytes((uint32_t)P + 1),
                         exclusive_scan_u32_bytes((uint32_t)WI + 1),
                         sort_pairs_u64_u32_bytes((uint32_t)ns, 32 + wbits),
                         kbits <= 32 ? sort_keys_u32_bytes((uint32_t)P, kbits)
                                     : sort_keys_u64_bytes((uint32_t)P, kbits)});
  HIP_TRY(s->d_sort_tmp.ensure(tmp));
  // snapshot of the world states and stats, and the copies to and from it
  const size_t wsb = nw * sizeof(WorldState), stb = nw * sizeof(vxp_stats);
  HIP_TRY(s->d_rsnap.ensure(wsb + stb));
  HIP_TRY(s->d_rsize.ensure(64));
  HIP_TRY(s->d_rdesc.ensure(4 * sizeof(CopyDesc)));
  char* snap = s->d_rsnap.as<char>();
  const CopyDesc d[4] = {{s->d_ws.p, snap, wsb}, {s->d_stats.p, snap + wsb, stb},
                         {snap, s->d_ws.p, wsb}, {snap + wsb, s->d_stats.p, stb}};
  if (std::memcmp(d, s->rec_desc, sizeof(d)) != 0) {  // only when an address changed
    HIP_TRY(hipMemcpyAsync(s->d_rdesc.p, d, sizeof(d), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::memcpy(s->rec_desc, d, sizeof(d));
  }
  return VXP_OK;
}

struct RecCtx {
  DevCfg cfg;
  DevBodies B;
  DevColliders C;
  WorldState* ws;
  vxp_stats* st;
  uint32_t nw, ns, wbits, id_bits, swv, P, WP, WI, CA;
  const uint32_t* npairs;  // device: the step's pairs
  const uint32_t* maxw;    // device: the most pairs of one world
  uint32_t* rsize;         // device: [0] probe work items, [1] contact candidates
  DevManifold *prev, *ml;
};

static RecCtx recCtx(vxp_system* s, int par) {
  RecCtx c;
  c.cfg = devCfg(s);
  c.cfg.pair_color_max = s->tune.use[vxp_system::kPairColorMax];
  c.B = devBodies(s);
  c.C = devColliders(s);
  c.ws = s->d_ws.as<WorldState>();
  c.st = s->d_stats.as<vxp_stats>();
  c.nw = s->nw;
  c.ns = (uint32_t)slots(s);
  c.wbits = worldBits(s);
  c.id_bits = idBits(s);
  c.swv = sweep_waves();
  c.P = s->rc_pairs;
  c.WP = s->rc_wp;
  c.WI = s->rc_wi;
  c.CA = s->rc_cand;
  c.npairs = s->d_swo.as<uint32_t>() + c.swv;
  c.maxw = c.nw > 1 ? s->d_maxw.as<uint32_t>() : c.npairs;
  c.rsize = s->d_rsize.as<uint32_t>();
  c.ml = s->d_ml[par].as<DevManifold>();
  c.prev = s->d_ml[par ^ 1].as<DevManifold>();
  return c;
}

// Segment 1: snapshot, step begin, fat boxes, sort-and-sweep pairs, pair keys,
// pair processing, the cache flags' scan; publishes the bounded counts.
static int recSeg1(vxp_system* s, const RecCtx& c) {
  hipStream_t st = s->stream;
  void* tmp = s->d_sort_tmp.p;
  const size_t tb = s->d_sort_tmp.bytes;
  uint32_t* ctr = s->d_step_ctr.as<uint32_t>();
  launch_copy_many(st, s->d_rdesc.as<CopyDesc>(), 2);
  launch_step_begin(st, c.cfg, c.B, c.ws, c.st, c.nw, ctr);
  launch_rec_carried(st, c.ws, c.nw, c.rsize);
  launch_fat_aabb(st, c.cfg, c.B, c.C, c.ws, c.ns, nullptr);
  if (!launch_sweep_sort(st, c.cfg, c.B, c.ws, s->d_svals2.as<uint32_t>(), c.nw)) {
    launch_sweep_keys(st, c.cfg, c.B, c.ws, s->d_skeys.as<uint64_t>(), s->d_svals.as<uint32_t>(),
                      c.ns, c.nw);
    HIP_TRY(sort_pairs_u64_u32(tmp, tb, s->d_skeys.as<uint64_t>(), s->d_skeys2.as<uint64_t>(),
                               s->d_svals.as<uint32_t>(), s->d_svals2.as<uint32_t>(), c.ns,
                               32 + c.wbits, st));
  }
  launch_sweep_count(st, c.cfg, c.B, c.ws, s->d_svals2.as<uint32_t>(), c.ns, s->d_slo.as<float4>(),
                     s->d_shi.as<float4>(), s->d_send.as<uint32_t>(), s->d_wcnt.as<uint32_t>());
  HIP_TRY(exclusive_scan_u32(tmp, tb, s->d_wcnt.as<uint32_t>(), s->d_woff.as<uint32_t>(),
                             c.ns + 1, st));
  const bool narrow = 2 * c.id_bits + c.wbits <= 32;
  launch_sweep_pairs(st, c.ws, s->d_slo.as<float4>(), s->d_shi.as<float4>(), c.ns,
                     s->d_send.as<uint32_t>(), s->d_woff.as<uint32_t>(), s->d_swc.as<uint32_t>(),
                     nullptr, nullptr, true, c.id_bits, narrow, c.nw > 1);
  HIP_TRY(exclusive_scan_u32(tmp, tb, s->d_swc.as<uint32_t>(), s->d_swo.as<uint32_t>(),
                             c.swv + 1, st));
  if (c.nw > 1) launch_world_pairs(st, c.ws, c.nw, nullptr, s->MB, s->d_maxw.as<uint32_t>());
  launch_sweep_pairs(st, c.ws, s->d_slo.as<float4>(), s->d_shi.as<float4>(), c.ns,
                     s->d_send.as<uint32_t>(), s->d_woff.as<uint32_t>(), nullptr,
                     s->d_swo.as<uint32_t>(), s->d_pairs_raw.as<uint64_t>(), false, c.id_bits,
                     narrow, false, c.P);
  // the pair keys in canonical order (as the eager step sorts them): per world
  // in one workgroup when every world holds at most 16384 pairs, device-wide
  // otherwise; both are recorded, each acting on a count gated on the device
  const uint32_t kbits = 2 * c.id_bits + c.wbits;
  uint64_t* keys = s->d_pairs.as<uint64_t>();
  const uint32_t* radix_n = c.npairs;
  if (2 * c.id_bits <= 32) {
    launch_rec_gate(st, c.npairs, c.maxw, 16384, c.rsize);
    launch_segment_sort(st, s->d_pairs_raw.p, kbits <= 32, c.P, c.nw, c.id_bits, 16384, keys,
                        c.rsize + 4);
    radix_n = c.rsize + 3;
  }
  if (kbits <= 32) {
    uint32_t* raw = s->d_pairs_raw.as<uint32_t>();
    HIP_TRY(sort_keys_u32_n(tmp, tb, raw, raw + c.P, c.P, radix_n, kbits, st));
    launch_expand_keys32(st, raw + c.P, keys, c.P, c.id_bits, radix_n);
  } else {
    HIP_TRY(sort_keys_u64_n(tmp, tb, s->d_pairs_raw.as<uint64_t>(), keys, c.P, radix_n, kbits,
                            st));
    launch_expand_keys(st, keys, c.P, c.id_bits, radix_n);
  }
  launch_pair_process(st, c.cfg, c.B, c.C, c.ws, c.st, keys, c.P, s->d_pr.as<PairRec>(), c.prev,
                      nullptr, s->d_cflag.as<uint32_t>(), s->d_produce.as<uint32_t>(),
                      s->d_pvchk.as<uint32_t>(), s->d_wlist.as<uint32_t>(), ctr,
                      s->d_counts.as<uint32_t>(), true, c.npairs);
  HIP_TRY(exclusive_scan_u32_n(tmp, s->d_cflag.as<uint32_t>(), s->d_cidx.as<uint32_t>(), c.P + 1,
                               c.npairs, 1, st));
  launch_publish4(st, s->d_read + kPub1, c.npairs, c.maxw, ctr + kCtrChunkBound, c.rsize + 2, 1);
  HIP_TRY(hipGetLastError());
  return VXP_OK;
}

// Segment 2: the brick-pair cache replay, the probe's work items and its count
// pass; publishes the work items and the candidates.
static int recSeg2(vxp_system* s, const RecCtx& c) {
  hipStream_t st = s->stream;
  void* tmp = s->d_sort_tmp.p;
  PairRec* pr = s->d_pr.as<PairRec>();
  launch_cache_replay(st, c.cfg, c.ws, c.st, pr, s->d_cache_keys.as<CacheKeyRec>(),
                      s->d_cflag.as<uint32_t>(), s->d_cidx.as<uint32_t>(), nullptr,
                      s->d_clist.as<uint32_t>(), c.P, c.nw, s->d_counts.as<uint32_t>(), c.npairs);
  HIP_TRY(exclusive_scan_u32_n(tmp, s->d_counts.as<uint32_t>(), s->d_wi_off.as<uint32_t>(), c.P,
                               c.npairs, 0, st));
  launch_rec_total_wi(st, s->d_wi_off.as<uint32_t>(), s->d_counts.as<uint32_t>(), c.npairs,
                      c.rsize);
  const ProbeShape shape{s->tune.use[vxp_system::kProbeThreads],
                         s->tune.use[vxp_system::kProbeXcdRun]};
  launch_wi_fill(st, s->d_counts.as<uint32_t>(), s->d_wi_off.as<uint32_t>(), c.P,
                 s->d_wi_pair.as<uint32_t>(), s->d_wi_ccnt.as<uint32_t>(),
                 c.npairs);
  launch_probe_count(st, c.cfg, c.B, c.C, pr, s->d_wi_pair.as<uint32_t>(),
                     s->d_wi_off.as<uint32_t>(), c.WI, s->d_wi_ccnt.as<uint32_t>(),
                     s->d_pvchk.as<uint32_t>(), s->d_lane_n.as<uint8_t>(),
                     s->d_wi_list.as<uint32_t>(), s->d_step_ctr.as<uint32_t>() + kCtrProbeItems,
                     shape, c.rsize);
  HIP_TRY(exclusive_scan_u32_n(tmp, s->d_wi_ccnt.as<uint32_t>(), s->d_wi_cbase.as<uint32_t>(),
                               c.WI + 1, c.rsize, 1, st));
  launch_rec_total_cand(st, s->d_wi_cbase.as<uint32_t>(), c.rsize);
  launch_publish4(st, s->d_read + kPub2, c.rsize, c.rsize + 1, nullptr, nullptr, 2);
  HIP_TRY(hipGetLastError());
  return VXP_OK;
}

// Segment 3: the probe's write pass, the manifold reduction and compaction,
// islands and the solver; publishes the manifold count.
static int recSeg3(vxp_system* s, const RecCtx& c) {
  hipStream_t st = s->stream;
  void* tmp = s->d_sort_tmp.p;
  PairRec* pr = s->d_pr.as<PairRec>();
  uint32_t* ctr = s->d_step_ctr.as<uint32_t>();
  const ProbeShape shape{s->tune.use[vxp_system::kProbeThreads],
                         s->tune.use[vxp_system::kProbeXcdRun]};
  launch_probe_write(st, c.cfg, c.B, c.C, pr, s->d_wi_pair.as<uint32_t>(),
                     s->d_wi_off.as<uint32_t>(), s->d_wi_list.as<uint32_t>(), c.WI * 64,
                     s->d_wi_cbase.as<uint32_t>(), s->d_lane_n.as<uint8_t>(),
                     s->d_cand.as<Cand>(), ctr + kCtrProbeItems);
  launch_reduce(st, c.cfg, c.B, c.C, c.ws, c.st, pr, s->d_pairs.as<uint64_t>(), c.P,
                s->d_wlist.as<uint32_t>(), ctr, s->d_wi_off.as<uint32_t>(),
                s->d_wi_cbase.as<uint32_t>(), s->d_wi_ccnt.as<uint32_t>(),
                s->d_pvchk.as<uint32_t>(), s->d_cand.as<Cand>(), c.prev,
                s->d_cache_keys.as<CacheKeyRec>(), s->d_cache_vals.as<CacheVal>(),
                s->d_cstage.as<CacheVal>(), s->d_staging.as<DevManifold>(),
                s->d_produce.as<uint32_t>(), s->d_rstat.as<RStat>(), s->d_coop.as<uint32_t>());
  HIP_TRY(exclusive_scan_u32_n(tmp, s->d_produce.as<uint32_t>(), s->d_pidx.as<uint32_t>(), c.P,
                               c.npairs, 0, st));
  launch_world_segments(st, c.ws, c.nw, s->d_total.as<uint32_t>());
  launch_scatter(st, c.ws, pr, c.P, s->d_produce.as<uint32_t>(), s->d_pidx.as<uint32_t>(),
                 s->d_staging.as<DevManifold>(), c.ml, c.npairs);
  launch_copy_skipped(st, c.ws, c.prev, c.ml, c.nw);
  if (s->cfg.enable_coloring)
    launch_ref_coloring(st, c.cfg, c.ws, c.ml, s->d_ref_masks.as<uint64_t>(),
                        s->d_ref_colors.as<uint8_t>(), c.nw);
  SolverScratch S;
  S.ws_pos = s->s_pos.as<float4>();
  S.ws_rot = s->s_rot.as<float4>();
  S.ws_lv = s->s_lv.as<float4>();
  S.ws_av = s->s_av.as<float4>();
  S.ws_ii = s->s_ii.as<float>();
  S.adj_cnt = s->s_adj_cnt.as<uint32_t>();
  S.adj_off = s->s_adj_off.as<uint32_t>();
  S.adj = s->s_adj.as<uint32_t>();
  S.color = s->s_color.as<int32_t>();
  S.order = s->s_order.as<uint32_t>();
  S.rec = s->s_rec.as<float4>();
  S.pts = s->s_pts.as<float4>();
  S.tstamp = nullptr;
  S.upar = s->s_upar.as<uint32_t>();
  S.ucnt = s->s_ucnt.as<uint32_t>();
  S.uoff = s->s_uoff.as<uint32_t>();
  S.ubody = s->s_ubody.as<uint32_t>();
  S.uorder = s->s_uorder.as<uint32_t>();
  S.units = s->s_units.as<uint4>();
  S.uqueue = s->s_uqueue.as<uint32_t>();
  S.pcnt = s->s_pcnt.as<uint32_t>();
  S.fq = s->s_fq.as<uint32_t>();
  S.cend = s->s_cend.as<uint4>();
  // island labels and the body working state beside the solver's prep
  HIP_TRY(hipEventRecord(s->ev_fork, st));
  HIP_TRY(hipStreamWaitEvent(s->side, s->ev_fork, 0));
  if (s->cfg.enable_islands)
    launch_islands(s->side, c.cfg, c.B, c.ws, c.st, c.ml, c.prev, islandArrays(s), c.nw, s->rc_mm,
                   s->lds_limit);
  launch_body_state(s->side, c.cfg, c.B, c.ws, S, c.ns);
  HIP_TRY(hipEventRecord(s->ev_join, s->side));
  launch_solve_prep(st, c.cfg, c.B, c.ws, c.st, c.ml, S, s->s_prio.as<uint32_t>(), ctr, c.nw,
                    s->rc_mm, s->lds_limit, s->d_total.as<uint32_t>());
  HIP_TRY(hipStreamWaitEvent(st, s->ev_join, 0));
  launch_unit_order(st, S, ctr, s->s_uqueue2.as<uint32_t>());
  launch_solve_units(st, c.cfg, c.B, c.ws, c.ml, S, ctr, c.nw, s->n_cu, s->lds_limit,
                     unitThreads(s));
  launch_solve_tail(st, c.cfg, c.B, c.ws, c.st, S, ctr, c.nw, s->n_cu);
  launch_publish4(st, s->d_read + kPub3, s->d_total.as<uint32_t>(), s->d_total.as<uint32_t>() + 1,
                  nullptr, nullptr, 3);
  HIP_TRY(hipGetLastError());
  return VXP_OK;
}

static int recRecord(vxp_system* s, int par) {
  auto& R = s->rec[par];
  const auto t0 = std::chrono::steady_clock::now();
  const RecCtx c = recCtx(s, par);
  for (int k = 0; k < 3; k++) {
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = k == 0 ? recSeg1(s, c) : k == 1 ? recSeg2(s, c) : recSeg3(s, c);
    const hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc != VXP_OK) {
      if (e == hipSuccess && g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_TRY(e);
    if (R.g[k]) {
      (void)hipGraphExecDestroy(R.g[k]);
      R.g[k] = nullptr;
    }
    const hipError_t ie = hipGraphInstantiate(&R.g[k], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(ie);
  }
  R.sig = recSignature(s, par);
  s->rec_records++;
  s->rec_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return VXP_OK;
}

// Launches one recorded segment and, when asked, waits for its publication
// (polling the pinned words, as readSizes does).
static int recRun(vxp_system* s, int par, int k, uint32_t slot, uint32_t seq, uint32_t* out,
                  int nout) {
  __atomic_store_n(&s->h_read[slot], 0u, __ATOMIC_RELEASE);
  HIP_TRY(hipGraphLaunch(s->rec[par].g[k], s->stream));
  if (!out) return VXP_OK;
  for (uint32_t spin = 1;; spin++) {
    if (__atomic_load_n(&s->h_read[slot], __ATOMIC_ACQUIRE) == seq) break;
    if ((spin & 1023u) == 0) {
      hipError_t e = hipStreamQuery(s->stream);
      if (e != hipSuccess && e != hipErrorNotReady) {
        g_err = std::string("hipStreamQuery: ") + hipGetErrorString(e);
        return VXP_ERR_DEVICE;
      }
      if (e == hipSuccess && __atomic_load_n(&s->h_read[slot], __ATOMIC_ACQUIRE) != seq) {
        g_err = "recorded step: the stream drained without the published counts";
        return VXP_ERR_DEVICE;
      }
    }
    __builtin_ia32_pause();
  }
  for (int i = 0; i < nout; i++) out[i] = __atomic_load_n(&s->h_read[slot + 1 + i], __ATOMIC_RELAXED);
  return VXP_OK;
}

// A capacity that holds `need`: twice it (the recorded launches loop over their
// items, so a larger capacity costs memory, not time, and fewer recordings).
static uint32_t grown(uint32_t need, uint32_t have) {
  const uint64_t g = std::max<uint64_t>(2 * (uint64_t)need + 1024, (uint64_t)have * 2);
  return (uint32_t)std::min<uint64_t>(g, 0xFFFFFFF0ull);
}

static int stepRecorded(vxp_system* s) {
  // first capacities: the last eager step's counts with headroom
  if (!s->rc_pairs) {
    s->rc_mm = s->last_max_man + s->last_max_man / 4 + 64;
    s->rc_carry = grown(s->n_manifolds, 256);
    s->rc_pairs = grown(s->n_pairs, 1024);
    s->rc_wp = s->nw > 1 ? std::min(s->rc_pairs, grown(s->n_pairs / s->nw, 1024)) : s->rc_pairs;
    s->rc_wi = grown(s->last_wi, 1024);
    s->rc_cand = grown(s->last_cand, 4096);
  }
  s->cur ^= 1;  // ContactManager.beginFrame: the current manifolds become the previous ones
  const int par = s->cur;
  for (int attempt = 0;; attempt++) {
    RC(recEnsure(s));
    if (!s->rec[par].g[0] || s->rec[par].sig != recSignature(s, par)) RC(recRecord(s, par));
    uint32_t c1[4];
    RC(recRun(s, par, 0, kPub1, 1, c1, 4));
    // the previous recorded step's most manifolds of one world (its segment 3
    // ran before this segment 1): the islands' and solver prep's LDS request
    // of the next recording
    if (__atomic_load_n(&s->h_read[kPub3], __ATOMIC_ACQUIRE) == 3) {
      s->last_max_man = __atomic_load_n(&s->h_read[kPub3 + 2], __ATOMIC_RELAXED);
      if (s->last_max_man > s->rc_mm) s->rc_mm = s->last_max_man + s->last_max_man / 4 + 64;
    }
    const uint32_t npairs = c1[0], maxw = c1[1], bound = c1[2], carried = c1[3];
    // the most pairs of one world only sizes the islands' and the solver
    // prep's LDS request (their kernels check the fit per world): a larger
    // value is taken for the next recording, without a redo
    if (maxw > s->rc_wp) s->rc_wp = std::min(s->rc_pairs, grown(maxw, s->rc_wp));
    if (npairs <= s->rc_pairs && bound <= s->rc_wi && carried <= s->rc_carry) {
      s->n_pairs = npairs;
      break;
    }
    // past a capacity: roll the world states and stats back, grow, record again
    launch_copy_many(s->stream, s->d_rdesc.as<CopyDesc>() + 2, 2);
    HIP_TRY(hipGetLastError());
    s->rec_redos++;
    if (npairs > s->rc_pairs) s->rc_pairs = grown(npairs, s->rc_pairs);
    if (bound > s->rc_wi) s->rc_wi = grown(bound, s->rc_wi);
    if (carried > s->rc_carry) s->rc_carry = grown(carried, s->rc_carry);
    if (attempt == 4) {
      g_err = "recorded step: the capacities did not settle";
      return VXP_ERR_DEVICE;
    }
  }
  uint32_t c2[2];
  RC(recRun(s, par, 1, kPub2, 2, c2, 2));
  s->last_wi = c2[0];
  s->last_cand = c2[1];
  if (c2[1] > s->rc_cand) {  // segment 3 only: nothing it reads has been written yet
    s->rc_cand = grown(c2[1], s->rc_cand);
    RC(recEnsure(s));
    RC(recRecord(s, par));
  }
  RC(recRun(s, par, 2, kPub3, 3, nullptr, 0));
  s->rec_nman_pending = true;
  s->rec_steps++;
  s->dev_newer = true;
  return VXP_OK;
}

// The manifold count of the last recorded step (published at its end).
int settleManifolds(vxp_system* s) {
  if (!s->rec_nman_pending) return VXP_OK;
  RC(streamWait(s));
  if (__atomic_load_n(&s->h_read[kPub3], __ATOMIC_ACQUIRE) != 3) {
    g_err = "recorded step: the manifold count was not published";
    return VXP_ERR_DEVICE;
  }
  s->n_manifolds = __atomic_load_n(&s->h_read[kPub3 + 1], __ATOMIC_RELAXED);
  s->rec_nman_pending = false;
  return VXP_OK;
}

static void recRelease(vxp_system* s) {
  for (auto& R : s->rec)
    for (auto& g : R.g)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
}

int checkErrors(vxp_system* s) {
  HIP_TRY(hipStreamSynchronize(s->stream));
  RC(settleManifolds(s));
  profFlush(s);
  std::vector<WorldState> ws(s->nw);
  HIP_TRY(hipMemcpy(ws.data(), s->d_ws.p, ws.size() * sizeof(WorldState), hipMemcpyDeviceToHost));
  for (auto& w : ws) {
    if (w.error & kErrColours) {
      g_err = "solver colouring exceeded 1024 colours";
      return VXP_ERR_CAPACITY;
    }
    if (w.error & kErrTreeStack) {
      g_err = "a broad-phase tree query exceeded its stack (tree deeper than 63 levels)";
      return VXP_ERR_CAPACITY;
    }
    if (w.error & kErrLbvh) {
      g_err = "an LBVH broad-phase query walked past the size of its tree";
      return VXP_ERR_DEVICE;
    }
  }
  return VXP_OK;
}

int updateOnce(vxp_system* s, float dt) {
  const vxp_config& c = s->cfg;
  RC(ensureDevice(s));
  s->acc += dt;
  launch_update_begin(s->stream, s->d_stats.as<vxp_stats>(), s->nw, s->acc);
  uint32_t n = 0;
  while (s->acc >= c.fixed_timestep && n < c.max_substeps) {
    RC(stepAll(s));
    s->acc -= c.fixed_timestep;
    n++;
  }
  if (s->acc > c.fixed_timestep * (float)c.max_substeps) s->acc = c.fixed_timestep;
  launch_update_end(s->stream, devCfg(s), devBodies(s), s->d_stats.as<vxp_stats>(), s->nw, n);
  HIP_TRY(hipGetLastError());
  return VXP_OK;
}

}  // namespace vxh

extern "C" {

const char* vxp_last_error(void) { return g_err.c_str(); }

int vxp_profile_enable(vxp_system* s, int on) {
  if (!s) return VXP_ERR_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(s->stream));
  profFlush(s);
  if (on != VXP_PROFILE_OFF && on != VXP_PROFILE_ALL && on != VXP_PROFILE_SOLVE_UNITS)
    return VXP_ERR_ARGUMENT;
  s->prof = on;
  if (s->prof)
    for (int i = 0; i < VXP_K_COUNT; i++) {
      s->prof_ms[i] = 0.0;
      s->prof_n[i] = 0;
    }
  return VXP_OK;
}

int vxp_graph_enable(vxp_system* s, int on) {
  if (!s) return VXP_ERR_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(s->stream));
  RC(settleManifolds(s));
  s->rec_on = on != 0;
  return VXP_OK;
}

int vxp_graph_info_get(vxp_system* s, vxp_graph_info* out) {
  if (!s || !out) return VXP_ERR_ARGUMENT;
  out->recorded_steps = s->rec_steps;
  out->redos = s->rec_redos;
  out->recordings = s->rec_records;
  out->recording_ms = s->rec_ms;
  out->cap_pairs = s->rc_pairs;
  out->cap_world_pairs = s->rc_wp;
  out->cap_work_items = s->rc_wi;
  out->cap_candidates = s->rc_cand;
  return VXP_OK;
}

int vxp_profile_solver_phases(vxp_system* s, uint32_t w, uint64_t* ts) {
  if (!s || !ts || w >= s->nw || !s->s_tstamp.p) return VXP_ERR_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(ts, s->s_tstamp.as<uint64_t>() + (size_t)w * 16, 16 * 8, hipMemcpyDeviceToHost));
  return VXP_OK;
}

int vxp_tuning_get(vxp_system* s, vxp_tuning* out) {
  if (!s || !out) return VXP_ERR_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  const auto& t = s->tune;
  auto val = [&](uint32_t k) {
    if (t.forced[k]) return t.forced[k];
    if (k == vxp_system::kUnitThreads && !t.pick[k]) return s->nw >= 64 ? 256u : 512u;
    return t.pick[k];
  };
  out->pair_color_max = t.pick[vxp_system::kPairColorMax];
  out->windows = t.windows;
  out->fixed = t.fixed;
  out->unit_threads = val(vxp_system::kUnitThreads);
  out->broad_phase = val(vxp_system::kBroadPhase);
  out->probe_threads = val(vxp_system::kProbeThreads);
  out->probe_xcd_run = val(vxp_system::kProbeXcdRun);
  return VXP_OK;
}
int vxp_tuning_set(vxp_system* s, const vxp_tuning* t) {
  if (!s || !t) return VXP_ERR_ARGUMENT;
  if (t->unit_threads != 0 && t->unit_threads != 256 && t->unit_threads != 512) {
    g_err = "unit_threads must be 0, 256 or 512";
    return VXP_ERR_ARGUMENT;
  }
  if (t->broad_phase > VXP_BROAD_LBVH) {
    g_err = "broad_phase must be VXP_BROAD_AUTO, VXP_BROAD_SWEEP or VXP_BROAD_LBVH";
    return VXP_ERR_ARGUMENT;
  }
  if (t->probe_threads != 0 && t->probe_threads != 64 && t->probe_threads != 128 &&
      t->probe_threads != 256) {
    g_err = "probe_threads must be 0, 64, 128 or 256";
    return VXP_ERR_ARGUMENT;
  }
  if (t->probe_xcd_run > 4096 || (t->probe_xcd_run & (t->probe_xcd_run - 1))) {
    g_err = "probe_xcd_run must be 0 or a power of two up to 4096";
    return VXP_ERR_ARGUMENT;
  }
  auto& u = s->tune;
  u.forced[vxp_system::kUnitThreads] = t->unit_threads;
  u.forced[vxp_system::kBroadPhase] = t->broad_phase;
  u.forced[vxp_system::kProbeThreads] = t->probe_threads;
  u.forced[vxp_system::kProbeXcdRun] = t->probe_xcd_run;
  u.window_left = 0;
  u.pending = false;
  u.knob = 0;
  if (t->pair_color_max == 0xFFFFFFFFu) {  // a new round from the next step
    u.fixed = false;
    u.next_window = u.frame;
  } else {
    u.fixed = true;
    u.pick[vxp_system::kPairColorMax] = t->pair_color_max;
  }
  return VXP_OK;
}
int vxp_profile_get(vxp_system* s, uint32_t id, double* total_ms, uint64_t* launches) {
  if (!s || id >= VXP_K_COUNT || !total_ms || !launches) return VXP_ERR_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(s->stream));
  profFlush(s);
  *total_ms = s->prof_ms[id];
  *launches = s->prof_n[id];
  return VXP_OK;
}

int vxp_config_default(vxp_config* c) {
  if (!c) return VXP_ERR_ARGUMENT;
  std::memset(c, 0, sizeof(*c));
  c->gravity[0] = 0.0f;
  c->gravity[1] = -9.81f;
  c->gravity[2] = 0.0f;
  c->fixed_timestep = 1.0f / 60.0f;
  c->max_substeps = 8;
  c->max_bodies = 1024;
  c->default_friction = 0.6f;
  c->default_restitution = 0.2f;
  c->default_density = 1.0f;
  c->enable_sleeping = 1;
  c->enable_warm_starting = 1;
  c->enable_islands = 1;
  c->enable_coloring = 0;
  c->substeps = 4;
  c->velocity_iterations = 1;
  c->position_iterations = 2;
  c->hertz = 30.0f;
  c->damping_ratio = 1.0f;
  c->position_correction_rate = 0.2f;
  c->max_position_correction = 0.04f;
  c->slop = 0.005f;
  c->warm_start_scale = 0.8f;
  c->disable_rotation = 0;
  c->warm_starting = 1;
  c->enable_friction = 1;
  c->enable_restitution = 1;
  c->use_brick_sat = 1;
  c->use_brick_pair_cache = 1;
  c->brick_pair_cache_size = 1024;
  c->pair_cap = 0;
  c->pair_order = VXP_PAIR_ORDER_KEY;
  c->solve_order = VXP_SOLVE_ORDER_COLOR;
  c->cpu_threads = 1;
  return VXP_OK;
}

int vxp_system_create(const vxp_config* cfg, uint32_t n_worlds, vxp_system** out) {
  if (!cfg || !out || n_worlds == 0 || cfg->max_bodies == 0 || cfg->substeps == 0) {
    g_err = "invalid configuration";
    return VXP_ERR_ARGUMENT;
  }
  if (cfg->pair_order > VXP_PAIR_ORDER_TREE) {
    g_err = "unknown pair_order";
    return VXP_ERR_ARGUMENT;
  }
  if (cfg->solve_order != VXP_SOLVE_ORDER_COLOR) {
    g_err = "the GPU pipeline solves in colour order (VXP_SOLVE_ORDER_COLOR)";
    return VXP_ERR_ARGUMENT;
  }
  uint32_t cap = cfg->brick_pair_cache_size ? cfg->brick_pair_cache_size : 1024;
  if (cap * sizeof(CacheKeyRec) > 60 * 1024) {
    g_err = "brick_pair_cache_size above 1536 entries is not supported (LDS-resident replay)";
    return VXP_ERR_ARGUMENT;
  }
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) {
    g_err = "no HIP device";
    return VXP_ERR_DEVICE;
  }
  auto* s = new vxp_system();
  s->cfg = *cfg;
  s->nw = n_worlds;
  s->MB = cfg->max_bodies;
  s->id_cap = cfg->max_bodies + 2;
  size_t ns = slots(s);
  if (treeMode(s)) {  // leaves <= max_bodies, internal nodes one fewer
    s->t_nc = 2 * s->MB + 2;
    s->t_nodes.assign((size_t)n_worlds * 2 * s->t_nc, TreeNode{});
    s->t_heads.assign((size_t)n_worlds * 2, TreeHead{kTreeNull, kTreeNull, 0, 0});
    s->t_leaf.assign(ns, kTreeNull);
    s->t_host_dirty = true;
  }
  s->pos.assign(ns, make_float4(0, 0, 0, 0));
  s->rot.assign(ns, make_float4(0, 0, 0, 1));
  s->lv.assign(ns, make_float4(0, 0, 0, 0));
  s->av.assign(ns, make_float4(0, 0, 0, 0));
  s->fat_lo.assign(ns, make_float4(0, 0, 0, 0));
  s->fat_hi.assign(ns, make_float4(0, 0, 0, 0));
  s->inv_mass.assign(ns, 0.0f);
  s->iil.assign(ns * 9, 0.0f);
  s->fric.assign(ns, 0.5f);
  s->rest.assign(ns, 0.0f);
  s->ldamp.assign(ns, 0.0f);
  s->adamp.assign(ns, 0.0f);
  s->sleep_timer.assign(ns, 0.0f);
  s->type.assign(ns, 0);
  s->sleeping.assign(ns, 0);
  s->active.assign(ns, 0);
  s->collider.assign(ns, -1);
  s->id.assign(ns, 0);
  s->idmap.assign((size_t)n_worlds * s->id_cap, -1);
  s->generation.assign(ns, 0);
  s->next_id.assign(n_worlds, 1);
  s->body_count.assign(n_worlds, 0);
  s->free_list.assign(n_worlds, {});
  s->n_total.assign(n_worlds, 0);
  s->n_dyn.assign(n_worlds, 0);
  s->n_static.assign(n_worlds, 0);
  int rc = VXP_OK;
  auto fail = [&](int r) {
    if (s->h_read) (void)hipHostFree(s->h_read);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return r;
  };
  {
    int dev = 0, lds = 0;
    (void)hipGetDevice(&dev);
    s->device = dev;
    // the allocation-free entry point's scratch address on this device
    if ((rc = noallocInit(dev))) return fail(rc);
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        lds > 24 * 1024)
      s->lds_limit = (size_t)lds - 20 * 1024;  // the solver's static LDS
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0)
      s->n_cu = (uint32_t)cu;
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) {
    g_err = "hipStreamCreate failed";
    return fail(VXP_ERR_DEVICE);
  }
  // coherent (fine-grained) so that a kernel's writes reach the host while it polls
  if (hipHostMalloc((void**)&s->h_read, 256, hipHostMallocCoherent | hipHostMallocMapped) !=
          hipSuccess ||
      hipHostGetDevicePointer((void**)&s->d_read, s->h_read, 0) != hipSuccess) {
    g_err = "pinned host allocation failed";
    return fail(VXP_ERR_OOM);
  }
  std::memset(s->h_read, 0, 256);
  std::vector<WorldState> ws(n_worlds);
  std::memset(ws.data(), 0, ws.size() * sizeof(WorldState));
  for (WorldState& x : ws) x.isl_need = 1;  // IslandManager.init (src/islands.zig:248-270)
  std::vector<vxp_stats> st(n_worlds);
  std::memset(st.data(), 0, st.size() * sizeof(vxp_stats));
  std::vector<CacheKeyRec> ck((size_t)n_worlds * cap);
  std::memset(ck.data(), 0, ck.size() * sizeof(CacheKeyRec));
  if ((rc = up(s->d_ws, ws)) || (rc = up(s->d_stats, st)) || (rc = up(s->d_cache_keys, ck)))
    return fail(rc);
  if (s->d_cache_vals.ensure((size_t)n_worlds * cap * sizeof(CacheVal)) != hipSuccess ||
      s->d_ml[0].ensure(sizeof(DevManifold)) != hipSuccess ||
      s->d_ml[1].ensure(sizeof(DevManifold)) != hipSuccess) {
    g_err = "device allocation failed";
    return fail(VXP_ERR_OOM);
  }
  *out = s;
  return VXP_OK;
}

int vxp_system_destroy(vxp_system* s) {
  if (!s) return VXP_ERR_ARGUMENT;
  (void)hipStreamSynchronize(s->stream);
  recRelease(s);
  for (DevBuf* b : allBufs(s)) b->release();
  profFlush(s);
  for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
  if (s->h_read) (void)hipHostFree(s->h_read);
  if (s->tune.a) (void)hipEventDestroy(s->tune.a);
  if (s->tune.b) (void)hipEventDestroy(s->tune.b);
  (void)hipStreamSynchronize(s->side);
  (void)hipEventDestroy(s->ev_fork);
  (void)hipEventDestroy(s->ev_join);
  (void)hipStreamDestroy(s->side);
  (void)hipStreamDestroy(s->stream);
  delete s;
  return VXP_OK;
}

// Voxel bytes from solid occupancy on the device (k_collider_classify: the
// classifyFromNeighbors rule of src/collision_voxel.zig:213-235 over the six
// face neighbours, across bricks by exact position), then vxp_collider_create.
int vxp_collider_create_solid(vxp_system* s, float voxel_size, uint32_t n, const int32_t* bp,
                              const uint64_t* solid, uint32_t* out) {
  if (!s || !out || (n && (!bp || !solid)) || !(voxel_size > 0.0f)) return VXP_ERR_ARGUMENT;
  std::unordered_map<uint64_t, int32_t> at;  // exact position -> input brick
  auto key = [](int32_t x, int32_t y, int32_t z) {
    return ((uint64_t)(uint32_t)x << 42) ^ ((uint64_t)(uint32_t)y << 21) ^ (uint64_t)(uint32_t)z;
  };
  for (uint32_t i = 0; i < n; i++) {
    const int32_t* p = bp + 3 * i;
    if (p[0] < -(1 << 20) || p[0] >= (1 << 20) || p[1] < -(1 << 20) || p[1] >= (1 << 20) ||
        p[2] < -(1 << 20) || p[2] >= (1 << 20) || !at.emplace(key(p[0], p[1], p[2]), (int32_t)i).second) {
      g_err = "collider bricks must have distinct positions within +-2^20";
      return VXP_ERR_ARGUMENT;
    }
  }
  std::vector<int32_t> nbr((size_t)n * 6, -1);
  const int32_t d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (uint32_t i = 0; i < n; i++)
    for (int k = 0; k < 6; k++) {
      auto it = at.find(key(bp[3 * i] + d[k][0], bp[3 * i + 1] + d[k][1], bp[3 * i + 2] + d[k][2]));
      if (it != at.end()) nbr[(size_t)i * 6 + k] = it->second;
    }
  std::vector<uint8_t> vox((size_t)n * 512);
  if (n) {
    HIP_TRY(s->d_cb_solid.ensure((size_t)n * 64));
    HIP_TRY(s->d_cb_nbr.ensure((size_t)n * 24));
    HIP_TRY(s->d_cb_vox.ensure((size_t)n * 512));
    HIP_TRY(hipMemcpyAsync(s->d_cb_solid.p, solid, (size_t)n * 64, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_cb_nbr.p, nbr.data(), (size_t)n * 24, hipMemcpyHostToDevice, s->stream));
    launch_collider_classify(s->stream, s->d_cb_solid.as<uint64_t>(), s->d_cb_nbr.as<int32_t>(), n,
                             s->d_cb_vox.as<uint8_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(vox.data(), s->d_cb_vox.p, (size_t)n * 512, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  return vxp_collider_create(s, voxel_size, n, bp, vox.data(), out);
}

int vxp_collider_info(vxp_system* s, uint32_t collider, int32_t bounds[6], float aabb[6],
                      uint32_t* solid) {
  if (!s || !bounds || !aabb || !solid || collider >= s->colliders.size()) return VXP_ERR_ARGUMENT;
  const HostCollider& c = s->colliders[collider];
  for (int k = 0; k < 3; k++) {
    bounds[k] = c.minb[k];
    bounds[3 + k] = c.maxb[k];
    aabb[k] = c.lo[k];
    aabb[3 + k] = c.hi[k];
  }
  *solid = c.solid;
  return VXP_OK;
}

int vxp_collider_bricks(vxp_system* s, uint32_t collider, int32_t* bp, uint8_t* vox, uint32_t cap,
                        uint32_t* count) {
  if (!s || !count || collider >= s->colliders.size()) return VXP_ERR_ARGUMENT;
  const HostCollider& c = s->colliders[collider];
  const uint32_t n = (uint32_t)(c.pos.size() / 3);
  *count = n;
  if (cap && bp && vox) {
    const uint32_t k = std::min(n, cap);
    std::memcpy(bp, c.pos.data(), (size_t)k * 12);
    std::memcpy(vox, c.vox.data(), (size_t)k * 512);
  }
  return VXP_OK;
}

int vxp_collider_mass_properties(vxp_system* s, uint32_t collider, float density, float out[13]) {
  if (!s || !out || collider >= s->colliders.size()) return VXP_ERR_ARGUMENT;
  MassProps m;
  V3 com;
  RC(massProperties(s, s->colliders[collider], density, &m, &com));
  out[0] = m.inv_mass;
  for (int i = 0; i < 9; i++) out[1 + i] = m.inv_inertia.m[i];
  out[10] = com.x;
  out[11] = com.y;
  out[12] = com.z;
  return VXP_OK;
}

int vxp_collider_create(vxp_system* s, float voxel_size, uint32_t n, const int32_t* bp,
                        const uint8_t* vox, uint32_t* out) {
  if (!s || !out || (n && (!bp || !vox)) || !(voxel_size > 0.0f)) return VXP_ERR_ARGUMENT;
  HostCollider c;
  c.vs = voxel_size;
  // setCollisionVoxel semantics (src/collider.zig:232-239,267-273): bricks keyed
  // by hashBrickPos; colliding hashes share the first brick, later writes win.
  struct Tmp {
    int32_t p[3];
    uint8_t v[512];
  };
  std::vector<Tmp> merged;
  std::unordered_map<uint64_t, size_t> idx;
  for (uint32_t i = 0; i < n; i++) {
    uint64_t h = hashBrickPos(bp[3 * i], bp[3 * i + 1], bp[3 * i + 2]);
    auto it = idx.find(h);
    if (it == idx.end()) {
      Tmp t;
      t.p[0] = bp[3 * i];
      t.p[1] = bp[3 * i + 1];
      t.p[2] = bp[3 * i + 2];
      std::memset(t.v, 0, 512);
      it = idx.emplace(h, merged.size()).first;
      merged.push_back(t);
    }
    Tmp& t = merged[it->second];
    for (int j = 0; j < 512; j++)
      if (vox[512ull * i + j] & 7u) t.v[j] = vox[512ull * i + j];
  }
  std::sort(merged.begin(), merged.end(), [](const Tmp& a, const Tmp& b) {
    if (a.p[2] != b.p[2]) return a.p[2] < b.p[2];
    if (a.p[1] != b.p[1]) return a.p[1] < b.p[1];
    return a.p[0] < b.p[0];
  });
  for (size_t i = 0; i < merged.size(); i++) {
    c.pos.insert(c.pos.end(), merged[i].p, merged[i].p + 3);
    c.vox.insert(c.vox.end(), merged[i].v, merged[i].v + 512);
    c.by_hash[hashBrickPos(merged[i].p[0], merged[i].p[1], merged[i].p[2])] = (int32_t)i;
  }
  RC(colliderMetadata(s, c));
  *out = (uint32_t)s->colliders.size();
  s->colliders.push_back(std::move(c));
  s->coll_dirty = true;
  return VXP_OK;
}

static int createBody(vxp_system* s, uint32_t w, uint32_t type, const float pos[3],
                      const float rot[4], uint32_t col, float density, const MassProps* given,
                      vxp_handle* out) {
  if (!s || !out || !pos || !rot || w >= s->nw || type > VXP_BODY_DYNAMIC) return VXP_ERR_ARGUMENT;
  if (col != 0xFFFFFFFFu && col >= s->colliders.size()) return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  // allocateBody (src/system.zig:475-499)
  uint32_t index;
  if (!s->free_list[w].empty()) {
    index = s->free_list[w].back();
    s->free_list[w].pop_back();
  } else {
    uint32_t i = 0;
    for (; i < s->MB; i++)
      if (!s->active[(size_t)w * s->MB + i]) break;
    if (i == s->MB) {
      g_err = "MaxBodiesReached";
      return VXP_ERR_CAPACITY;
    }
    index = i;
  }
  size_t g = (size_t)w * s->MB + index;
  uint32_t bid = s->next_id[w]++;
  growIds(s, bid + 1);
  s->active[g] = 1;
  s->body_count[w]++;
  s->id[g] = bid;
  s->type[g] = (uint8_t)type;
  s->pos[g] = make_float4(pos[0], pos[1], pos[2], 0.0f);
  s->rot[g] = make_float4(rot[0], rot[1], rot[2], rot[3]);
  s->lv[g] = make_float4(0, 0, 0, 0);
  s->av[g] = make_float4(0, 0, 0, 0);
  s->collider[g] = col == 0xFFFFFFFFu ? -1 : (int32_t)col;
  s->sleep_timer[g] = 0.0f;
  s->fric[g] = s->cfg.default_friction;
  s->rest[g] = s->cfg.default_restitution;
  if (type == VXP_BODY_DYNAMIC) {
    MassProps mp{0.0f, M3{{0, 0, 0, 0, 0, 0, 0, 0, 0}}};
    if (given) {
      mp = *given;
    } else if (col != 0xFFFFFFFFu) {
      HostCollider& hc = s->colliders[col];
      uint32_t key;
      std::memcpy(&key, &density, 4);
      auto it = hc.mass_cache.find(key);
      if (it != hc.mass_cache.end())
        mp = it->second;
      else  // computed with every other new body's before anything reads it
        s->mass_pending.push_back({g, col, density});
    }
    s->inv_mass[g] = mp.inv_mass;
    for (int i = 0; i < 9; i++) s->iil[g * 9 + i] = mp.inv_inertia.m[i];
    s->ldamp[g] = 0.3f;
    s->adamp[g] = 0.5f;
    s->sleeping[g] = 0;
  } else {
    s->inv_mass[g] = 0.0f;
    for (int i = 0; i < 9; i++) s->iil[g * 9 + i] = 0.0f;
    s->ldamp[g] = 0.0f;
    s->adamp[g] = 0.0f;
    s->sleeping[g] = type == VXP_BODY_STATIC ? 1 : 0;
  }
  // getWorldAABB + insertBody inflate (src/body.zig:430-435, src/broad_phase.zig:150-163)
  V3 p{pos[0], pos[1], pos[2]};
  Box wb{p, p};
  if (col != 0xFFFFFFFFu) {
    const HostCollider& hc = s->colliders[col];
    wb = boxTransform(Box{V3{hc.lo[0], hc.lo[1], hc.lo[2]}, V3{hc.hi[0], hc.hi[1], hc.hi[2]}}, p,
                      Q4{rot[0], rot[1], rot[2], rot[3]});
  }
  Box fat = boxInflate(wb, 0.1f);
  s->fat_lo[g] = make_float4(fat.lo.x, fat.lo.y, fat.lo.z, 0.0f);
  s->fat_hi[g] = make_float4(fat.hi.x, fat.hi.y, fat.hi.z, 0.0f);
  if (treeMode(s)) {  // addStaticBody / addDynamicBody (src/broad_phase.zig:580-590)
    RC(treesToHost(s));
    const float lo[3] = {fat.lo.x, fat.lo.y, fat.lo.z}, hi[3] = {fat.hi.x, fat.hi.y, fat.hi.z};
    s->t_leaf[g] = treeInsertBody(hostTree(s, w, type == VXP_BODY_STATIC ? 0 : 1), bid, lo, hi);
    s->t_host_dirty = true;
  }
  s->idmap[(size_t)w * s->id_cap + bid] = (int32_t)index;
  s->n_total[w]++;
  if (type == VXP_BODY_DYNAMIC) s->n_dyn[w]++;
  if (type == VXP_BODY_STATIC) s->n_static[w]++;
  s->host_dirty = true;
  out->index = index;
  out->generation = s->generation[g];
  return VXP_OK;
}

int vxp_body_create_w(vxp_system* s, uint32_t w, uint32_t type, const float pos[3], const float rot[4],
                      uint32_t col, float density, vxp_handle* out) {
  return createBody(s, w, type, pos, rot, col, density, nullptr, out);
}

int vxp_body_create(vxp_system* s, uint32_t type, const float pos[3], const float rot[4], uint32_t col,
                    float density, vxp_handle* out) {
  return vxp_body_create_w(s, 0, type, pos, rot, col, density, out);
}

int vxp_body_create_with_mass_w(vxp_system* s, uint32_t w, const float pos[3], const float rot[4],
                                uint32_t col, float inv_mass, const float inv_inertia[9],
                                vxp_handle* out) {
  if (!inv_inertia) return VXP_ERR_ARGUMENT;
  MassProps mp;
  mp.inv_mass = inv_mass;
  for (int i = 0; i < 9; i++) mp.inv_inertia.m[i] = inv_inertia[i];
  return createBody(s, w, VXP_BODY_DYNAMIC, pos, rot, col, 1.0f, &mp, out);
}

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

// setGravity / setSolverConfig (src/system.zig:1034-1052)
int vxp_system_set_config(vxp_system* s, const vxp_config* cfg) {
  if (!s || !cfg) return VXP_ERR_ARGUMENT;
  if (!configChangeAllowed(s->cfg, *cfg)) {
    g_err = "only gravity and the SolverConfig fields can change after creation";
    return VXP_ERR_ARGUMENT;
  }
  s->cfg = *cfg;
  return VXP_OK;
}

static bool validHandle(vxp_system* s, uint32_t w, vxp_handle h) {
  if (w >= s->nw || h.index >= s->MB) return false;
  size_t g = (size_t)w * s->MB + h.index;
  return s->active[g] && s->generation[g] == h.generation;
}

int vxp_body_destroy_w(vxp_system* s, uint32_t w, vxp_handle h) {
  if (!s) return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  RC(resolveMasses(s));  // no pending entry may outlive its body's slot
  if (!validHandle(s, w, h)) return VXP_ERR_INVALID_HANDLE;
  size_t g = (size_t)w * s->MB + h.index;
  s->n_total[w]--;
  if (s->type[g] == VXP_BODY_DYNAMIC) s->n_dyn[w]--;
  if (s->type[g] == VXP_BODY_STATIC) s->n_static[w]--;
  s->idmap[(size_t)w * s->id_cap + s->id[g]] = -1;
  if (treeMode(s)) {  // BroadPhase.removeBody (src/broad_phase.zig:592-597)
    RC(treesToHost(s));
    treeRemoveBody(hostTree(s, w, s->type[g] == VXP_BODY_STATIC ? 0 : 1), s->t_leaf[g]);
    s->t_leaf[g] = kTreeNull;
    s->t_host_dirty = true;
  }
  s->active[g] = 0;
  s->generation[g]++;
  s->body_count[w]--;
  s->free_list[w].push_back(h.index);
  s->host_dirty = true;
  return VXP_OK;
}

int vxp_system_update(vxp_system* s, float dt) {
  if (!s) return VXP_ERR_ARGUMENT;
  RC(updateOnce(s, dt));
  return checkErrors(s);
}

int vxp_system_update_frames(vxp_system* s, float dt, uint32_t frames) {
  if (!s) return VXP_ERR_ARGUMENT;
  for (uint32_t f = 0; f < frames; f++) RC(updateOnce(s, dt));
  return checkErrors(s);
}

int vxp_system_step(vxp_system* s, float dt) {
  if (!s) return VXP_ERR_ARGUMENT;
  if (dt != s->cfg.fixed_timestep) {
    g_err = "step(dt) runs the fixed-timestep pipeline: dt must equal fixed_timestep";
    return VXP_ERR_ARGUMENT;
  }
  RC(stepAll(s));
  return checkErrors(s);
}

int vxp_bodies_get_w(vxp_system* s, uint32_t w, vxp_body_state* out, uint32_t cap, uint32_t* count) {
  if (!s || !count || w >= s->nw) return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  uint32_t n = std::min(cap, s->MB);
  for (uint32_t i = 0; i < n && out; i++) {
    size_t g = (size_t)w * s->MB + i;
    vxp_body_state& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.active = s->active[g];
    if (!o.active) continue;
    o.position[0] = s->pos[g].x; o.position[1] = s->pos[g].y; o.position[2] = s->pos[g].z;
    o.rotation[0] = s->rot[g].x; o.rotation[1] = s->rot[g].y; o.rotation[2] = s->rot[g].z;
    o.rotation[3] = s->rot[g].w;
    o.linear_velocity[0] = s->lv[g].x; o.linear_velocity[1] = s->lv[g].y; o.linear_velocity[2] = s->lv[g].z;
    o.angular_velocity[0] = s->av[g].x; o.angular_velocity[1] = s->av[g].y; o.angular_velocity[2] = s->av[g].z;
    o.sleep_timer = s->sleep_timer[g];
    o.id = s->id[g];
    o.body_type = s->type[g];
    o.is_sleeping = s->sleeping[g];
  }
  *count = n;
  return VXP_OK;
}

int vxp_body_set_state_w(vxp_system* s, uint32_t w, vxp_handle h, const float pos[3], const float rot[4],
                         const float lv[3], const float av[3]) {
  if (!s) return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  if (!validHandle(s, w, h)) return VXP_ERR_INVALID_HANDLE;
  size_t g = (size_t)w * s->MB + h.index;
  if (pos) s->pos[g] = make_float4(pos[0], pos[1], pos[2], 0.0f);
  if (rot) s->rot[g] = make_float4(rot[0], rot[1], rot[2], rot[3]);
  if (s->type[g] == VXP_BODY_DYNAMIC) {
    if (lv) s->lv[g] = make_float4(lv[0], lv[1], lv[2], 0.0f);
    if (av) s->av[g] = make_float4(av[0], av[1], av[2], 0.0f);
  }
  if (s->type[g] != VXP_BODY_STATIC) {  // wakeUp (src/body.zig:400-404)
    s->sleeping[g] = 0;
    s->sleep_timer[g] = 0.0f;
  }
  s->host_dirty = true;
  return VXP_OK;
}

// VoxelBody mutators (src/body.zig:266-414) on the host mirror, uploaded
// before the next step.  Same arithmetic as the reference: Vec3 add/scale/
// cross, Mat3.mulVec, getWorldInvInertia = (R * I^-1) * R^T.
int vxp_body_op_w(vxp_system* s, uint32_t w, vxp_handle h, uint32_t op, const float* v,
                  const float point[3]) {
  if (!s || op > VXP_OP_SLEEP) return VXP_ERR_ARGUMENT;
  if (op != VXP_OP_WAKE_UP && op != VXP_OP_SLEEP && !v) return VXP_ERR_ARGUMENT;
  if ((op == VXP_OP_APPLY_FORCE_AT_POINT || op == VXP_OP_APPLY_IMPULSE_AT_POINT) && !point)
    return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  RC(resolveMasses(s));  // the mutators read the inverse mass and inertia
  if (!validHandle(s, w, h)) return VXP_ERR_INVALID_HANDLE;
  const size_t g = (size_t)w * s->MB + h.index;
  const uint8_t type = s->type[g];
  const bool dyn = type == VXP_BODY_DYNAMIC;
  V3 lv{s->lv[g].x, s->lv[g].y, s->lv[g].z}, av{s->av[g].x, s->av[g].y, s->av[g].z};
  V3 pos{s->pos[g].x, s->pos[g].y, s->pos[g].z};
  Q4 rot{s->rot[g].x, s->rot[g].y, s->rot[g].z, s->rot[g].w};
  const float im = s->inv_mass[g];
  auto wake = [&] {  // wakeUp (src/body.zig:400-404)
    if (type == VXP_BODY_STATIC) return;
    s->sleeping[g] = 0;
    s->sleep_timer[g] = 0.0f;
  };
  auto linear = [&](V3 f) {  // applyForce / applyImpulse (src/body.zig:327-333,355-361)
    if (!dyn || im == 0.0f) return;
    lv = add(lv, scale(f, im));
    wake();
  };
  auto angular = [&](V3 t) {  // applyTorque / applyAngularImpulse (src/body.zig:346-353,374-381)
    if (!dyn) return;
    M3 ii{};
    if (im != 0.0f) {  // getWorldInvInertia (src/body.zig:384-390)
      M3 l;
      for (int i = 0; i < 9; i++) l.m[i] = s->iil[g * 9 + i];
      ii = worldInvInertia(rot, l);
    }
    av = add(av, mulVec(ii, t));
    wake();
  };
  const V3 a = v ? V3{v[0], v[1], v[2]} : V3{0.0f, 0.0f, 0.0f};
  switch (op) {
    case VXP_OP_APPLY_FORCE:
    case VXP_OP_APPLY_IMPULSE:
      linear(a);
      break;
    case VXP_OP_APPLY_FORCE_AT_POINT:
    case VXP_OP_APPLY_IMPULSE_AT_POINT:
      if (!dyn) break;
      linear(a);
      angular(cross(sub(V3{point[0], point[1], point[2]}, pos), a));
      break;
    case VXP_OP_APPLY_TORQUE:
    case VXP_OP_APPLY_ANGULAR_IMPULSE:
      angular(a);
      break;
    case VXP_OP_SET_LINEAR_VELOCITY:
      if (!dyn) break;
      lv = a;
      wake();
      break;
    case VXP_OP_SET_ANGULAR_VELOCITY:
      if (!dyn) break;
      av = a;
      wake();
      break;
    case VXP_OP_SET_POSITION:
      pos = a;
      wake();
      break;
    case VXP_OP_SET_ROTATION:
      rot = Q4{v[0], v[1], v[2], v[3]};
      wake();
      break;
    case VXP_OP_WAKE_UP:
      wake();
      break;
    case VXP_OP_SLEEP:  // sleep (src/body.zig:407-412)
      if (type == VXP_BODY_STATIC) break;
      s->sleeping[g] = 1;
      lv = V3{0.0f, 0.0f, 0.0f};
      av = V3{0.0f, 0.0f, 0.0f};
      break;
  }
  s->lv[g] = make_float4(lv.x, lv.y, lv.z, 0.0f);
  s->av[g] = make_float4(av.x, av.y, av.z, 0.0f);
  s->pos[g] = make_float4(pos.x, pos.y, pos.z, 0.0f);
  s->rot[g] = make_float4(rot.x, rot.y, rot.z, rot.w);
  s->host_dirty = true;
  return VXP_OK;
}

int vxp_body_world_aabb_w(vxp_system* s, uint32_t w, vxp_handle h, float out[6]) {
  if (!s || !out) return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  if (!validHandle(s, w, h)) return VXP_ERR_INVALID_HANDLE;
  const size_t g = (size_t)w * s->MB + h.index;
  V3 p{s->pos[g].x, s->pos[g].y, s->pos[g].z};
  Box b{p, p};
  if (s->collider[g] >= 0) {
    const HostCollider& hc = s->colliders[s->collider[g]];
    b = boxTransform(Box{V3{hc.lo[0], hc.lo[1], hc.lo[2]}, V3{hc.hi[0], hc.hi[1], hc.hi[2]}}, p,
                     Q4{s->rot[g].x, s->rot[g].y, s->rot[g].z, s->rot[g].w});
  }
  const float r[6] = {b.lo.x, b.lo.y, b.lo.z, b.hi.x, b.hi.y, b.hi.z};
  std::memcpy(out, r, sizeof(r));
  return VXP_OK;
}

int vxp_body_find_w(vxp_system* s, uint32_t w, uint32_t id, vxp_handle* out) {
  if (!s || !out || w >= s->nw) return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  if (id >= s->id_cap) return VXP_ERR_INVALID_HANDLE;
  const int32_t slot = s->idmap[(size_t)w * s->id_cap + id];
  if (slot < 0) return VXP_ERR_INVALID_HANDLE;
  const size_t g = (size_t)w * s->MB + (uint32_t)slot;
  if (!s->active[g] || s->id[g] != id) return VXP_ERR_INVALID_HANDLE;
  out->index = (uint32_t)slot;
  out->generation = s->generation[g];
  return VXP_OK;
}

int vxp_stats_get_w(vxp_system* s, uint32_t w, vxp_stats* out) {
  if (!s || !out || w >= s->nw) return VXP_ERR_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(out, s->d_stats.as<vxp_stats>() + w, sizeof(vxp_stats), hipMemcpyDeviceToHost));
  out->total_bodies = s->n_total[w];
  out->dynamic_bodies = s->n_dyn[w];
  out->static_bodies = s->n_static[w];
  return VXP_OK;
}

int vxp_stats_get_all(vxp_system* s, vxp_stats* out, uint32_t cap) {
  if (!s || !out || cap < s->nw) return VXP_ERR_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(out, s->d_stats.p, (size_t)s->nw * sizeof(vxp_stats), hipMemcpyDeviceToHost));
  for (uint32_t w = 0; w < s->nw; w++) {
    out[w].total_bodies = s->n_total[w];
    out[w].dynamic_bodies = s->n_dyn[w];
    out[w].static_bodies = s->n_static[w];
  }
  return VXP_OK;
}

static int readWorlds(vxp_system* s, std::vector<WorldState>& ws) {
  HIP_TRY(hipStreamSynchronize(s->stream));
  ws.resize(s->nw);
  HIP_TRY(hipMemcpy(ws.data(), s->d_ws.p, ws.size() * sizeof(WorldState), hipMemcpyDeviceToHost));
  return VXP_OK;
}

int vxp_manifolds_get_w(vxp_system* s, uint32_t w, vxp_manifold* out, uint32_t cap, uint32_t* count) {
  if (!s || !count || w >= s->nw) return VXP_ERR_ARGUMENT;
  std::vector<WorldState> ws;
  RC(readWorlds(s, ws));
  uint32_t n = ws[w].man_count;
  *count = n;
  if (!out || !cap || !n) return VXP_OK;
  std::vector<DevManifold> tmp(std::min(n, cap));
  HIP_TRY(hipMemcpy(tmp.data(), s->d_ml[s->cur].as<DevManifold>() + ws[w].man_start,
                    tmp.size() * sizeof(DevManifold), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < tmp.size(); i++) out[i] = tmp[i].m;
  return VXP_OK;
}

int vxp_pairs_get_w(vxp_system* s, uint32_t w, uint64_t* out, uint32_t cap, uint32_t* count) {
  if (!s || !count || w >= s->nw) return VXP_ERR_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(s->stream));
  std::vector<uint64_t> all(s->n_pairs);
  const DevBuf& src = treeMode(s) ? s->d_pairs_tree : s->d_pairs;  // the narrow phase's order
  if (s->n_pairs)
    HIP_TRY(hipMemcpy(all.data(), src.p, all.size() * 8, hipMemcpyDeviceToHost));
  uint32_t n = 0;
  for (uint64_t k : all) {
    if (keyWorld(k) != w) continue;
    if (out && n < cap) out[n] = exportKey(k);
    n++;
  }
  *count = n;
  return VXP_OK;
}

int vxp_islands_get_w(vxp_system* s, uint32_t w, uint32_t* ids, uint32_t* labels, uint32_t cap,
                      uint32_t* count) {
  if (!s || !count || w >= s->nw) return VXP_ERR_ARGUMENT;
  if (!s->cfg.enable_islands || !s->d_isl_timer.p) {
    *count = 0;
    return VXP_OK;
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  const uint32_t ic = s->isl_cap;
  std::vector<uint32_t> par(ic), end(ic);
  HIP_TRY(hipMemcpy(par.data(), s->d_isl_parent.as<uint32_t>() + (size_t)w * ic, ic * 4,
                    hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(end.data(), s->d_isl_end.as<uint32_t>() + (size_t)w * ic, ic * 4,
                    hipMemcpyDeviceToHost));
  uint32_t n = 0;
  for (uint32_t i = 0; i < ic; i++) {
    if (!end[i]) continue;
    if (ids && labels && n < cap) {
      ids[n] = i;
      labels[n] = par[i];
    }
    n++;
  }
  *count = n;
  return VXP_OK;
}

int vxp_island_stats_get_w(vxp_system* s, uint32_t w, vxp_island_stats* out) {
  if (!s || !out || w >= s->nw || !s->cfg.enable_islands) return VXP_ERR_ARGUMENT;
  std::vector<WorldState> ws;
  RC(readWorlds(s, ws));
  *out = ws[w].isl;
  return VXP_OK;
}

// Label of body `id` of world `w` in the last island build, or kNone when the
// body was no manifold endpoint (body_to_island, src/islands.zig:204-205).
static int islandOf(vxp_system* s, uint32_t w, uint32_t id, uint32_t* label) {
  *label = 0xFFFFFFFFu;
  if (!s->d_isl_timer.p || id >= s->isl_cap) return VXP_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  const size_t i = (size_t)w * s->isl_cap + id;
  uint32_t end = 0, par = 0;
  HIP_TRY(hipMemcpy(&end, s->d_isl_end.as<uint32_t>() + i, 4, hipMemcpyDeviceToHost));
  if (!end) return VXP_OK;
  HIP_TRY(hipMemcpy(&par, s->d_isl_parent.as<uint32_t>() + i, 4, hipMemcpyDeviceToHost));
  *label = par;
  return VXP_OK;
}

int vxp_body_in_sleeping_island_w(vxp_system* s, uint32_t w, vxp_handle h, int* out) {
  if (!s || !out || w >= s->nw) return VXP_ERR_ARGUMENT;
  *out = 0;
  RC(ensureHost(s));
  if (!validHandle(s, w, h)) return VXP_ERR_INVALID_HANDLE;
  if (!s->cfg.enable_islands) return VXP_OK;
  uint32_t label;
  RC(islandOf(s, w, s->id[(size_t)w * s->MB + h.index], &label));
  if (label == 0xFFFFFFFFu) return VXP_OK;
  uint32_t z = 0;
  HIP_TRY(hipMemcpy(&z, s->d_isl_sleep.as<uint32_t>() + (size_t)w * s->isl_cap + label, 4,
                    hipMemcpyDeviceToHost));
  *out = z ? 1 : 0;
  return VXP_OK;
}

int vxp_body_wake_island_w(vxp_system* s, uint32_t w, vxp_handle h) {
  if (!s || w >= s->nw) return VXP_ERR_ARGUMENT;
  RC(ensureHost(s));
  if (!validHandle(s, w, h)) return VXP_ERR_INVALID_HANDLE;
  if (s->cfg.enable_islands) {
    // wakeIsland (src/islands.zig:497-507): the island's flag and timer
    uint32_t label;
    RC(islandOf(s, w, s->id[(size_t)w * s->MB + h.index], &label));
    if (label != 0xFFFFFFFFu) {
      const uint32_t zero = 0;
      const size_t i = (size_t)w * s->isl_cap + label;
      HIP_TRY(hipMemcpy(s->d_isl_sleep.as<uint32_t>() + i, &zero, 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s->d_isl_timer.as<float>() + i, &zero, 4, hipMemcpyHostToDevice));
    }
    // needs_rebuild, set on both paths (src/islands.zig:506,510-516)
    HIP_TRY(hipStreamSynchronize(s->stream));
    const uint32_t one = 1;
    HIP_TRY(hipMemcpy((uint8_t*)(s->d_ws.as<WorldState>() + w) + offsetof(WorldState, isl_need), &one,
                      4, hipMemcpyHostToDevice));
  }
  return vxp_body_op_w(s, w, h, VXP_OP_WAKE_UP, nullptr, nullptr);
}

int vxp_tree_get_w(vxp_system* s, uint32_t w, uint32_t which, vxp_tree_node* out, uint32_t cap,
                   uint32_t* count, uint32_t* root) {
  if (!s || !count || !root || w >= s->nw || which > 1 || !treeMode(s)) return VXP_ERR_ARGUMENT;
  RC(treesToHost(s));
  const Tree t = hostTree(s, w, which);
  static_assert(sizeof(vxp_tree_node) == sizeof(TreeNode), "tree node export");
  *count = t.h->len;
  *root = t.h->root;
  if (out) std::memcpy(out, t.n, (size_t)std::min(cap, t.h->len) * sizeof(TreeNode));
  return VXP_OK;
}

int vxp_coloring_get_w(vxp_system* s, uint32_t w, uint8_t* colors, uint32_t cap, uint32_t* count) {
  if (!s || !count || w >= s->nw) return VXP_ERR_ARGUMENT;
  if (!s->cfg.enable_coloring || !s->d_ref_colors.p) {
    *count = 0;
    return VXP_OK;
  }
  std::vector<WorldState> ws;
  RC(readWorlds(s, ws));
  uint32_t n = ws[w].man_count;
  *count = n;
  if (colors && cap && n)
    HIP_TRY(hipMemcpy(colors, s->d_ref_colors.as<uint8_t>() + ws[w].man_start, std::min(n, cap),
                      hipMemcpyDeviceToHost));
  return VXP_OK;
}

}  // extern "C"
