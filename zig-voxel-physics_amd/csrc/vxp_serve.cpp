// vxp_serve.cpp — continuous batching over resident worlds (vxp_server_* in
// include/vxp.h).
//
// Two systems with the same layout: the pool holds every resident world in
// device memory and is never stepped; the batch holds `batch_slots` worlds
// and is stepped one update(fixed_timestep) frame per tick for all slots at
// once.  Admitting a request moves its world pool -> free slot, retiring it
// moves the world back; each move is one descriptor list executed by a single
// copy kernel (body slot arrays up to the world's high-water slot, id map,
// counters, brick-pair cache tables) plus the world's manifolds, which are
// appended to the batch's current manifold list and re-keyed to the slot.
// Every per-world quantity that survives a step travels with the world, so a
// world advanced in the mix ends bit-identical to the same world advanced alone.
#include <time.h>

#include <deque>

#include "vxp_host.hpp"

namespace {

double nowS() {
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

struct Req {
  uint64_t tag;
  uint32_t world, frames, left;
  double submit_s, admit_s;
  uint64_t body_substeps;
};

struct Resident {
  bool in_slot = false;
  uint32_t frame = 0;   // BrickPairCache frame counter
  uint32_t mcount = 0;  // parked manifolds
  uint32_t hw = 0;      // body slots in use (high-water mark)
  DevBuf ml;            // parked manifolds (DevManifold)
  uint32_t isl_need = 1;  // island manager bookkeeping while parked
  vxp_island_stats isl{};
};

// Per-body-slot device arrays of a system and their bytes per slot.
struct SlotArray {
  DevBuf vxp_system::*buf;
  size_t elem;
};
const SlotArray kSlotArrays[] = {
    {&vxp_system::b_pos, 16},     {&vxp_system::b_rot, 16},        {&vxp_system::b_lv, 16},
    {&vxp_system::b_av, 16},      {&vxp_system::b_fat_lo, 16},     {&vxp_system::b_fat_hi, 16},
    {&vxp_system::b_inv_mass, 4}, {&vxp_system::b_iil, 36},        {&vxp_system::b_fric, 4},
    {&vxp_system::b_rest, 4},     {&vxp_system::b_ldamp, 4},       {&vxp_system::b_adamp, 4},
    {&vxp_system::b_sleep_timer, 4}, {&vxp_system::b_type, 1},     {&vxp_system::b_sleeping, 1},
    {&vxp_system::b_active, 1},   {&vxp_system::b_collider, 4},    {&vxp_system::b_id, 4}};

size_t slotBytes() {
  size_t b = 0;
  for (const SlotArray& a : kSlotArrays) b += a.elem;
  return b;
}

uint32_t cacheCap(const vxp_config& c) {
  return c.brick_pair_cache_size ? c.brick_pair_cache_size : 1024;
}

}  // namespace

struct vxp_server {
  vxp_system* pool = nullptr;
  vxp_system* batch = nullptr;
  uint32_t capacity = 0, S = 0, n_worlds = 0;
  uint64_t budget = 0, per_world = 0, batch_bytes = 0;
  std::vector<Resident> res;
  std::deque<Req> queue;
  std::vector<Req> slot;
  std::vector<uint8_t> busy;
  std::deque<vxp_request_done> done;
  uint64_t ticks = 0, completed = 0, body_substeps = 0;
  // pinned staging (copy descriptors, world-state records, slot lists) and
  // its device copy
  uint8_t* h_stage = nullptr;
  size_t h_cap = 0;
  DevBuf d_stage;
  WorldState* h_ws = nullptr;  // pinned per-slot reads after each tick
  vxp_stats* h_st = nullptr;
};

namespace {

// Descriptors moving world `wf` of `from` onto world `wt` of `to`.
void worldCopies(std::vector<CopyDesc>& d, vxp_system* from, uint32_t wf, vxp_system* to,
                 uint32_t wt, uint32_t hw) {
  const size_t MB = from->MB;
  for (const SlotArray& a : kSlotArrays)
    d.push_back({(from->*a.buf).as<uint8_t>() + wf * MB * a.elem,
                 (to->*a.buf).as<uint8_t>() + wt * MB * a.elem, (uint64_t)hw * a.elem});
  const size_t ic = from->id_cap;
  d.push_back({from->b_idmap.as<int32_t>() + wf * ic, to->b_idmap.as<int32_t>() + wt * ic, ic * 4});
  d.push_back({from->d_stats.as<vxp_stats>() + wf, to->d_stats.as<vxp_stats>() + wt,
               sizeof(vxp_stats)});
  const size_t cap = cacheCap(from->cfg);
  d.push_back({from->d_cache_keys.as<CacheKeyRec>() + wf * cap,
               to->d_cache_keys.as<CacheKeyRec>() + wt * cap, cap * sizeof(CacheKeyRec)});
  d.push_back({from->d_cache_vals.as<CacheVal>() + wf * cap,
               to->d_cache_vals.as<CacheVal>() + wt * cap, cap * sizeof(CacheVal)});
  if (from->cfg.enable_islands)  // labels, endpoint flags and island sleep state
    for (DevBuf vxp_system::*a : {&vxp_system::d_isl_parent, &vxp_system::d_isl_end,
                                  &vxp_system::d_isl_timer, &vxp_system::d_isl_sleep})
      d.push_back({(from->*a).as<uint8_t>() + wf * ic * 4, (to->*a).as<uint8_t>() + wt * ic * 4,
                   ic * 4});
}

// Uploads descriptors + world-state records + extra u32 words through the
// pinned staging buffer and runs the copies.  World-state record i goes to
// ws_dst[i]; the u32 words are returned as a device pointer for the caller.
int runCopies(vxp_server* v, std::vector<CopyDesc>& d, const std::vector<WorldState>& wsrec,
              const std::vector<WorldState*>& ws_dst, const std::vector<uint32_t>& words,
              const uint32_t** dev_words) {
  vxp_system* b = v->batch;
  const size_t nd = d.size() + wsrec.size();
  const size_t off_ws = nd * sizeof(CopyDesc);
  const size_t off_w = off_ws + wsrec.size() * sizeof(WorldState);
  const size_t bytes = off_w + words.size() * 4 + 16;
  RC(streamWait(b));  // the staging buffer is reused
  if (bytes > v->h_cap) {
    if (v->h_stage) (void)hipHostFree(v->h_stage);
    v->h_stage = nullptr;
    size_t nc = std::max(bytes, v->h_cap * 2);
    HIP_TRY(hipHostMalloc((void**)&v->h_stage, nc, hipHostMallocDefault));
    v->h_cap = nc;
  }
  HIP_TRY(v->d_stage.ensure(bytes));
  uint8_t* dbase = v->d_stage.as<uint8_t>();
  for (size_t i = 0; i < wsrec.size(); i++)
    d.push_back({dbase + off_ws + i * sizeof(WorldState), ws_dst[i], sizeof(WorldState)});
  std::memcpy(v->h_stage, d.data(), d.size() * sizeof(CopyDesc));
  if (!wsrec.empty())
    std::memcpy(v->h_stage + off_ws, wsrec.data(), wsrec.size() * sizeof(WorldState));
  if (!words.empty()) std::memcpy(v->h_stage + off_w, words.data(), words.size() * 4);
  HIP_TRY(hipMemcpyAsync(dbase, v->h_stage, bytes, hipMemcpyHostToDevice, b->stream));
  launch_copy_many(b->stream, (const CopyDesc*)dbase, (uint32_t)d.size());
  if (dev_words) *dev_words = (const uint32_t*)(dbase + off_w);
  HIP_TRY(hipGetLastError());
  return VXP_OK;
}

int syncBatch(vxp_server* v) { return streamWait(v->batch); }

// Pool -> slots for the admitted (world, slot) pairs.
int admit(vxp_server* v, const std::vector<std::pair<uint32_t, uint32_t>>& moves) {
  if (moves.empty()) return VXP_OK;
  vxp_system* b = v->batch;
  std::vector<CopyDesc> d;
  std::vector<WorldState> wsrec;
  std::vector<WorldState*> wsdst;
  std::vector<uint32_t> rekey;  // (begin, count, slot) triples
  RC(settleManifolds(b));
  uint32_t total = b->n_manifolds;
  size_t add = 0;
  for (auto& m : moves) add += v->res[m.first].mcount;
  HIP_TRY(b->d_ml[b->cur].ensure((total + add + 1) * sizeof(DevManifold), true, b->stream));
  DevManifold* ml = b->d_ml[b->cur].as<DevManifold>();
  for (auto& m : moves) {
    Resident& r = v->res[m.first];
    worldCopies(d, v->pool, m.first, b, m.second, r.hw);
    if (r.mcount) {
      d.push_back({r.ml.p, ml + total, (uint64_t)r.mcount * sizeof(DevManifold)});
      rekey.insert(rekey.end(), {total, r.mcount, m.second});
    }
    WorldState w;
    std::memset(&w, 0, sizeof(w));
    w.man_start = total;
    w.man_count = r.mcount;
    w.frame = r.frame;
    w.isl_need = r.isl_need;
    w.isl = r.isl;
    wsrec.push_back(w);
    wsdst.push_back(b->d_ws.as<WorldState>() + m.second);
    total += r.mcount;
    r.in_slot = true;
  }
  b->n_manifolds = total;
  const uint32_t* dwords = nullptr;
  RC(runCopies(v, d, wsrec, wsdst, rekey, &dwords));
  launch_rekey_manifolds(b->stream, ml, dwords, (uint32_t)(rekey.size() / 3));
  HIP_TRY(hipGetLastError());
  return VXP_OK;
}

// Slots -> pool for the retired (world, slot) pairs; the slots are emptied.
int retire(vxp_server* v, const std::vector<std::pair<uint32_t, uint32_t>>& moves) {
  if (moves.empty()) return VXP_OK;
  vxp_system* b = v->batch;
  std::vector<CopyDesc> d;
  std::vector<WorldState> wsrec;
  std::vector<WorldState*> wsdst;
  const DevManifold* ml = b->d_ml[b->cur].as<DevManifold>();
  for (auto& m : moves) {
    Resident& r = v->res[m.first];
    const WorldState& w = v->h_ws[m.second];
    worldCopies(d, b, m.second, v->pool, m.first, r.hw);
    r.frame = w.frame;
    r.mcount = w.man_count;
    r.isl_need = w.isl_need;
    r.isl = w.isl;
    if (r.mcount) {
      HIP_TRY(r.ml.ensure((size_t)r.mcount * sizeof(DevManifold)));
      d.push_back({ml + w.man_start, r.ml.p, (uint64_t)r.mcount * sizeof(DevManifold)});
    }
    WorldState z;
    std::memset(&z, 0, sizeof(z));
    wsrec.push_back(z);
    wsdst.push_back(b->d_ws.as<WorldState>() + m.second);
    r.in_slot = false;
  }
  std::vector<uint32_t> slots;
  for (auto& m : moves) slots.push_back(m.second);
  const uint32_t* dslots = nullptr;
  RC(runCopies(v, d, wsrec, wsdst, slots, &dslots));
  launch_clear_slots(b->stream, devCfg(b), devBodies(b), dslots, (uint32_t)slots.size());
  HIP_TRY(hipGetLastError());
  v->pool->dev_newer = true;
  return VXP_OK;
}

}  // namespace

extern "C" {

int vxp_server_create(const vxp_config* cfg, uint32_t max_worlds, uint32_t batch_slots,
                      uint64_t hbm_budget_bytes, vxp_server** out) {
  if (!cfg || !out || !max_worlds || !batch_slots) return VXP_ERR_ARGUMENT;
  if (cfg->pair_order != VXP_PAIR_ORDER_KEY) {
    g_err = "the server keeps no broad-phase trees: pair_order must be VXP_PAIR_ORDER_KEY";
    return VXP_ERR_ARGUMENT;
  }
  const uint64_t MB = cfg->max_bodies, ic = MB + 2, cap = cacheCap(*cfg);
  // resident bytes of one world: slot arrays, id map, counters, cache tables
  // and a reserve of two manifolds per body slot
  const uint64_t per_world = MB * slotBytes() + ic * 4 + sizeof(vxp_stats) +
                             cap * (sizeof(CacheKeyRec) + sizeof(CacheVal)) +
                             2 * MB * sizeof(DevManifold);
  // the batch also holds per-step buffers: budget them as many slot worlds again
  const uint64_t batch_bytes = 2 * (uint64_t)batch_slots * per_world;
  if (hbm_budget_bytes <= batch_bytes + per_world) {
    g_err = "HBM budget below one batch plus one resident world";
    return VXP_ERR_OOM;
  }
  const uint64_t fit = (hbm_budget_bytes - batch_bytes) / per_world;
  auto* v = new vxp_server();
  v->capacity = (uint32_t)std::min<uint64_t>(max_worlds, fit);
  v->S = batch_slots;
  v->budget = hbm_budget_bytes;
  v->per_world = per_world;
  v->batch_bytes = batch_bytes;
  int rc;
  if ((rc = vxp_system_create(cfg, v->capacity, &v->pool)) ||
      (rc = vxp_system_create(cfg, batch_slots, &v->batch))) {
    if (v->pool) vxp_system_destroy(v->pool);
    delete v;
    return rc;
  }
  // the batch's (empty) host mirror goes up once; from here on the device is
  // authoritative for the batch and only colliders are uploaded again
  if ((rc = ensureDevice(v->batch))) {
    vxp_system_destroy(v->pool);
    vxp_system_destroy(v->batch);
    delete v;
    return rc;
  }
  if (hipHostMalloc((void**)&v->h_ws, batch_slots * sizeof(WorldState), hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&v->h_st, batch_slots * sizeof(vxp_stats), hipHostMallocDefault) !=
          hipSuccess) {
    g_err = "pinned host allocation failed";
    vxp_server_destroy(v);
    return VXP_ERR_OOM;
  }
  v->res.resize(v->capacity);
  v->slot.resize(batch_slots);
  v->busy.assign(batch_slots, 0);
  *out = v;
  return VXP_OK;
}

int vxp_server_destroy(vxp_server* v) {
  if (!v) return VXP_ERR_ARGUMENT;
  (void)hipStreamSynchronize(v->batch->stream);
  for (Resident& r : v->res) r.ml.release();
  v->d_stage.release();
  if (v->h_stage) (void)hipHostFree(v->h_stage);
  if (v->h_ws) (void)hipHostFree(v->h_ws);
  if (v->h_st) (void)hipHostFree(v->h_st);
  vxp_system_destroy(v->pool);
  vxp_system_destroy(v->batch);
  delete v;
  return VXP_OK;
}

int vxp_server_collider_create(vxp_server* v, float voxel_size, uint32_t n, const int32_t* bp,
                               const uint8_t* vox, uint32_t* out) {
  if (!v || !out) return VXP_ERR_ARGUMENT;
  uint32_t a = 0, b = 0;
  RC(vxp_collider_create(v->pool, voxel_size, n, bp, vox, &a));
  RC(vxp_collider_create(v->batch, voxel_size, n, bp, vox, &b));
  *out = a;
  return VXP_OK;
}

int vxp_server_world_create(vxp_server* v, uint32_t* out) {
  if (!v || !out) return VXP_ERR_ARGUMENT;
  if (v->n_worlds >= v->capacity) {
    g_err = "HBM budget exhausted: no resident world slot left";
    return VXP_ERR_OOM;
  }
  *out = v->n_worlds++;
  return VXP_OK;
}

int vxp_server_body_create(vxp_server* v, uint32_t w, uint32_t type, const float pos[3],
                           const float rot[4], uint32_t col, float density, vxp_handle* out) {
  if (!v || w >= v->n_worlds || !out) return VXP_ERR_ARGUMENT;
  if (v->res[w].in_slot) {
    g_err = "world is being stepped";
    return VXP_ERR_ARGUMENT;
  }
  RC(syncBatch(v));
  RC(vxp_body_create_w(v->pool, w, type, pos, rot, col, density, out));
  if (v->pool->id_cap != v->batch->id_cap) {
    g_err = "body ids exceed the id map of a batch slot";
    return VXP_ERR_CAPACITY;
  }
  v->res[w].hw = std::max(v->res[w].hw, out->index + 1);
  return VXP_OK;
}

int vxp_server_submit(vxp_server* v, uint32_t w, uint32_t frames, uint64_t tag, double submit_s) {
  if (!v || w >= v->n_worlds || frames == 0) return VXP_ERR_ARGUMENT;
  Req r{tag, w, frames, frames, submit_s < 0.0 ? nowS() : submit_s, 0.0, 0};
  v->queue.push_back(r);
  return VXP_OK;
}

int vxp_server_graph_enable(vxp_server* v, int on) {
  if (!v) return VXP_ERR_ARGUMENT;
  return vxp_graph_enable(v->batch, on);
}

int vxp_server_graph_info_get(vxp_server* v, vxp_graph_info* out) {
  if (!v) return VXP_ERR_ARGUMENT;
  return vxp_graph_info_get(v->batch, out);
}

int vxp_server_tick(vxp_server* v, uint32_t* n_done) {
  if (!v) return VXP_ERR_ARGUMENT;
  if (n_done) *n_done = 0;
  vxp_system* b = v->batch;
  RC(ensureDevice(v->pool));
  if (b->cfg.enable_islands) {
    RC(ensureIslands(v->pool));
    RC(ensureIslands(b));
  }
  // ---- admission: FIFO, one request per world in flight ----------------------
  std::vector<std::pair<uint32_t, uint32_t>> in;
  {
    std::vector<uint32_t> free_slots;
    for (uint32_t s = v->S; s-- > 0;)
      if (!v->busy[s]) free_slots.push_back(s);
    std::vector<uint8_t> held(v->n_worlds, 0);  // a request of this world waits ahead
    const double t = nowS();
    for (auto it = v->queue.begin(); it != v->queue.end() && !free_slots.empty();) {
      if (v->res[it->world].in_slot || held[it->world]) {
        held[it->world] = 1;
        ++it;
        continue;
      }
      const uint32_t s = free_slots.back();
      free_slots.pop_back();
      Req r = *it;
      r.admit_s = t;
      v->slot[s] = r;
      v->busy[s] = 1;
      held[r.world] = 1;
      in.push_back({r.world, s});
      it = v->queue.erase(it);
    }
  }
  RC(admit(v, in));
  bool any = false;
  for (uint32_t s = 0; s < v->S; s++) any |= v->busy[s] != 0;
  if (!any) return syncBatch(v);
  // ---- one frame for every slot -------------------------------------------------
  RC(updateOnce(b, b->cfg.fixed_timestep));
  v->ticks++;
  HIP_TRY(hipMemcpyAsync(v->h_ws, b->d_ws.p, v->S * sizeof(WorldState), hipMemcpyDeviceToHost,
                         b->stream));
  HIP_TRY(hipMemcpyAsync(v->h_st, b->d_stats.p, v->S * sizeof(vxp_stats), hipMemcpyDeviceToHost,
                         b->stream));
  RC(streamWait(b));
  for (uint32_t s = 0; s < v->S; s++)
    if (v->h_ws[s].error) {
      g_err = "solver colouring exceeded its colour limit";
      return VXP_ERR_CAPACITY;
    }
  // ---- retirement -----------------------------------------------------------------
  std::vector<std::pair<uint32_t, uint32_t>> outm;
  std::vector<uint32_t> outs;
  for (uint32_t s = 0; s < v->S; s++) {
    if (!v->busy[s]) continue;
    Req& r = v->slot[s];
    uint64_t bs = (uint64_t)v->h_st[s].awake_bodies_stepped * b->cfg.substeps;
    r.body_substeps += bs;
    v->body_substeps += bs;
    if (--r.left == 0) {
      outm.push_back({r.world, s});
      outs.push_back(s);
    }
  }
  RC(retire(v, outm));
  RC(syncBatch(v));
  const double t = nowS();
  for (uint32_t s : outs) {
    const Req& r = v->slot[s];
    vxp_request_done dn;
    dn.tag = r.tag;
    dn.world = r.world;
    dn.frames = r.frames;
    dn.submit_s = r.submit_s;
    dn.admit_s = r.admit_s;
    dn.done_s = t;
    dn.body_substeps = r.body_substeps;
    v->done.push_back(dn);
    v->busy[s] = 0;
    v->completed++;
  }
  if (n_done) *n_done = (uint32_t)outs.size();
  return VXP_OK;
}

int vxp_server_poll(vxp_server* v, vxp_request_done* out, uint32_t cap, uint32_t* count) {
  if (!v || !count) return VXP_ERR_ARGUMENT;
  uint32_t n = 0;
  while (n < cap && out && !v->done.empty()) {
    out[n++] = v->done.front();
    v->done.pop_front();
  }
  *count = n;
  return VXP_OK;
}

int vxp_server_bodies_get(vxp_server* v, uint32_t w, vxp_body_state* out, uint32_t cap,
                          uint32_t* count) {
  if (!v || w >= v->n_worlds) return VXP_ERR_ARGUMENT;
  if (v->res[w].in_slot) {
    g_err = "world is being stepped";
    return VXP_ERR_ARGUMENT;
  }
  RC(syncBatch(v));
  return vxp_bodies_get_w(v->pool, w, out, cap, count);
}

int vxp_server_stats_get(vxp_server* v, uint32_t w, vxp_stats* out) {
  if (!v || w >= v->n_worlds) return VXP_ERR_ARGUMENT;
  if (v->res[w].in_slot) {
    g_err = "world is being stepped";
    return VXP_ERR_ARGUMENT;
  }
  RC(syncBatch(v));
  return vxp_stats_get_w(v->pool, w, out);
}

int vxp_server_manifolds_get(vxp_server* v, uint32_t w, vxp_manifold* out, uint32_t cap,
                             uint32_t* count) {
  if (!v || w >= v->n_worlds || !count) return VXP_ERR_ARGUMENT;
  const Resident& r = v->res[w];
  if (r.in_slot) {
    g_err = "world is being stepped";
    return VXP_ERR_ARGUMENT;
  }
  RC(syncBatch(v));
  *count = r.mcount;
  if (!out || !cap || !r.mcount) return VXP_OK;
  std::vector<DevManifold> tmp(std::min(r.mcount, cap));
  HIP_TRY(hipMemcpy(tmp.data(), r.ml.p, tmp.size() * sizeof(DevManifold), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < tmp.size(); i++) out[i] = tmp[i].m;
  return VXP_OK;
}

int vxp_server_info_get(vxp_server* v, vxp_server_info* out) {
  if (!v || !out) return VXP_ERR_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  out->worlds = v->n_worlds;
  out->capacity = v->capacity;
  out->slots = v->S;
  for (uint32_t s = 0; s < v->S; s++) out->occupied += v->busy[s];
  out->queued = (uint32_t)v->queue.size();
  out->ticks = v->ticks;
  out->completed = v->completed;
  out->body_substeps = v->body_substeps;
  uint64_t arenas = 0;
  for (const Resident& r : v->res) arenas += r.ml.bytes;
  out->resident_bytes = (uint64_t)v->capacity * v->per_world + v->batch_bytes + arenas;
  out->budget_bytes = v->budget;
  return VXP_OK;
}

}  // extern "C"
