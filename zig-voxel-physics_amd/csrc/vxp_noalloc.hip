// vxp_noalloc.hip — an allocation-free variant of vxp_collider_mass_properties
// for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops that admit new worlds between ticks, harnesses that
// forbid hipMalloc).  It has no counterpart in the reference.
//
// The staged bricks, the job record and the results live in one file-scope
// device array, `scratch`, of kScratchBytes per GPU.  Host code never passes
// the array's name to a kernel or a copy: each device's address is taken once
// with hipGetSymbolAddress (at system creation, with that device current) and
// kept in g_scratch[device].  Every call on a GPU shares that GPU's array and
// overwrites what it needs (bricks, job record, results are all written before
// they are read, nothing must start at zero); calls that overlap on one GPU —
// from two threads, or two systems' streams — race on it.
#include <mutex>

#include "vxp_host.hpp"
#include "vxp_kernels.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)4 << 20;  // 4 MiB per GPU
constexpr size_t kJobOff = 0;                      // one MassJob
constexpr size_t kOutOff = 256;                    // kMassOut floats
constexpr size_t kPosOff = 512;                    // 12 bytes per brick, then the voxel bytes
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices] = {};

inline size_t roundUp(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Bytes of scratch a collider of nb bricks needs.
inline size_t scratchNeed(size_t nb) { return roundUp(kPosOff + nb * 12, 256) + nb * 512; }

}  // namespace

namespace vxh {

// Device address of the current device's scratch array, taken once per device
// (called by vxp_system_create with the system's device current).
int noallocInit(int device) {
  if (device < 0 || device >= kMaxDevices) {
    g_err = "device index beyond the allocation-free entry point's table";
    return VXP_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return VXP_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return VXP_ERR_DEVICE;
  }
  g_scratch[device] = (uint8_t*)p;
  return VXP_OK;
}

}  // namespace vxh

extern "C" {

uint32_t vxp_noalloc_max_bricks(void) {
  size_t nb = (kScratchBytes - kPosOff) / (512 + 12);
  while (nb && scratchNeed(nb) > kScratchBytes) nb--;
  return (uint32_t)nb;
}

int vxp_collider_mass_properties_noalloc(vxp_system* s, uint32_t collider, float density,
                                         float out[13]) {
  if (!s || !out || collider >= s->colliders.size()) return VXP_ERR_ARGUMENT;
  const HostCollider& c = s->colliders[collider];
  const size_t nb = c.pos.size() / 3;
  if (nb > vxp_noalloc_max_bricks()) {
    g_err = "collider of " + std::to_string(nb) + " bricks exceeds the allocation-free limit of " +
            std::to_string(vxp_noalloc_max_bricks()) + " bricks (4 MiB of device scratch)";
    return VXP_ERR_CAPACITY;
  }
  if (c.solid == 0) {  // as computeMassProperties of an empty collider: zero
    for (int i = 0; i < 13; i++) out[i] = 0.0f;
    return VXP_OK;
  }
  uint8_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (s->device >= 0 && s->device < kMaxDevices) base = g_scratch[s->device];
  }
  if (!base) {
    g_err = "allocation-free scratch not set up for this system's device";
    return VXP_ERR_DEVICE;
  }
  MassJob job{};
  job.brick_off = 0;
  job.nb = (uint32_t)nb;
  job.solid = c.solid;
  job.vs = c.vs;
  job.density = density;
  int32_t* pos = (int32_t*)(base + kPosOff);
  uint8_t* vox = base + roundUp(kPosOff + nb * 12, 256);
  float* res = (float*)(base + kOutOff);
  HIP_TRY(hipMemcpyAsync(base + kJobOff, &job, sizeof(job), hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(pos, c.pos.data(), nb * 12, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(vox, c.vox.data(), nb * 512, hipMemcpyHostToDevice, s->stream));
  launch_collider_mass(s->stream, (const MassJob*)(base + kJobOff), 1, pos, vox, res);
  HIP_TRY(hipGetLastError());
  float r[kMassOut];
  HIP_TRY(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int i = 0; i < 13; i++) out[i] = r[i];
  return VXP_OK;
}

}  // extern "C"
