// vxp_prims.hip — device-wide exclusive scans and stable LSD radix sorts,
// written for wave64 gfx950 workgroups.  Used by the broad phase (sweep order
// of worlds past one block, packed pair keys) and the narrow-phase work
// expansion (scans of chunk counts and manifold flags).
//
// Scan: tiles of 16384 words, one 1024-thread workgroup each.  One tile: a
// single launch.  More: the tile sums, then every tile scanned again with its
// offset, the sum of the earlier tiles' sums, which the workgroup adds up
// itself (two launches, two reads of the data).
//
// Sort: LSD radix over the key bits the caller names, in the fewest passes of
// at most 9-bit digits, tiles of 2048 items per 256-thread workgroup.  The
// first pass's digit counts per tile come from a counting kernel, every later
// pass's from the previous pass's scatter (each item counted for the tile it
// lands in); a pass scans its digit-major table, then each tile ranks its
// items chunk by chunk (vxp_sort.hpp: stable) and writes them to their
// digit's place.  Keys and values ping-pong between the caller's arrays and the
// temporary storage; the result lands in the output arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "vxp_kernels.hpp"
#include "vxp_sort.hpp"

namespace vxd {

constexpr uint32_t kScanT = 1024, kScanIPT = 16, kScanTile = kScanT * kScanIPT;
// LDS index of word i of a tile: one pad word per 32, so that a thread's run
// of consecutive words and a wave's strided accesses spread over the banks
__device__ __forceinline__ uint32_t padIdx(uint32_t i) { return i + (i >> 5); }
constexpr uint32_t kSortT = 256, kSortIPT = 8, kSortTile = kSortT * kSortIPT;
constexpr uint32_t kDevMaxBits = 9, kDevBins = 1u << kDevMaxBits;  // digits of the device-wide sort

static inline uint32_t cdivU(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }
// The item count of a launch: `cap` (the grid's size), or, when a device word
// is given (launches whose count is known only on the device, as in a
// captured step), min(*n_dev + n_add, cap).
__device__ __forceinline__ uint32_t itemCount(uint32_t cap, const uint32_t* n_dev, uint32_t n_add) {
  return n_dev ? min(*n_dev + n_add, cap) : cap;
}
__device__ __forceinline__ uint32_t cdivD(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
// The words of a sort's digit table for a device count: bins x tiles of the
// actual items (the table scan's count).
__global__ void k_table_count(const uint32_t* n_dev, uint32_t cap, uint32_t* out) {
  if (threadIdx.x == 0) out[0] = kDevBins * cdivD(min(*n_dev, cap), kSortTile);
}
static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Exclusive scan of one value per thread over the block; *total = the sum.
__device__ __forceinline__ uint32_t blockExclusive(uint32_t v, uint32_t* s_wave, uint32_t* total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6, nw = blockDim.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)incl, d, 64);
    if (lane >= d) incl += u;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < nw; w++) {
    const uint32_t x = s_wave[w];
    before += w < wave ? x : 0u;
    all += x;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// Exclusive scan of each tile (coalesced in and out through LDS; each thread
// scans 16 consecutive words), plus the sum of sums[0..tile) when sums is
// given.  In
// place is allowed: a tile is read whole before it is written.
// Sum of each tile of `in` (zero past n).
__global__ __launch_bounds__(kScanT) void k_scan_sums(const uint32_t* __restrict__ in, uint32_t n,
                                                      uint32_t* __restrict__ sums,
                                                      const uint32_t* n_dev, uint32_t n_add) {
  __shared__ uint32_t s_wave[kScanT / 64];
  n = itemCount(n, n_dev, n_add);
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t v = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanIPT; k++) {
    const size_t i = base + k * kScanT + threadIdx.x;
    if (i < n) v += in[i];
  }
  uint32_t total;
  (void)blockExclusive(v, s_wave, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanT) void k_scan_tile(const uint32_t* in, uint32_t* out, uint32_t n,
                                                      const uint32_t* __restrict__ sums,
                                                      const uint32_t* n_dev, uint32_t n_add) {
  __shared__ uint32_t s_d[kScanTile + kScanTile / 32];
  __shared__ uint32_t s_wave[kScanT / 64];
  n = itemCount(n, n_dev, n_add);
  const uint32_t tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t offset = 0;
  if (sums) {  // the sum of the earlier tiles
    uint32_t part = 0;
    for (uint32_t t = tid; t < blockIdx.x; t += kScanT) part += sums[t];
    (void)blockExclusive(part, s_wave, &offset);
  }
#pragma unroll
  for (uint32_t k = 0; k < kScanIPT; k++) {
    const size_t i = base + k * kScanT + tid;
    s_d[padIdx(k * kScanT + tid)] = i < n ? in[i] : 0u;
  }
  __syncthreads();
  uint32_t v[kScanIPT], t = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanIPT; k++) {
    v[k] = s_d[padIdx(tid * kScanIPT + k)];
    t += v[k];
  }
  uint32_t total;
  uint32_t acc = blockExclusive(t, s_wave, &total) + offset;
#pragma unroll
  for (uint32_t k = 0; k < kScanIPT; k++) {
    s_d[padIdx(tid * kScanIPT + k)] = acc;
    acc += v[k];
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kScanIPT; k++) {
    const size_t i = base + k * kScanT + tid;
    if (i < n) out[i] = s_d[padIdx(k * kScanT + tid)];
  }
}

template <class K>
__device__ __forceinline__ uint32_t digitOf(K key, uint32_t shift, uint32_t mask) {
  return (uint32_t)(key >> shift) & mask;
}

// Digit counts of each tile for one pass (digit key bits [shift, shift +
// bits)), digit-major: hist[d * n_tiles + tile].
template <class K>
__global__ __launch_bounds__(kSortT) void k_radix_hist(const K* __restrict__ keys, uint32_t n,
                                                       uint32_t shift, uint32_t bits,
                                                       uint32_t n_tiles,
                                                       uint32_t* __restrict__ hist,
                                                       const uint32_t* n_dev) {
  __shared__ uint32_t s_h[kDevBins];
  n = itemCount(n, n_dev, 0);
  // with a device count the table holds the tiles of the actual items only
  // (digit-major over those tiles); the grid's further tiles have no column
  if (n_dev) {
    n_tiles = cdivD(n, kSortTile);
    if (blockIdx.x >= n_tiles) return;
  }
  const uint32_t tid = threadIdx.x, mask = (1u << bits) - 1u;
  for (uint32_t d = tid; d < kDevBins; d += kSortT) s_h[d] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll
  for (uint32_t k = 0; k < kSortIPT; k++) {
    const size_t i = base + k * kSortT + tid;
    if (i < n) atomicAdd(&s_h[digitOf(keys[i], shift, mask)], 1u);
  }
  __syncthreads();
  for (uint32_t d = tid; d < kDevBins; d += kSortT) hist[(size_t)d * n_tiles + blockIdx.x] = s_h[d];
}

// Every item of the tile to off[d * n_tiles + tile] + its stable rank among
// the tile's items of digit d.  (The next pass's per-tile digit counts are a
// pass of their own over the output: counting them here needs device-scope
// atomics that serialize whenever many tiles feed one destination tile.)
template <class K, bool kVals>
__global__ __launch_bounds__(kSortT) void k_radix_scatter(const K* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          uint32_t n, uint32_t shift, uint32_t bits,
                                                          uint32_t n_tiles,
                                                          const uint32_t* __restrict__ off,
                                                          const uint32_t* n_dev) {
  __shared__ uint32_t s_wcnt[2 * (kSortT / 64) * kDevBins];
  __shared__ uint32_t s_run[2 * kDevBins];
  const uint32_t tid = threadIdx.x, mask = (1u << bits) - 1u;
  n = itemCount(n, n_dev, 0);
  if (n_dev) {
    n_tiles = cdivD(n, kSortTile);
    if (blockIdx.x >= n_tiles) return;
  }
  for (uint32_t d = tid; d < 2 * (kSortT / 64) * kDevBins; d += kSortT) s_wcnt[d] = 0;
  for (uint32_t d = tid; d < kDevBins; d += kSortT)
    s_run[d] = off[(size_t)d * n_tiles + blockIdx.x];  // digit d's first output slot
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
  K key[kSortIPT];  // the whole tile requested before the first chunk is ranked
  uint32_t val[kSortIPT];
#pragma unroll
  for (uint32_t k = 0; k < kSortIPT; k++) {
    const size_t i = base + k * kSortT + tid;
    key[k] = i < n ? kin[i] : (K)0;
    val[k] = kVals && i < n ? vin[i] : 0u;
  }
  Ranker<kSortT, kDevBins> rk(s_wcnt, s_run);
#pragma unroll
  for (uint32_t k = 0; k < kSortIPT; k++) {
    if (base + (size_t)k * kSortT >= n) break;  // uniform over the block
    const bool valid = base + k * kSortT + tid < n;
    const uint32_t o = rk.rank(valid ? digitOf(key[k], shift, mask) : 0u, valid, bits);
    if (valid) {
      kout[o] = key[k];
      if (kVals) vout[o] = val[k];
    }
  }
}

static size_t scanTmpBytes(uint32_t n) { return align256((size_t)cdivU(n, kScanTile) * 4); }

// n: the item count, or with n_dev the capacity the launch is sized for.
static hipError_t scanImpl(void* tmp, const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t s,
                           const uint32_t* n_dev = nullptr, uint32_t n_add = 0) {
  if (n == 0) return hipSuccess;
  const uint32_t nb = cdivU(n, kScanTile);
  if (nb == 1) {
    k_scan_tile<<<1, kScanT, 0, s>>>(in, out, n, nullptr, n_dev, n_add);
    return hipGetLastError();
  }
  uint32_t* sums = (uint32_t*)tmp;
  k_scan_sums<<<nb, kScanT, 0, s>>>(in, n, sums, n_dev, n_add);
  k_scan_tile<<<nb, kScanT, 0, s>>>(in, out, n, sums, n_dev, n_add);
  return hipGetLastError();
}

// Passes of a sort over `end_bit` key bits: the fewest of at most kDevMaxBits
// each, their widths balanced.
static uint32_t sortPasses(uint32_t end_bit) {
  return std::max(1u, (end_bit + kDevMaxBits - 1) / kDevMaxBits);
}

// Temporary storage of a sort: the digit table and its scan storage, then
// the ping-pong key and value arrays.
template <class K>
static size_t sortTmpBytes(uint32_t n, uint32_t end_bit, bool vals) {
  (void)end_bit;
  const uint32_t tiles = std::max(1u, cdivU(n, kSortTile));
  const size_t table = (size_t)kDevBins * tiles;
  return align256(table * 4) + scanTmpBytes((uint32_t)table) +
         align256((size_t)n * sizeof(K)) + (vals ? align256((size_t)n * 4) : 0) + 256;
}

template <class K, bool kVals>
static hipError_t sortImpl(void* tmp, size_t bytes, const K* kin, K* kout, const uint32_t* vin,
                           uint32_t* vout, uint32_t n, uint32_t end_bit, hipStream_t s,
                           const uint32_t* n_dev = nullptr) {
  if (bytes < sortTmpBytes<K>(n, end_bit, kVals)) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  const uint32_t tiles = cdivU(n, kSortTile);
  const uint32_t table = kDevBins * tiles;
  const uint32_t passes = sortPasses(end_bit);
  const uint32_t width = std::max(1u, (end_bit + passes - 1) / passes);
  char* p = (char*)tmp;
  uint32_t* hist = (uint32_t*)p;  // the digit table (`table` words), reused by every pass
  p += align256((size_t)table * 4);
  void* scan_tmp = p;
  p += scanTmpBytes(table);
  K* kalt = (K*)p;
  p += align256((size_t)n * sizeof(K));
  uint32_t* valt = kVals ? (uint32_t*)p : nullptr;
  p += kVals ? align256((size_t)n * 4) : 0;
  uint32_t* table_n = nullptr;  // the digit table's words for a device count
  if (n_dev) {
    table_n = (uint32_t*)p;
    k_table_count<<<1, 64, 0, s>>>(n_dev, n, table_n);
  }
  // the last pass must write kout: start from kout or the spare array so that
  // the passes alternate into it
  K* ka = (passes & 1) ? kout : kalt;
  K* kb = (passes & 1) ? kalt : kout;
  uint32_t* va = (passes & 1) ? vout : valt;
  uint32_t* vb = (passes & 1) ? valt : vout;
  const K* src = kin;
  const uint32_t* vsrc = vin;
  auto bitsOf = [&](uint32_t pass) {
    const uint32_t sh = pass * width;
    return sh >= end_bit ? 1u : std::min(width, end_bit - sh);
  };
  for (uint32_t pass = 0; pass < passes; pass++) {
    k_radix_hist<K><<<tiles, kSortT, 0, s>>>(src, n, pass * width, bitsOf(pass), tiles, hist,
                                             n_dev);
    hipError_t e = scanImpl(scan_tmp, hist, hist, table, s, table_n, 0);
    if (e != hipSuccess) return e;
    K* dst = (pass & 1) ? kb : ka;
    uint32_t* vdst = (pass & 1) ? vb : va;
    k_radix_scatter<K, kVals><<<tiles, kSortT, 0, s>>>(src, vsrc, dst, vdst, n, pass * width,
                                                       bitsOf(pass), tiles, hist, n_dev);
    src = dst;
    vsrc = vdst;
  }
  return hipGetLastError();
}

// Sorts on key bits [0, end_bit): the callers' keys are zero above it.
size_t sort_pairs_u64_u32_bytes(uint32_t n, uint32_t end_bit) {
  return sortTmpBytes<uint64_t>(n, end_bit, true);
}
hipError_t sort_pairs_u64_u32(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout,
                              const uint32_t* vin, uint32_t* vout, uint32_t n, uint32_t end_bit,
                              hipStream_t s) {
  return sortImpl<uint64_t, true>(tmp, bytes, kin, kout, vin, vout, n, end_bit, s);
}
size_t sort_keys_u64_bytes(uint32_t n, uint32_t end_bit) {
  return sortTmpBytes<uint64_t>(n, end_bit, false);
}
hipError_t sort_keys_u64(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout, uint32_t n,
                         uint32_t end_bit, hipStream_t s) {
  return sortImpl<uint64_t, false>(tmp, bytes, kin, kout, nullptr, nullptr, n, end_bit, s);
}
size_t sort_keys_u32_bytes(uint32_t n, uint32_t end_bit) {
  return sortTmpBytes<uint32_t>(n, end_bit, false);
}
hipError_t sort_keys_u32(void* tmp, size_t bytes, const uint32_t* kin, uint32_t* kout, uint32_t n,
                         uint32_t end_bit, hipStream_t s) {
  return sortImpl<uint32_t, false>(tmp, bytes, kin, kout, nullptr, nullptr, n, end_bit, s);
}
size_t exclusive_scan_u32_bytes(uint32_t n) { return scanTmpBytes(n); }
// The device-count forms: sized for `cap` items, they act on min(*n_dev
// (+ n_add), cap) items, so a launch sequence can be recorded before the
// counts exist.
hipError_t exclusive_scan_u32_n(void* tmp, const uint32_t* in, uint32_t* out, uint32_t cap,
                                const uint32_t* n_dev, uint32_t n_add, hipStream_t s) {
  return scanImpl(tmp, in, out, cap, s, n_dev, n_add);
}
hipError_t sort_keys_u64_n(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout,
                           uint32_t cap, const uint32_t* n_dev, uint32_t end_bit, hipStream_t s) {
  return sortImpl<uint64_t, false>(tmp, bytes, kin, kout, nullptr, nullptr, cap, end_bit, s, n_dev);
}
hipError_t sort_keys_u32_n(void* tmp, size_t bytes, const uint32_t* kin, uint32_t* kout,
                           uint32_t cap, const uint32_t* n_dev, uint32_t end_bit, hipStream_t s) {
  return sortImpl<uint32_t, false>(tmp, bytes, kin, kout, nullptr, nullptr, cap, end_bit, s, n_dev);
}
hipError_t exclusive_scan_u32(void* tmp, size_t /*bytes*/, const uint32_t* in, uint32_t* out,
                              uint32_t n, hipStream_t s) {
  return scanImpl(tmp, in, out, n, s);
}

}  // namespace vxd
