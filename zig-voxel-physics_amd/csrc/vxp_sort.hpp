// vxp_sort.hpp — stable radix ranking for wave64 workgroups, shared by the
// block sort of the sweep order (vxp_kernels.hip) and the device-wide sorts
// (vxp_prims.hip).
//
// Items are processed in chunks of one item per thread, in thread order (the
// "striped" arrangement: item i of a tile sits in chunk i / T, thread i % T),
// so ranking chunk after chunk keeps equal digits in input order (stable).
// Inside a chunk a wave finds the lanes that share its digit with one ballot
// per digit bit; the first such lane publishes the wave's count for the digit
// in LDS; the rank of an item is the digit's running count from earlier
// chunks + the counts of earlier waves + the earlier lanes of its own wave.
// The per-wave counts and the running counts are double-buffered between
// consecutive chunks, so a chunk needs one workgroup barrier.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vxd {

constexpr uint32_t kRadixBits = 8;  // digit width of the in-LDS block sort
constexpr uint32_t kRadixBins = 1u << kRadixBits;

__device__ __forceinline__ uint64_t laneMaskLt() {
  const uint32_t lane = threadIdx.x & 63u;
  return lane ? (~0ull >> (64u - lane)) : 0ull;
}

// Ranking state of one workgroup of kT threads over kBins digit values.  LDS
// it works in: wcnt [2][kT / 64][kBins], zero when a ranking starts (and left
// zero by finish()); run [2][kBins], run[0] = the start of every digit when a
// ranking starts.
template <uint32_t kT, uint32_t kBins = kRadixBins>
struct Ranker {
  static constexpr uint32_t kWaves = kT / 64;
  uint32_t* wcnt;
  uint32_t* run;
  uint32_t b = 0, last_digit = 0;
  bool last_leader = false;

  __device__ __forceinline__ Ranker(uint32_t* wcnt_, uint32_t* run_) : wcnt(wcnt_), run(run_) {}

  // Position of this thread's item among the items of its digit (the digit's
  // start included).  `valid` false for padding lanes (no rank).  Every
  // thread of the workgroup calls it.
  __device__ __forceinline__ uint32_t rank(uint32_t digit, bool valid, uint32_t bits) {
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    uint64_t m = __ballot(valid);
    for (uint32_t k = 0; k < bits; k++) {
      const bool on = (digit >> k) & 1u;
      const uint64_t bb = __ballot(on);
      m &= on ? bb : ~bb;
    }
    const uint32_t below = (uint32_t)__popcll(m & laneMaskLt());
    const bool leader = valid && below == 0;
    uint32_t* cur = wcnt + b * kWaves * kBins;
    uint32_t* prev = wcnt + (b ^ 1u) * kWaves * kBins;
    if (leader) cur[wave * kBins + digit] = (uint32_t)__popcll(m);
    __syncthreads();
    // the previous chunk's counts were read by every thread before the
    // barrier: this wave clears its own entry (it alone writes that row)
    if (last_leader) prev[wave * kBins + last_digit] = 0;
    const uint32_t* rc = run + b * kBins;
    uint32_t r = 0;
    if (valid) {
      r = rc[digit] + below;
      for (uint32_t w = 0; w < wave; w++) r += cur[w * kBins + digit];
    }
    uint32_t* rn = run + (b ^ 1u) * kBins;
    for (uint32_t d = tid; d < kBins; d += kT) {
      uint32_t c = rc[d];
#pragma unroll
      for (uint32_t w = 0; w < kWaves; w++) c += cur[w * kBins + d];
      rn[d] = c;
    }
    last_leader = leader;
    last_digit = digit;
    b ^= 1u;
    return r;
  }

  // After the last chunk: every count read, the buffers left zero.
  __device__ __forceinline__ void finish() {
    __syncthreads();
    if (last_leader) wcnt[(b ^ 1u) * kWaves * kBins + (threadIdx.x >> 6) * kBins + last_digit] = 0;
    last_leader = false;
    __syncthreads();
  }
};

// Exclusive scan of s[0..kRadixBins) in place by the first wave (four bins
// per lane, lane totals by shuffles); starts and ends with a barrier.
__device__ __forceinline__ void scanBins(uint32_t* s) {
  static_assert(kRadixBins == 256, "four bins per lane of one wave");
  const uint32_t tid = threadIdx.x;
  __syncthreads();
  if (tid < 64) {
    uint32_t v[4], t = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      v[k] = s[tid * 4 + k];
      t += v[k];
    }
    uint32_t incl = t;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)incl, d, 64);
      if (tid >= d) incl += u;
    }
    uint32_t acc = incl - t;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      s[tid * 4 + k] = acc;
      acc += v[k];
    }
  }
  __syncthreads();
}

// Stable sort of the first n (<= kT * kIPT) items held striped in registers
// (item i = k * kT + tid in keys[k], vals[k]) by key bits [0, key_bits),
// entirely in LDS, kRadixBits per pass.  On return keys[k] (and vals[k])
// hold item k * kT + tid of the sorted order (items past n undefined).
// LDS: s_k [kT * kIPT], s_v [kT * kIPT] when kVals, wcnt [2][kT / 64][kRadixBins]
// (zero), run [2][kRadixBins].
template <uint32_t kT, uint32_t kIPT, bool kVals, class K>
__device__ void blockRadixSort(K (&keys)[kIPT], uint32_t (&vals)[kIPT], uint32_t n,
                               uint32_t key_bits, K* s_k, uint32_t* s_v, uint32_t* wcnt,
                               uint32_t* run) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t shift = 0; shift < key_bits; shift += kRadixBits) {
    const uint32_t bits = min(kRadixBits, key_bits - shift);
    const uint32_t mask = (1u << bits) - 1u;
    for (uint32_t d = tid; d < kRadixBins; d += kT) run[d] = 0;
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kIPT; k++)
      if (k * kT + tid < n) atomicAdd(&run[(uint32_t)(keys[k] >> shift) & mask], 1u);
    scanBins(run);
    Ranker<kT> rk(wcnt, run);
    uint32_t pos[kIPT];
#pragma unroll
    for (uint32_t k = 0; k < kIPT; k++) {
      pos[k] = 0;
      if (k * kT < n)  // uniform: chunks past n hold padding only
        pos[k] = rk.rank((uint32_t)(keys[k] >> shift) & mask, k * kT + tid < n, bits);
    }
    rk.finish();
#pragma unroll
    for (uint32_t k = 0; k < kIPT; k++) {
      if (k * kT + tid < n) {
        s_k[pos[k]] = keys[k];
        if (kVals) s_v[pos[k]] = vals[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kIPT; k++) {
      if (k * kT + tid < n) {
        keys[k] = s_k[k * kT + tid];
        if (kVals) vals[k] = s_v[k * kT + tid];
      }
    }
    __syncthreads();
  }
}

}  // namespace vxd
