// vxp_prims.hip — device-wide sort and scan primitives (rocPRIM) used by the
// broad phase (sort-and-sweep order, pair keys) and the narrow-phase work
// expansion (exclusive scans of chunk counts and manifold flags).
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/rocprim.hpp>

#include "vxp_kernels.hpp"

namespace vxd {

// Sorts on key bits [0, end_bit): the callers' keys are zero above it.
size_t sort_pairs_u64_u32_bytes(uint32_t n, uint32_t end_bit) {
  size_t b = 0;
  (void)rocprim::radix_sort_pairs(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                  (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0, end_bit);
  return b;
}
hipError_t sort_pairs_u64_u32(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout,
                              const uint32_t* vin, uint32_t* vout, uint32_t n, uint32_t end_bit,
                              hipStream_t s) {
  return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, s);
}
size_t sort_keys_u64_bytes(uint32_t n, uint32_t end_bit) {
  size_t b = 0;
  (void)rocprim::radix_sort_keys(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr, n, 0,
                                 end_bit);
  return b;
}
hipError_t sort_keys_u64(void* tmp, size_t bytes, const uint64_t* kin, uint64_t* kout, uint32_t n,
                         uint32_t end_bit, hipStream_t s) {
  return rocprim::radix_sort_keys(tmp, bytes, kin, kout, n, 0, end_bit, s);
}
size_t sort_keys_u32_bytes(uint32_t n, uint32_t end_bit) {
  size_t b = 0;
  (void)rocprim::radix_sort_keys(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0,
                                 end_bit);
  return b;
}
hipError_t sort_keys_u32(void* tmp, size_t bytes, const uint32_t* kin, uint32_t* kout, uint32_t n,
                         uint32_t end_bit, hipStream_t s) {
  return rocprim::radix_sort_keys(tmp, bytes, kin, kout, n, 0, end_bit, s);
}
size_t exclusive_scan_u32_bytes(uint32_t n) {
  size_t b = 0;
  (void)rocprim::exclusive_scan(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr, 0u, n,
                          rocprim::plus<uint32_t>());
  return b;
}
hipError_t exclusive_scan_u32(void* tmp, size_t bytes, const uint32_t* in, uint32_t* out,
                              uint32_t n, hipStream_t s) {
  return rocprim::exclusive_scan(tmp, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
}

}  // namespace vxd
