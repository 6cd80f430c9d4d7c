// Device helpers shared by the stiffness kernels: raw buffer loads / stores
// through an SGPR resource (base pointer), one 32-bit VGPR byte offset and an
// SGPR byte offset, so that the many plane-strided loads of one thread share
// a single address register.  Extents behind one resource stay below 2 GiB
// (checked on the host when a problem is created).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mjw {

typedef __amdgpu_buffer_rsrc_t rsrc_t;

// raw (unstructured) buffer, no swizzle; dword3 as used for CDNA buffers
__device__ __forceinline__ rsrc_t mkres(const void *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, 0x7fffffff,
                                           0x00020000);
}

template <typename T>
__device__ __forceinline__ T bld(rsrc_t r, uint32_t vo, uint32_t so);
template <>
__device__ __forceinline__ float bld<float>(rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
template <>
__device__ __forceinline__ int32_t bld<int32_t>(rsrc_t r, uint32_t vo, uint32_t so) {
  return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0);
}
template <>
__device__ __forceinline__ double bld<double>(rsrc_t r, uint32_t vo, uint32_t so) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}

template <typename T>
__device__ __forceinline__ void bst(T v, rsrc_t r, uint32_t vo, uint32_t so);
template <>
__device__ __forceinline__ void bst<float>(float v, rsrc_t r, uint32_t vo, uint32_t so) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, vo, so, 0);
}
template <>
__device__ __forceinline__ void bst<double>(double v, rsrc_t r, uint32_t vo, uint32_t so) {
  using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, vo, so, 0);
}

// Logical block index such that the blocks the hardware places on one XCD
// (block b runs on XCD b % nxcd) process one contiguous range of logical
// blocks: neighbouring elements then share that XCD's L2.  nxcd <= 1: identity.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb, unsigned nxcd) {
  if (nxcd <= 1) return b;
  const unsigned q = nb / nxcd, r = nb % nxcd, x = b % nxcd, i = b / nxcd;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

}  // namespace mjw
