// bufio.hpp — raw buffer addressing for plane-marching kernels on gfx9: a per-plane buffer
// descriptor (scalar base, the plane's byte range) plus a per-thread in-plane byte offset, so a
// plane step needs no 64-bit vector address arithmetic.  An offset at or past the descriptor's
// range reads 0 and discards stores (the hardware range check), which masks lanes without
// exec-mask branches.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace pdehip {

inline constexpr unsigned kBufRsrcWord3 = 0x00020000u;   // gfx9 raw buffer: 32-bit, no swizzle
inline constexpr unsigned kDropOffset = 0x80000000u;     // past any plane's range: store discarded

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, long long elem,
                                                             int esize, int bytes) {
  char* p = const_cast<char*>(static_cast<const char*>(base)) + elem * esize;
  return __builtin_amdgcn_make_buffer_rsrc(p, static_cast<short>(0), bytes,
                                           static_cast<int>(kBufRsrcWord3));
}

template <typename S> struct BufIO;
template <> struct BufIO<float> {
  static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, unsigned byte_off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
  }
  static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, unsigned byte_off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, byte_off, 0, 0);
  }
};
template <> struct BufIO<__half> {
  static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t r, unsigned byte_off) {
    const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(r, byte_off, 0, 0);
    return __half2float(__ushort_as_half(h));
  }
  static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, unsigned byte_off, __half v) {
    __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(v), r, byte_off, 0, 0);
  }
};

}  // namespace pdehip
