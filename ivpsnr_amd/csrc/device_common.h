// Device helpers shared by the gfx950 kernel files: buffer-resource loads and stores, wave
// reductions, clamping and the 24-bit multiply-add kept as one instruction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qmiv {
namespace {

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Buffer-resource view of one packed frame: 32-bit sample offsets, one VGPR per address, and
// hardware range checking (reads beyond `bytes` return 0, writes are dropped).
constexpr int32_t kBufDword3 = 0x00020000;  // raw buffer, gfx9-family data format
template <typename T>
struct Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const void* p, uint32_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, kBufDword3)) {}
  __device__ __forceinline__ uint32_t ld(uint32_t i) const {
    if constexpr (sizeof(T) == 2) return __builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, 0, 0);
    else return __builtin_amdgcn_raw_buffer_load_b8(r, i, 0, 0);
  }
  __device__ __forceinline__ void st(uint32_t i, uint32_t v) const {
    if constexpr (sizeof(T) == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, i * 2u, 0, 0);
    else __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v, r, i, 0, 0);
  }
};

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}

__device__ __forceinline__ int32_t mad24(int32_t a, int32_t b, int32_t c) {
  // Kept as one instruction: the compiler would otherwise reassociate a chain of these into
  // mul + mul + add3. ALU only.
  int32_t d;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

}  // namespace
}  // namespace qmiv
