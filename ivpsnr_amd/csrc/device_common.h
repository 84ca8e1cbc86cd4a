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

// The SSIM window F is symmetric: F[a][b] = F[b][a], and F[5 + d][b] = F[5 - d][b] (a Gaussian of
// the offsets -5 .. 4). Its 100 taps hold 21 distinct weights: tap (dy, dx) has weight sym_id(dy, dx),
// and sym_rep(u) is a tap index (dy * 10 + dx) of weight u. The launchers check the table's symmetry.
constexpr int kSsimSym = 21;
__host__ __device__ constexpr int sym_off(int i) { return i < 5 ? 5 - i : i - 5; }  // |offset| of tap index i
__host__ __device__ constexpr int sym_id(int dy, int dx) {
  const int a = sym_off(dy) < sym_off(dx) ? sym_off(dy) : sym_off(dx);
  const int b = sym_off(dy) < sym_off(dx) ? sym_off(dx) : sym_off(dy);
  return b * (b + 1) / 2 + a;  // pairs a <= b of 0 .. 5
}
__host__ __device__ constexpr int sym_rep(int u) {
  int b = 0;
  while ((b + 1) * (b + 2) / 2 <= u) b++;
  const int a = u - b * (b + 1) / 2;
  return (5 - a) * 10 + (5 - b);
}
// Whether a 10 x 10 table has that symmetry (every tap equals its weight's representative).
inline bool ssim_window_symmetric(const double* f) {
  for (int i = 0; i < 100; i++)
    if (f[i] != f[sym_rep(sym_id(i / 10, i % 10))]) return false;
  return true;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b to XCD b mod 8). Renumbered by this
// bijection of [0, n), XCD x takes the contiguous run [x q + min(x, r), ...) of q + (x < r) tiles
// (q = n / 8, r = n mod 8), so neighbouring tiles, whose halos overlap, share one L2.
__device__ __forceinline__ int xcd_order(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, j = b >> 3;
  return x * q + min(x, r) + j;
}

}  // namespace
}  // namespace qmiv
