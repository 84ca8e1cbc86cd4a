// gfx950 kernels on whole pictures: 4:2:0 -> 4:4:4 nearest-neighbour upsampling, the unpack of raw
// file frames (4:2:0 of any width, 4:2:2, 4:0:0) into 4:4:4 pictures, the store of the engine's
// 16-bit shift-compensated planes into a caller's picture, RGB <-> YCbCr conversion, plane exchange
// and the clip of out-of-range samples.
#include "device_common.h"
#include "qmiv_internal.h"

namespace qmiv {
namespace {

// xSeqBase::xUnpackFrame with CvtUpsampleHV / UpsampleHV (libs/pmbb-core/src/xSeq.cpp:197-213,
// libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49,68-87): dst[2y + i][2x + j] = src[y][x] for the chroma
// planes, a copy for luma (and for 4:4:4 sources). One thread writes 4 consecutive samples of
// one destination row; grid = (ceil(w / 4 / 256), h, 3 planes).
template <typename T>
__global__ __launch_bounds__(256) void k_upsample(const T* __restrict__ src, Geom gs, T* __restrict__ dst) {
  const int x = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int y = blockIdx.y;
  const int c = blockIdx.z;
  if (x >= gs.w) return;
  const int sh = c == 0 ? 0 : gs.cs;
  const int sw = c == 0 ? gs.w : gs.cw;
  const T* s = src + (c == 0 ? 0 : c == 1 ? gs.cb_off : gs.cr_off) + (int64_t)(y >> sh) * sw;
  T* d = dst + (int64_t)c * gs.w * gs.h + (int64_t)y * gs.w;
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (x + i < gs.w) d[x + i] = s[(x + i) >> sh];
}

// dst (8-bit picture, planar 4:4:4) = src (16-bit planes), values already within 8 bits.
__global__ __launch_bounds__(256) void k_narrow(const uint16_t* __restrict__ src, int64_t n, uint8_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (uint8_t)src[i];
}

// Eight consecutive samples of one plane, moved as one 8- or 16-byte access.
template <typename T>
struct alignas(8 * sizeof(T)) Vec8 {
  T v[8];
};

// int32 arithmetic that wraps like the reference's int32 expressions (they overflow at 16 bits,
// where 65536 * sample exceeds 2^31); >> is arithmetic on int32.
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
__device__ __forceinline__ int32_t wshl(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }

// One pel of xColorSpaceSTD::ConvertRGB2YCbCr_I32 / ConvertYCbCr2RGB_I32
// (libs/pmbb-core/src/xColorSpaceSTD.cpp:195-272): Q16 products, + 2^15, arithmetic >> 16, chroma
// offset by 2^(bd-1), clipped to [0, 2^bd - 1]. The forward matrix's 0.5 taps are shifts by 15.
template <bool TO_RGB>
__device__ __forceinline__ void colour_pel(int32_t a, int32_t b, int32_t c, const ColourArgs& k, int32_t& o0,
                                           int32_t& o1, int32_t& o2) {
  constexpr int32_t kAdd = 1 << 15;
  if constexpr (!TO_RGB) {
    const int32_t y = wadd(wadd(wadd(wmul(k.m[0], a), wmul(k.m[1], b)), wmul(k.m[2], c)), kAdd) >> 16;
    const int32_t u = wadd(wadd(wadd(wmul(k.m[3], a), wmul(k.m[4], b)), wshl(c, 15)), kAdd) >> 16;
    const int32_t v = wadd(wadd(wadd(wshl(a, 15), wmul(k.m[7], b)), wmul(k.m[8], c)), kAdd) >> 16;
    o0 = clampi(y, 0, k.maxv);
    o1 = clampi(u + k.mid, 0, k.maxv);
    o2 = clampi(v + k.mid, 0, k.maxv);
  } else {
    const int32_t u = b - k.mid, v = c - k.mid;
    const int32_t sy = wadd(wshl(a, 16), kAdd);
    o0 = clampi(wadd(sy, wmul(k.m[2], v)) >> 16, 0, k.maxv);
    o1 = clampi(wadd(wadd(sy, wmul(k.m[4], u)), wmul(k.m[5], v)) >> 16, 0, k.maxv);
    o2 = clampi(wadd(sy, wmul(k.m[7], u)) >> 16, 0, k.maxv);
  }
}

// One thread converts 8 consecutive pels (vector loads and stores when every plane is aligned to
// 8 samples). Each pel is read before it is written by the same thread, so dst may equal src.
template <typename T, bool TO_RGB>
__global__ __launch_bounds__(256) void k_colour(const T* src, T* dst, int64_t plane, ColourArgs k, int vec) {
  const int64_t i0 = 8 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (i0 >= plane) return;
  if (vec) {
    const Vec8<T> a = *(const Vec8<T>*)(src + i0);
    const Vec8<T> b = *(const Vec8<T>*)(src + plane + i0);
    const Vec8<T> c = *(const Vec8<T>*)(src + 2 * plane + i0);
    Vec8<T> x, y, z;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      int32_t o0, o1, o2;
      colour_pel<TO_RGB>(a.v[j], b.v[j], c.v[j], k, o0, o1, o2);
      x.v[j] = (T)o0;
      y.v[j] = (T)o1;
      z.v[j] = (T)o2;
    }
    *(Vec8<T>*)(dst + i0) = x;
    *(Vec8<T>*)(dst + plane + i0) = y;
    *(Vec8<T>*)(dst + 2 * plane + i0) = z;
  } else {
    const int n = plane - i0 < 8 ? (int)(plane - i0) : 8;
    for (int j = 0; j < n; j++) {
      int32_t o0, o1, o2;
      colour_pel<TO_RGB>(src[i0 + j], src[plane + i0 + j], src[2 * plane + i0 + j], k, o0, o1, o2);
      dst[i0 + j] = (T)o0;
      dst[plane + i0 + j] = (T)o1;
      dst[2 * plane + i0 + j] = (T)o2;
    }
  }
}

// Exchanges two planes of n samples, 8 samples per thread.
template <typename T>
__global__ __launch_bounds__(256) void k_swap(T* __restrict__ a, T* __restrict__ b, int64_t n, int vec) {
  const int64_t i0 = 8 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (i0 >= n) return;
  if (vec) {
    const Vec8<T> x = *(const Vec8<T>*)(a + i0);
    const Vec8<T> y = *(const Vec8<T>*)(b + i0);
    *(Vec8<T>*)(a + i0) = y;
    *(Vec8<T>*)(b + i0) = x;
  } else {
    const int m = n - i0 < 8 ? (int)(n - i0) : 8;
    for (int j = 0; j < m; j++) {
      const T t = a[i0 + j];
      a[i0 + j] = b[i0 + j];
      b[i0 + j] = t;
    }
  }
}

// xSeqBase::xUnpackFrame (libs/pmbb-core/src/xSeq.cpp:179-249) of one packed file frame (bytes at
// src, src_bytes of them) into a 4:4:4 picture of the same sample size: luma copied; 4:2:0 and 4:2:2
// chroma upsampled nearest-neighbour from rows of W >> 1 samples (UpsampleHV / UpsampleH,
// libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49, with the file stride of xSeq.cpp:200,216), so for an odd
// width the last column takes sample W >> 1 of its row, which is the next row's first; 4:4:4 chroma
// copied; 4:0:0 chroma set to the mid value 2^(bd - 1) (xSeq.cpp:399-400). Chroma planes begin at
// byte offsets cb_byte / cr_byte, which may be odd for 16-bit samples, so samples are read bytewise;
// a sample past the end of the frame reads 0. grid = (ceil(w / 256), h, 3).
template <typename T>
__global__ __launch_bounds__(256) void k_unpack(const uint8_t* __restrict__ src, uint64_t src_bytes, int32_t w,
                                                int32_t h, int32_t fcf, uint64_t cb_byte, uint64_t cr_byte,
                                                uint32_t mid, T* __restrict__ dst) {
  const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t y = blockIdx.y;
  const int c = blockIdx.z;
  if (x >= w) return;
  uint32_t v;
  if (c != 0 && fcf == 400) {
    v = mid;
  } else {
    int64_t idx;
    uint64_t base = 0;
    if (c == 0 || fcf == 444) idx = (int64_t)y * w + x;
    else if (fcf == 422) idx = (int64_t)y * (w >> 1) + (x >> 1);
    else idx = (int64_t)(y >> 1) * (w >> 1) + (x >> 1);
    if (c == 1) base = cb_byte;
    if (c == 2) base = cr_byte;
    const uint64_t b = base + (uint64_t)idx * sizeof(T);
    v = b < src_bytes ? (uint32_t)src[b] : 0u;
    if constexpr (sizeof(T) == 2) v |= (b + 1 < src_bytes ? (uint32_t)src[b + 1] : 0u) << 8;
  }
  dst[(int64_t)c * w * h + (int64_t)y * w + x] = (T)v;
}

// dst (16-bit planes) = src (8-bit planes), n samples: a picture read back into the reference's
// uint16 layout.
__global__ __launch_bounds__(256) void k_widen(const uint8_t* __restrict__ src, int64_t n, uint16_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// The reference's interleaved picture (xPicI: 4 x uint16 per pel, Y Cb Cr A, libs/pmbb-core/src/xPic.h:79-113,
// xPic.cpp:201-207), its rows packed one after another (4 w samples each), to and from a planar 4:4:4
// frame of w x h samples; A is 0 in the interleaved form. grid = (ceil(w / 256), h).
template <typename T>
__global__ __launch_bounds__(256) void k_deinterleave(const uint16_t* __restrict__ src, int32_t w, int32_t h,
                                                      T* __restrict__ dst) {
  const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t y = blockIdx.y;
  if (x >= w) return;
  const ushort4 v = *reinterpret_cast<const ushort4*>(src + 4 * ((int64_t)y * w + x));
  const int64_t plane = (int64_t)w * h, o = (int64_t)y * w + x;
  dst[o] = (T)v.x;
  dst[plane + o] = (T)v.y;
  dst[2 * plane + o] = (T)v.z;
}
template <typename T>
__global__ __launch_bounds__(256) void k_interleave(const T* __restrict__ src, int32_t w, int32_t h,
                                                    uint16_t* __restrict__ dst) {
  const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t y = blockIdx.y;
  if (x >= w) return;
  const int64_t plane = (int64_t)w * h, o = (int64_t)y * w + x;
  ushort4 v;
  v.x = src[o];
  v.y = src[plane + o];
  v.z = src[2 * plane + o];
  v.w = 0;
  *reinterpret_cast<ushort4*>(dst + 4 * o) = v;
}

// ClipToRange of n 16-bit samples, 8 per thread, when *max_pel (the frame's largest sample from the
// statistics pass; null: unconditional) exceeds maxv. Every thread reads the flag and leaves at once
// when the frame is in range, so the launch costs almost nothing then.
__global__ __launch_bounds__(256) void k_clip(uint16_t* __restrict__ p, int64_t n, int32_t maxv,
                                              const int32_t* __restrict__ max_pel, int vec) {
  if (max_pel && *max_pel <= maxv) return;
  const int64_t i0 = 8 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (i0 >= n) return;
  const uint16_t m = (uint16_t)maxv;
  if (vec) {
    Vec8<uint16_t> x = *(const Vec8<uint16_t>*)(p + i0);
#pragma unroll
    for (int j = 0; j < 8; j++) x.v[j] = x.v[j] > m ? m : x.v[j];
    *(Vec8<uint16_t>*)(p + i0) = x;
  } else {
    const int k = n - i0 < 8 ? (int)(n - i0) : 8;
    for (int j = 0; j < k; j++)
      if (p[i0 + j] > m) p[i0 + j] = m;
  }
}

}  // namespace

hipError_t launch_clip(void* pic, const Geom& g, int32_t maxv, const int32_t* max_pel, hipStream_t stream) {
  if (g.bytes_per_sample != 2) return hipSuccess;  // 8-bit samples never exceed an 8-bit maximum
  const int64_t n = (int64_t)g.frame_bytes / 2;
  if (n <= 0) return hipSuccess;
  const int vec = (n % 8 == 0) && ((uintptr_t)pic % 16 == 0);
  const dim3 grid((unsigned)((n + 8 * 256 - 1) / (8 * 256)));
  k_clip<<<grid, 256, 0, stream>>>((uint16_t*)pic, n, maxv, max_pel, vec);
  return hipGetLastError();
}

hipError_t launch_unpack(const void* src, uint64_t src_bytes, int32_t w, int32_t h, int32_t bd, int32_t fcf,
                         void* dst, hipStream_t stream) {
  const uint64_t cmp = (uint64_t)w * h * (bd <= 8 ? 1 : 2);
  uint64_t cb = cmp, cr = cmp;
  if (fcf == 420) cr = cmp + (cmp >> 2);
  else if (fcf == 422) cr = cmp + (cmp >> 1);
  else if (fcf == 444) cr = 2 * cmp;
  else if (fcf != 400) return hipErrorInvalidValue;
  const uint32_t mid = 1u << (bd - 1);
  const dim3 grid((unsigned)((w + 255) / 256), (unsigned)h, 3);
  if (bd <= 8)
    k_unpack<uint8_t><<<grid, 256, 0, stream>>>((const uint8_t*)src, src_bytes, w, h, fcf, cb, cr, mid, (uint8_t*)dst);
  else
    k_unpack<uint16_t><<<grid, 256, 0, stream>>>((const uint8_t*)src, src_bytes, w, h, fcf, cb, cr, mid, (uint16_t*)dst);
  return hipGetLastError();
}

hipError_t launch_colour(const void* src, void* dst, const Geom& g, const ColourArgs& k, bool to_rgb,
                         hipStream_t stream) {
  const int64_t plane = (int64_t)g.w * g.h;
  if (plane <= 0) return hipSuccess;
  const int vec = (plane % 8 == 0) && ((uintptr_t)src % (8 * g.bytes_per_sample) == 0) &&
                  ((uintptr_t)dst % (8 * g.bytes_per_sample) == 0);
  const dim3 grid((unsigned)((plane + 8 * 256 - 1) / (8 * 256)));
  if (g.bytes_per_sample == 1) {
    if (to_rgb)
      k_colour<uint8_t, true><<<grid, 256, 0, stream>>>((const uint8_t*)src, (uint8_t*)dst, plane, k, vec);
    else
      k_colour<uint8_t, false><<<grid, 256, 0, stream>>>((const uint8_t*)src, (uint8_t*)dst, plane, k, vec);
  } else {
    if (to_rgb)
      k_colour<uint16_t, true><<<grid, 256, 0, stream>>>((const uint16_t*)src, (uint16_t*)dst, plane, k, vec);
    else
      k_colour<uint16_t, false><<<grid, 256, 0, stream>>>((const uint16_t*)src, (uint16_t*)dst, plane, k, vec);
  }
  return hipGetLastError();
}

hipError_t launch_swap_planes(void* pic, const Geom& g, int a, int b, hipStream_t stream) {
  const int64_t plane = (int64_t)g.w * g.h;
  if (plane <= 0 || a == b) return hipSuccess;
  const int vec = (plane % 8 == 0) && ((uintptr_t)pic % (8 * g.bytes_per_sample) == 0);
  const dim3 grid((unsigned)((plane + 8 * 256 - 1) / (8 * 256)));
  if (g.bytes_per_sample == 1) {
    uint8_t* p = (uint8_t*)pic;
    k_swap<uint8_t><<<grid, 256, 0, stream>>>(p + a * plane, p + b * plane, plane, vec);
  } else {
    uint16_t* p = (uint16_t*)pic;
    k_swap<uint16_t><<<grid, 256, 0, stream>>>(p + a * plane, p + b * plane, plane, vec);
  }
  return hipGetLastError();
}

hipError_t launch_upsample(const void* src, const Geom& gs, void* dst, hipStream_t stream) {
  const dim3 grid((unsigned)((gs.w + 1023) / 1024), (unsigned)gs.h, 3);
  if (gs.bytes_per_sample == 1)
    hipLaunchKernelGGL(k_upsample<uint8_t>, grid, dim3(256), 0, stream, (const uint8_t*)src, gs, (uint8_t*)dst);
  else
    hipLaunchKernelGGL(k_upsample<uint16_t>, grid, dim3(256), 0, stream, (const uint16_t*)src, gs, (uint16_t*)dst);
  return hipGetLastError();
}

hipError_t launch_widen(const uint8_t* src, int64_t n, uint16_t* dst, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, n, dst);
  return hipGetLastError();
}

hipError_t launch_interleave(const void* planar, const Geom& g, uint16_t* inter, bool to_interleaved,
                             hipStream_t stream) {
  const dim3 grid((unsigned)((g.w + 255) / 256), (unsigned)g.h);
  if (to_interleaved) {
    if (g.bytes_per_sample == 1) k_interleave<uint8_t><<<grid, 256, 0, stream>>>((const uint8_t*)planar, g.w, g.h, inter);
    else k_interleave<uint16_t><<<grid, 256, 0, stream>>>((const uint16_t*)planar, g.w, g.h, inter);
  } else {
    if (g.bytes_per_sample == 1) k_deinterleave<uint8_t><<<grid, 256, 0, stream>>>(inter, g.w, g.h, (uint8_t*)planar);
    else k_deinterleave<uint16_t><<<grid, 256, 0, stream>>>(inter, g.w, g.h, (uint16_t*)planar);
  }
  return hipGetLastError();
}

hipError_t launch_narrow(const uint16_t* src, int64_t n, uint8_t* dst, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_narrow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, n, dst);
  return hipGetLastError();
}

}  // namespace qmiv
