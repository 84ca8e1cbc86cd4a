// gfx950 kernels on whole pictures: 4:2:0 -> 4:4:4 nearest-neighbour upsampling, the store of
// the engine's 16-bit shift-compensated planes into a caller's picture, RGB <-> YCbCr conversion
// and plane exchange.
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

}  // namespace

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

hipError_t launch_narrow(const uint16_t* src, int64_t n, uint8_t* dst, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_narrow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, n, dst);
  return hipGetLastError();
}

}  // namespace qmiv
