// gfx950 kernels on whole pictures: 4:2:0 -> 4:4:4 nearest-neighbour upsampling and the store of
// the engine's 16-bit shift-compensated planes into a caller's picture.
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

}  // namespace

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
