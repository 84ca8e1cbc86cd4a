// gfx950 kernels of the SSIM family: windowed moments (SSIM, MS-SSIM scales, IV-SSIM), the
// integer 2x2 pyramid of MS-SSIM and the per-row reduction of tile partial sums.
#include "qmiv_internal.h"

namespace qmiv {
namespace {

constexpr int kSsimTW = 64;            // output columns per workgroup (one per lane)
constexpr int kSsimK = 8;              // output rows per lane
constexpr int kSsimTH = 4 * kSsimK;    // output rows per workgroup
constexpr int kSsimLW = kSsimTW + 9;   // window columns dx in [-5, 4]
constexpr int kSsimLH = kSsimTH + 9;

constexpr int32_t kBufDword3 = 0x00020000;

// One plane read through its 4:4:4 view: value(x, y) = plane[(y >> shift) * stride + (x >> shift)].
struct ViewBuf {
  __amdgpu_buffer_rsrc_t r;
  int32_t stride, shift, bytes;
  __device__ __forceinline__ explicit ViewBuf(const PlaneView& v)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(v.base), 0, (int)v.nbytes, kBufDword3)),
        stride(v.stride), shift(v.shift), bytes(v.bytes) {}
  __device__ __forceinline__ uint32_t at(int x, int y) const {
    const uint32_t i = (uint32_t)((y >> shift) * stride + (x >> shift));
    return bytes == 2 ? (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, 0, 0)
                      : (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, i, 0, 0);
  }
};

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// ---------------------------------------------------------------------------------------------
// SSIM of one component: xSSIM::xCalcRowSSIM / xStructSim<flt64, CalcL>::CalcPel
// (libs/pmbb-ivqm/src/xSSIM.cpp:114-141, libs/pmbb-ivqm/src/xStructSim.cpp:15-60).
// Interior pels x in [5, w - 5), y in [5, h - 5); window taps dy, dx in [-5, 4] with the fp32
// Gaussian weights widened to fp64; moments sum R, T, R^2, T^2, RT in fp64, accumulated per pel
// in the reference's order (dy outer, dx inner; fused multiply-add). Output: one partial row sum
// per (tile column, row).
// ---------------------------------------------------------------------------------------------
template <int CALC_L>
__global__ __launch_bounds__(256) void k_ssim(SsimArgs a) {
  __shared__ float s_t[kSsimLH * kSsimLW];
  __shared__ float s_r[kSsimLH * kSsimLW];
  // Weights transposed and zero-padded: s_w[dx * kWRow + dy + 7] = F[dy][dx] for dy in [0, 9], so
  // the weights F[r - k][dx] of the lane's pels k = 0..7 for loaded row r are 8 consecutive doubles.
  constexpr int kWRow = 24;
  __shared__ double s_w[10 * kWRow];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tx = blockIdx.x;
  const int ty = blockIdx.y;
  const int x0 = 5 + tx * kSsimTW;  // first output column of the tile
  const int y0 = 5 + ty * kSsimTH;
  const ViewBuf T(a.tst), R(a.ref);

  for (int i = threadIdx.x; i < 10 * kWRow; i += 256) {
    const int dx = i / kWRow;
    const int dy = i - dx * kWRow - 7;
    s_w[i] = (dy >= 0 && dy <= 9) ? a.f[dy * 10 + dx] : 0.0;
  }
  // Stage the window rows [y0 - 5, y0 + TH + 4) x [x0 - 5, x0 + TW + 4), clamped inside the
  // picture (clamped values only feed pels outside the interior, which are discarded).
  for (int i = threadIdx.x; i < kSsimLH * kSsimLW; i += 256) {
    const int ly = i / kSsimLW;
    const int lx = i - ly * kSsimLW;
    const int gy = min(y0 - 5 + ly, a.h - 1);
    const int gx = min(x0 - 5 + lx, a.w - 1);
    s_t[i] = (float)T.at(gx, gy);
    s_r[i] = (float)R.at(gx, gy);
  }
  __syncthreads();

  double acc[kSsimK][5];
#pragma unroll
  for (int k = 0; k < kSsimK; k++)
#pragma unroll
    for (int m = 0; m < 5; m++) acc[k][m] = 0.0;

#pragma unroll
  for (int r = 0; r < kSsimK + 9; r++) {
    const int base = (kSsimK * wave + r) * kSsimLW + lane;
#pragma unroll
    for (int dx = 0; dx < 10; dx++) {
      const double tv = (double)s_t[base + dx];
      const double rv = (double)s_r[base + dx];
      const double rr = rv * rv, tt = tv * tv, rt = rv * tv;
      const double* wcol = s_w + dx * kWRow + r + 7;  // wcol[-k] = F[r - k][dx]
#pragma unroll
      for (int k = 0; k < kSsimK; k++) {
        const int dy = r - k;  // window row of pel k that this loaded row is
        if (dy < 0 || dy > 9) continue;
        const double c = wcol[-k];
        acc[k][0] = __builtin_fma(rv, c, acc[k][0]);
        acc[k][1] = __builtin_fma(tv, c, acc[k][1]);
        acc[k][2] = __builtin_fma(rr, c, acc[k][2]);
        acc[k][3] = __builtin_fma(tt, c, acc[k][3]);
        acc[k][4] = __builtin_fma(rt, c, acc[k][4]);
      }
      // One loaded value at a time: bounds the live products and weights.
      __builtin_amdgcn_sched_barrier(0);
    }
    // Re-read weights and samples for every loaded row instead of keeping them in registers.
    asm volatile("" ::: "memory");
  }

  const int x = x0 + lane;
#pragma unroll
  for (int k = 0; k < kSsimK; k++) {
    const int y = y0 + kSsimK * wave + k;
    double v = 0.0;
    if (x < a.w - 5 && y < a.h - 5) {
      const double aR = acc[k][0], aT = acc[k][1];
      const double vR = acc[k][2] - aR * aR;
      const double vT = acc[k][3] - aT * aT;
      const double cov = acc[k][4] - aR * aT;
      const double cs = (2 * cov + a.c2) / (vR + vT + a.c2);
      if (CALC_L) {
        const double l = (2 * aR * aT + a.c1) / (aR * aR + aT * aT + a.c1);
        v = l * cs;
      } else {
        v = cs;
      }
    }
    v = wave_sum_f64(v);
    if (lane == 0 && y < a.h - 5) a.partials[(int64_t)tx * a.h + y] = v;
  }
}

// rows[y] = sum over tile columns of the partials (rows outside the interior are zero).
__global__ __launch_bounds__(256) void k_ssim_rows(const double* __restrict__ part, int32_t tiles_x, int32_t h,
                                                   double* __restrict__ rows) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  double s = 0.0;
  if (y >= 5 && y < h - 5)
    for (int t = 0; t < tiles_x; t++) s += part[(int64_t)t * h + y];
  rows[y] = s;
}

// xPixelOpsSTD::DownsampleHV (libs/pmbb-core/src/xPixelOpsSTD.cpp:50-67): dst = (a + b + c + d + 2) >> 2
// over 2x2 blocks of the source view; dst is dw x dh uint16 (stride dw).
__global__ __launch_bounds__(256) void k_downsample(PlaneView src, int32_t dw, int32_t dh, uint16_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)dw * dh) return;
  const int y = (int)(i / dw);
  const int x = (int)(i - (int64_t)y * dw);
  const ViewBuf S(src);
  const uint32_t v = S.at(2 * x, 2 * y) + S.at(2 * x + 1, 2 * y) + S.at(2 * x, 2 * y + 1) + S.at(2 * x + 1, 2 * y + 1);
  dst[i] = (uint16_t)((v + 2) >> 2);
}

}  // namespace

hipError_t launch_ssim(const SsimArgs& a, hipStream_t stream) {
  const int tiles_x = a.w > 10 ? (a.w - 10 + kSsimTW - 1) / kSsimTW : 0;
  const int tiles_y = a.h > 10 ? (a.h - 10 + kSsimTH - 1) / kSsimTH : 0;
  if (tiles_x != a.tiles_x) return hipErrorInvalidValue;
  if (a.h <= 0) return hipSuccess;  // empty pyramid level of a small picture: no rows
  if (tiles_x > 0 && tiles_y > 0) {
    if (a.calc_l)
      hipLaunchKernelGGL(k_ssim<1>, dim3(tiles_x, tiles_y), dim3(256), 0, stream, a);
    else
      hipLaunchKernelGGL(k_ssim<0>, dim3(tiles_x, tiles_y), dim3(256), 0, stream, a);
    return hipGetLastError();
  }
  return hipSuccess;
}

hipError_t launch_ssim_rows(const SsimArgs& a, double* rows, hipStream_t stream) {
  if (a.h <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ssim_rows, dim3((a.h + 255) / 256), dim3(256), 0, stream, a.partials, a.tiles_x, a.h, rows);
  return hipGetLastError();
}

int32_t ssim_tiles_x(int32_t w) { return w > 10 ? (w - 10 + kSsimTW - 1) / kSsimTW : 0; }

hipError_t launch_downsample(const PlaneView& src, int32_t dw, int32_t dh, uint16_t* dst, hipStream_t stream) {
  const int64_t n = (int64_t)dw * dh;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_downsample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, dw, dh, dst);
  return hipGetLastError();
}

}  // namespace qmiv
