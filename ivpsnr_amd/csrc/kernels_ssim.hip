// gfx950 kernels of the SSIM family: windowed moments (SSIM, MS-SSIM scales, IV-SSIM), the
// integer 2x2 pyramid of MS-SSIM and the per-row reduction of tile partial sums.
#include <utility>

#include "qmiv_internal.h"

namespace qmiv {
namespace {

// Tile shape: each lane owns kSsimC adjacent output columns x kSsimK output rows; 4 waves stack
// their rows. A loaded window sample feeds up to kSsimC x kSsimK pels.
#ifndef QMIV_SSIM_COLS
#define QMIV_SSIM_COLS 1
#endif
#ifndef QMIV_SSIM_ROWS
#define QMIV_SSIM_ROWS 4
#endif
constexpr int kSsimC = QMIV_SSIM_COLS;
constexpr int kSsimK = QMIV_SSIM_ROWS;
constexpr int kSsimTW = 64 * kSsimC;                 // output columns per workgroup
constexpr int kSsimTH = 4 * kSsimK;                  // output rows per workgroup
constexpr int kSsimLW = (kSsimTW + 9 + 1) & ~1;      // window columns dx in [-5, 4], even row stride
constexpr int kSsimLH = kSsimTH + 9;
constexpr int kSsimWRow = kSsimK + 16;               // stride of one transposed weight column in LDS
static_assert(kSsimC == 1 || kSsimC == 2, "one or two columns per lane");

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

// The window is walked as steps (R, c): loaded window row R = 0 .. K + 8 of the lane's K x C output
// pels, sample column c = 0 .. 8 + C relative to the lane's first window column. Sample column c
// feeds output column j with window column dxi = c - j and output row k with window row dy = R - k,
// weight F[dy][dxi]; k runs over max(0, R - 9) .. min(K - 1, R).
constexpr int kSsimSW = 9 + kSsimC;               // sample columns per loaded row
constexpr int kSsimSteps = (kSsimK + 9) * kSsimSW;

// One step's operands: the two samples and the weights of every (j, k) they feed.
struct SsimStep {
  float t, r;
  double w[kSsimC][kSsimK];
};

template <int S>
__device__ __forceinline__ void ssim_load(const float* s_t, const float* s_r, const double* s_w, int base,
                                          SsimStep& d) {
  constexpr int R = S / kSsimSW, c = S % kSsimSW;
  constexpr int k_lo = R > 9 ? R - 9 : 0;
  constexpr int k_hi = R < kSsimK - 1 ? R : kSsimK - 1;
  d.t = s_t[base + R * kSsimLW + c];
  d.r = s_r[base + R * kSsimLW + c];
#pragma unroll
  for (int j = 0; j < kSsimC; j++) {
    constexpr int unused = 0;
    (void)unused;
    const int dxi = c - j;
    if (dxi < 0 || dxi > 9) continue;
#pragma unroll
    for (int k = k_lo; k <= k_hi; k++) d.w[j][k] = s_w[dxi * kSsimWRow + R - k + 7];  // F[R - k][dxi]
  }
}

template <int S>
__device__ __forceinline__ void ssim_fma(const SsimStep& d, double (&acc)[kSsimC][kSsimK][5]) {
  constexpr int R = S / kSsimSW, c = S % kSsimSW;
  constexpr int k_lo = R > 9 ? R - 9 : 0;
  constexpr int k_hi = R < kSsimK - 1 ? R : kSsimK - 1;
  const double tv = (double)d.t, rv = (double)d.r;
  const double rr = rv * rv, tt = tv * tv, rt = rv * tv;
#pragma unroll
  for (int j = 0; j < kSsimC; j++) {
    const int dxi = c - j;
    if (dxi < 0 || dxi > 9) continue;
#pragma unroll
    for (int k = k_lo; k <= k_hi; k++) {
      const double w = d.w[j][k];
      acc[j][k][0] = __builtin_fma(rv, w, acc[j][k][0]);
      acc[j][k][1] = __builtin_fma(tv, w, acc[j][k][1]);
      acc[j][k][2] = __builtin_fma(rr, w, acc[j][k][2]);
      acc[j][k][3] = __builtin_fma(tt, w, acc[j][k][3]);
      acc[j][k][4] = __builtin_fma(rt, w, acc[j][k][4]);
    }
  }
  // Passing the step's accumulators through an ordered no-op pins its FMAs in place: the body is
  // one basic block, and without these the instruction DAG would hoist every load and product above
  // every FMA (the memory clobber keeps the next loads below, the weights out of registers).
#pragma unroll
  for (int j = 0; j < kSsimC; j++) {
    const int dxi = c - j;
    if (dxi < 0 || dxi > 9) continue;
#pragma unroll
    for (int k = k_lo; k <= k_hi; k++)
      asm volatile("" : "+v"(acc[j][k][0]), "+v"(acc[j][k][1]), "+v"(acc[j][k][2]), "+v"(acc[j][k][3]),
                   "+v"(acc[j][k][4])::"memory");
  }
  __builtin_amdgcn_sched_barrier(0);
}

// Software pipeline over the steps: the operands of step S + 1 are loaded from LDS before the FMAs
// of step S, so the LDS latency of every step overlaps the previous step's arithmetic.
template <int S>
__device__ __forceinline__ void ssim_steps(const float* s_t, const float* s_r, const double* s_w, int base,
                                           double (&acc)[kSsimC][kSsimK][5], const SsimStep& cur) {
  if constexpr (S + 1 < kSsimSteps) {
    SsimStep nxt;
    ssim_load<S + 1>(s_t, s_r, s_w, base, nxt);
    ssim_fma<S>(cur, acc);
    ssim_steps<S + 1>(s_t, s_r, s_w, base, acc, nxt);
  } else {
    ssim_fma<S>(cur, acc);
  }
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
  // Weights transposed: s_w[dx * kSsimWRow + dy + 7] = F[dy][dx] for dy in [0, 9], so the weights
  // F[r - k][dx] of the lane's pels k for loaded row r are consecutive doubles.
  __shared__ __attribute__((aligned(16))) double s_w[10 * kSsimWRow];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tx = blockIdx.x;
  const int ty = blockIdx.y;
  const int x0 = 5 + tx * kSsimTW;  // first output column of the tile
  const int y0 = 5 + ty * kSsimTH;
  const ViewBuf T(a.tst), R(a.ref);

  for (int i = threadIdx.x; i < 10 * kSsimWRow; i += 256) {
    const int dx = i / kSsimWRow;
    const int dy = i - dx * kSsimWRow - 7;
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

  double acc[kSsimC][kSsimK][5];
#pragma unroll
  for (int j = 0; j < kSsimC; j++)
#pragma unroll
    for (int k = 0; k < kSsimK; k++)
#pragma unroll
      for (int m = 0; m < 5; m++) acc[j][k][m] = 0.0;

  // Every step as its own compile-time instance (the loop unroller gives up on this many
  // instructions): straight-line code, constant LDS offsets, no branches.
  {
    const int base = (kSsimK * wave) * kSsimLW + kSsimC * lane;
    SsimStep first;
    ssim_load<0>(s_t, s_r, s_w, base, first);
    ssim_steps<0>(s_t, s_r, s_w, base, acc, first);
  }

#pragma unroll
  for (int k = 0; k < kSsimK; k++) {
    const int y = y0 + kSsimK * wave + k;
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < kSsimC; j++) {
      const int x = x0 + kSsimC * lane + j;
      if (x < a.w - 5 && y < a.h - 5) {
        const double aR = acc[j][k][0], aT = acc[j][k][1];
        const double vR = acc[j][k][2] - aR * aR;
        const double vT = acc[j][k][3] - aT * aT;
        const double cov = acc[j][k][4] - aR * aT;
        const double cs = (2 * cov + a.c2) / (vR + vT + a.c2);
        if (CALC_L) {
          const double l = (2 * aR * aT + a.c1) / (aR * aR + aT * aT + a.c1);
          v += l * cs;
        } else {
          v += cs;
        }
      }
    }
    v = wave_sum_f64(v);
    if (lane == 0 && y < a.h - 5) a.partials[(int64_t)tx * a.h + y] = v;
  }
}

// Row sums of every evaluation of a frame: grid (ceil(max_h / 256), evaluations).
__global__ __launch_bounds__(256) void k_ssim_rows(SsimRowsJobs j) {
  const int e = blockIdx.y;
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = j.h[e];
  if (y >= h) return;
  const double* part = j.part[e];
  const int tiles_x = j.tiles_x[e];
  double s = 0.0;
  if (y >= 5 && y < h - 5)
    for (int t = 0; t < tiles_x; t++) s += part[(int64_t)t * h + y];
  j.rows[e][y] = s;
}

// xPixelOpsSTD::DownsampleHV (libs/pmbb-core/src/xPixelOpsSTD.cpp:50-67): dst = (a + b + c + d + 2) >> 2
// over 2x2 blocks of the source view; dst is dw x dh uint16 (stride dw). grid.z = plane.
__global__ __launch_bounds__(256) void k_downsample(DownsampleJobs j) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)j.dw * j.dh) return;
  const int y = (int)(i / j.dw);
  const int x = (int)(i - (int64_t)y * j.dw);
  const ViewBuf S(j.src[blockIdx.z]);
  const uint32_t v = S.at(2 * x, 2 * y) + S.at(2 * x + 1, 2 * y) + S.at(2 * x, 2 * y + 1) + S.at(2 * x + 1, 2 * y + 1);
  j.dst[blockIdx.z][i] = (uint16_t)((v + 2) >> 2);
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

hipError_t launch_ssim_rows(const SsimRowsJobs& j, hipStream_t stream) {
  if (j.n <= 0 || j.max_h <= 0) return hipSuccess;
  if (j.n > kMaxSsimEvals) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ssim_rows, dim3((unsigned)((j.max_h + 255) / 256), (unsigned)j.n), dim3(256), 0, stream, j);
  return hipGetLastError();
}

int32_t ssim_tiles_x(int32_t w) { return w > 10 ? (w - 10 + kSsimTW - 1) / kSsimTW : 0; }

hipError_t launch_downsample(const DownsampleJobs& j, hipStream_t stream) {
  const int64_t n = (int64_t)j.dw * j.dh;
  if (n <= 0 || j.n <= 0) return hipSuccess;
  if (j.n > 6) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_downsample, dim3((unsigned)((n + 255) / 256), 1, (unsigned)j.n), dim3(256), 0, stream, j);
  return hipGetLastError();
}

}  // namespace qmiv
