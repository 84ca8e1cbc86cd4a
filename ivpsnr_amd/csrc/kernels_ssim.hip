// gfx950 kernels of the SSIM family: windowed moments (SSIM, MS-SSIM scales, IV-SSIM), the
// integer 2x2 pyramid of MS-SSIM and the per-row reduction of tile partial sums.
#include <utility>

#include "device_common.h"
#include "qmiv_internal.h"

namespace qmiv {
namespace {

// Tile shape: each lane owns one output column x kSsimK output rows; 4 waves stack their rows. A
// loaded window sample feeds up to kSsimK pels.
#ifndef QMIV_SSIM_ROWS
#define QMIV_SSIM_ROWS 4
#endif
constexpr int kSsimK = QMIV_SSIM_ROWS;
constexpr int kSsimTW = 64;                          // output columns per workgroup
constexpr int kSsimTH = 4 * kSsimK;                  // output rows per workgroup
constexpr int kSsimLW = (kSsimTW + 9 + 1) & ~1;      // window columns dx in [-5, 4], even row stride
constexpr int kSsimLH = kSsimTH + 9;

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
  // The same with the sample size known at compile time (no branch around the load).
  template <int BYTES>
  __device__ __forceinline__ uint32_t at_t(int x, int y) const {
    const uint32_t i = (uint32_t)((y >> shift) * stride + (x >> shift));
    if constexpr (BYTES == 2) return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, 0, 0);
    else return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, i, 0, 0);
  }
};

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// The window's 21 distinct weights are kept in scalar registers for the whole kernel: tap (dy, dx)
// uses weight sym_id(dy, dx) (the symmetric-window helpers of device_common.h).

// The window is walked as steps (R, u): loaded window row R = 0 .. K + 8 of the lane's K output
// rows and column group u of the window: u = 0 the column of offset 0 (window column 5), u = 1 .. 4
// the two columns of offsets -u and +u (window columns 5 - u and 5 + u, one weight for both), u = 5
// the column of offset -5 (window column 0). Window row R feeds output row k with window row
// dy = R - k, k over max(0, R - 9) .. min(K - 1, R). A pair's two samples are added first, their
// squares and cross products with one multiply and one fused multiply-add: samples and products are
// integers below 2^53, so these sums are exact and each moment is the reference's sum of weighted
// terms (xStructSim.cpp:15-60) regrouped by equal weight, a rounding-level change.
constexpr int kSsimGroups = 6;
__host__ __device__ constexpr bool grp_pair(int u) { return u >= 1 && u <= 4; }
constexpr int kSsimSteps = (kSsimK + 9) * kSsimGroups;

// One step's operands: the samples of the group's one or two columns.
struct SsimStep {
  float t[2], r[2];
};

template <int S>
__device__ __forceinline__ void ssim_load(const float* s_t, const float* s_r, int base, SsimStep& d) {
  constexpr int R = S / kSsimGroups, u = S % kSsimGroups;
  const int o = base + R * kSsimLW;
  d.t[0] = s_t[o + 5 - u];
  d.r[0] = s_r[o + 5 - u];
  if constexpr (grp_pair(u)) {
    d.t[1] = s_t[o + 5 + u];
    d.r[1] = s_r[o + 5 + u];
  }
}

template <int S>
__device__ __forceinline__ void ssim_fma(const SsimStep& d, const double (&wv)[kSsimSym], double (&acc)[kSsimK][5]) {
  constexpr int R = S / kSsimGroups, u = S % kSsimGroups;
  constexpr int k_lo = R > 9 ? R - 9 : 0;
  constexpr int k_hi = R < kSsimK - 1 ? R : kSsimK - 1;
  double tv = (double)d.t[0], rv = (double)d.r[0];
  double rr = rv * rv, tt = tv * tv, rt = rv * tv;
  if constexpr (grp_pair(u)) {
    const double tb = (double)d.t[1], rb = (double)d.r[1];
    rr = __builtin_fma(rb, rb, rr);
    tt = __builtin_fma(tb, tb, tt);
    rt = __builtin_fma(rb, tb, rt);
    rv += rb;
    tv += tb;
  }
#pragma unroll
  for (int k = k_lo; k <= k_hi; k++) {
    const double w = wv[sym_id(R - k, 5 - u)];  // F[R - k][5 - u] = F[R - k][5 + u]
    acc[k][0] = __builtin_fma(rv, w, acc[k][0]);
    acc[k][1] = __builtin_fma(tv, w, acc[k][1]);
    acc[k][2] = __builtin_fma(rr, w, acc[k][2]);
    acc[k][3] = __builtin_fma(tt, w, acc[k][3]);
    acc[k][4] = __builtin_fma(rt, w, acc[k][4]);
  }
  // Passing the step's accumulators through an ordered no-op pins its FMAs in place: the body is
  // one basic block, and without these the instruction DAG would hoist every load and product above
  // every FMA (the memory clobber keeps the next loads below).
#pragma unroll
  for (int k = k_lo; k <= k_hi; k++)
    asm volatile("" : "+v"(acc[k][0]), "+v"(acc[k][1]), "+v"(acc[k][2]), "+v"(acc[k][3]), "+v"(acc[k][4])::"memory");
  __builtin_amdgcn_sched_barrier(0);
}

// Software pipeline over the steps: the operands of step S + 1 are loaded from LDS before the FMAs
// of step S, so the LDS latency of every step overlaps the previous step's arithmetic.
template <int S>
__device__ __forceinline__ void ssim_steps(const float* s_t, const float* s_r, int base, const double (&wv)[kSsimSym],
                                           double (&acc)[kSsimK][5], const SsimStep& cur) {
  if constexpr (S + 1 < kSsimSteps) {
    SsimStep nxt;
    ssim_load<S + 1>(s_t, s_r, base, nxt);
    ssim_fma<S>(cur, wv, acc);
    ssim_steps<S + 1>(s_t, s_r, base, wv, acc, nxt);
  } else {
    ssim_fma<S>(cur, wv, acc);
  }
}

// The tile fill of one plane pair with their sample size BYTES known: every load of a thread issued
// before its first LDS store (one load latency per tile instead of one per element; a load under a
// branch on the sample size is waited for at the branch's end). Loads past the tile's last element
// read that element again and are not stored.
template <int BYTES>
__device__ __forceinline__ void pair_fill(const ViewBuf& T, const ViewBuf& R, int x0, int y0, int w, int h,
                                          float* s_t, float* s_r) {
  constexpr int N = kSsimLH * kSsimLW, IT = (N + 255) / 256;
  uint32_t v[IT][2];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const int i = min((int)threadIdx.x + 256 * j, N - 1);
    const int ly = i / kSsimLW;
    const int lx = i - ly * kSsimLW;
    const int gy = min(y0 - 5 + ly, h - 1);
    const int gx = min(x0 - 5 + lx, w - 1);
    v[j][0] = T.at_t<BYTES>(gx, gy);
    v[j][1] = R.at_t<BYTES>(gx, gy);
  }
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const int i = (int)threadIdx.x + 256 * j;
    if (i < N) {
      s_t[i] = (float)v[j][0];
      s_r[i] = (float)v[j][1];
    }
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
// One tile (kSsimTW x kSsimTH output pels) of one evaluation: stage the window, accumulate the
// moments, write the tile's per-row partial sums.
__device__ __forceinline__ void ssim_tile(const PlaneView& tst, const PlaneView& ref, int32_t w, int32_t h,
                                          bool calc_l, double c1, double c2, const double* f, double* partials,
                                          int tx, int ty, float* s_t, float* s_r) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int x0 = 5 + tx * kSsimTW;  // first output column of the tile
  const int y0 = 5 + ty * kSsimTH;
  const ViewBuf T(tst), R(ref);

  // The 21 distinct weights, read once from the kernel arguments and held in scalar registers.
  double wv[kSsimSym];
#pragma unroll
  for (int u = 0; u < kSsimSym; u++) {
    wv[u] = f[sym_rep(u)];
    asm volatile("" : "+s"(wv[u]));
  }
  // Stage the window rows [y0 - 5, y0 + TH + 4) x [x0 - 5, x0 + TW + 4), clamped inside the
  // picture (clamped values only feed pels outside the interior, which are discarded).
  if (T.bytes == 2 && R.bytes == 2) pair_fill<2>(T, R, x0, y0, w, h, s_t, s_r);
  else if (T.bytes == 1 && R.bytes == 1) pair_fill<1>(T, R, x0, y0, w, h, s_t, s_r);
  else
    for (int i = threadIdx.x; i < kSsimLH * kSsimLW; i += 256) {
      const int ly = i / kSsimLW;
      const int lx = i - ly * kSsimLW;
      const int gy = min(y0 - 5 + ly, h - 1);
      const int gx = min(x0 - 5 + lx, w - 1);
      s_t[i] = (float)T.at(gx, gy);
      s_r[i] = (float)R.at(gx, gy);
    }
  __syncthreads();

  double acc[kSsimK][5];
#pragma unroll
  for (int k = 0; k < kSsimK; k++)
#pragma unroll
    for (int m = 0; m < 5; m++) acc[k][m] = 0.0;

  // Every step as its own compile-time instance (the loop unroller gives up on this many
  // instructions): straight-line code, constant LDS offsets, no branches.
  {
    const int base = (kSsimK * wave) * kSsimLW + lane;
    SsimStep first;
    ssim_load<0>(s_t, s_r, base, first);
    ssim_steps<0>(s_t, s_r, base, wv, acc, first);
  }

#pragma unroll
  for (int k = 0; k < kSsimK; k++) {
    const int y = y0 + kSsimK * wave + k;
    double v = 0.0;
    const int x = x0 + lane;
    if (x < w - 5 && y < h - 5) {
      const double aR = acc[k][0], aT = acc[k][1];
      const double vR = acc[k][2] - aR * aR;
      const double vT = acc[k][3] - aT * aT;
      const double cov = acc[k][4] - aR * aT;
      const double cs = (2 * cov + c2) / (vR + vT + c2);
      if (calc_l) {
        const double l = (2 * aR * aT + c1) / (aR * aR + aT * aT + c1);
        v = l * cs;
      } else {
        v = cs;
      }
    }
    v = wave_sum_f64(v);
    if (lane == 0 && y < h - 5) partials[(int64_t)tx * h + y] = v;
  }
}

template <int CALC_L>
__global__ __launch_bounds__(256) void k_ssim(SsimArgs a) {
  __shared__ float s_t[kSsimLH * kSsimLW];
  __shared__ float s_r[kSsimLH * kSsimLW];
  const int t = xcd_order(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const int ty = t / gridDim.x, tx = t - ty * gridDim.x;
  ssim_tile(a.tst, a.ref, a.w, a.h, CALC_L != 0, a.c1, a.c2, a.f, a.partials, tx, ty, s_t, s_r);
}

// Many small evaluations (the MS-SSIM pyramid levels of all components) in one launch: the
// workgroups are the evaluations' tiles back to back, job j owning tiles [tile0, tile0 + tiles).
__global__ __launch_bounds__(256) void k_ssim_batch(SsimBatchArgs a) {
  __shared__ float s_t[kSsimLH * kSsimLW];
  __shared__ float s_r[kSsimLH * kSsimLW];
  const int t = xcd_order(blockIdx.x, gridDim.x);
  int j = 0;
  while (j + 1 < a.n && t >= a.job[j + 1].tile0) j++;
  const SsimBatchJob& J = a.job[j];
  const int local = t - J.tile0;
  const int ty = local / J.tiles_x;
  const int tx = local - ty * J.tiles_x;
  ssim_tile(J.tst, J.ref, J.w, J.h, J.calc_l != 0, a.c1, a.c2, a.f, J.partials, tx, ty, s_t, s_r);
}

// ---------------------------------------------------------------------------------------------
// Several full-resolution evaluations of one component in one pass. SSIM(P0, P1) with L·CS, its
// CS alone (MS-SSIM level 0, or SSIM without L), IV-SSIM's SSIM(P0, SCP1) and SSIM(P1, SCP0)
// (xSSIM.cpp:54-66,67-87,192-212) share window moments: μ and E[x²] of P0 and P1 serve all of
// them, so the set needs 5 + 3 + 3 moment sums where the evaluations one by one need 5 each. Every
// moment is accumulated with the same fused multiply-adds in the same order as in k_ssim, so each
// output is bit-identical to its separate evaluation. Outputs with a null partials pointer are
// skipped; HAS_C / HAS_D add the SCP1 / SCP0 moments.
// ---------------------------------------------------------------------------------------------
template <int KM>
struct MultiTile {
  static constexpr int TH = 4 * KM;
  static constexpr int LH = TH + 9;
  static constexpr int Steps = (KM + 9) * kSsimGroups;
};

// The samples of one window column of each staged picture: a step's first column is loaded a step
// ahead, a pair's second column at the start of its own step (behind the first column's products).
template <bool HAS_C, bool HAS_D>
struct MultiStep {
  float p0, p1, sc, sd;
};

template <bool HAS_C, bool HAS_D>
__device__ __forceinline__ void multi_col(const float* s0, const float* s1, const float* sc, const float* sd, int o,
                                          MultiStep<HAS_C, HAS_D>& d) {
  d.p0 = s0[o];
  d.p1 = s1[o];
  if (HAS_C) d.sc = sc[o];
  if (HAS_D) d.sd = sd[o];
}

template <int KM, bool HAS_C, bool HAS_D, int S>
__device__ __forceinline__ void multi_load(const float* s0, const float* s1, const float* sc, const float* sd,
                                           int base, MultiStep<HAS_C, HAS_D>& d) {
  constexpr int R = S / kSsimGroups, u = S % kSsimGroups;
  multi_col<HAS_C, HAS_D>(s0, s1, sc, sd, base + R * kSsimLW + 5 - u, d);
}

// Moment slots per output row k: 0 Σw·P0, 1 Σw·P1, 2 Σw·P0², 3 Σw·P1², 4 Σw·P0·P1, then for SCP1
// 5 Σw·S1, 6 Σw·S1², 7 Σw·S1·P0, then for SCP0 8 (or 5) Σw·S0, Σw·S0², Σw·S0·P1.
template <bool HAS_C, bool HAS_D>
constexpr int multi_moments() { return 5 + (HAS_C ? 3 : 0) + (HAS_D ? 3 : 0); }

template <int KM, bool HAS_C, bool HAS_D, int S>
__device__ __forceinline__ void multi_fma(const float* s0, const float* s1, const float* sc, const float* sd, int base,
                                          const MultiStep<HAS_C, HAS_D>& d, const double (&wv)[kSsimSym],
                                          double (&acc)[KM][multi_moments<HAS_C, HAS_D>()]) {
  constexpr int R = S / kSsimGroups, u = S % kSsimGroups;
  constexpr int k_lo = R > 9 ? R - 9 : 0;
  constexpr int k_hi = R < KM - 1 ? R : KM - 1;
  constexpr int oD = HAS_C ? 8 : 5;
  MultiStep<HAS_C, HAS_D> e;
  if constexpr (grp_pair(u)) multi_col<HAS_C, HAS_D>(s0, s1, sc, sd, base + R * kSsimLW + 5 + u, e);
  // The column values (and, for a pair, the second column's), then each moment's value formed just
  // before its fused multiply-adds, so that only the column values stay live across the step.
  const double a0 = (double)d.p0, a1 = (double)d.p1;
  const double ac = HAS_C ? (double)d.sc : 0.0, ad = HAS_D ? (double)d.sd : 0.0;
  double b0 = 0, b1 = 0, bc = 0, bd = 0;
  if constexpr (grp_pair(u)) {
    b0 = (double)e.p0;
    b1 = (double)e.p1;
    if (HAS_C) bc = (double)e.sc;
    if (HAS_D) bd = (double)e.sd;
  }
  constexpr bool P = grp_pair(u);
  // Sample sum and product sum of the group: x + x' and x y + x' y' (exact).
  auto sum1 = [&](double x, double xp) { return P ? x + xp : x; };
  auto sum2 = [&](double x, double y, double xp, double yp) { return P ? __builtin_fma(xp, yp, x * y) : x * y; };
  auto put = [&](int m, double v) {
#pragma unroll
    for (int k = k_lo; k <= k_hi; k++) acc[k][m] = __builtin_fma(v, wv[sym_id(R - k, 5 - u)], acc[k][m]);
  };
  // k_ssim's rr, tt, rt as (ref, tst) = (P1, P0); (SCP1, P0); (SCP0, P1)
  put(0, sum1(a0, b0));
  put(1, sum1(a1, b1));
  put(2, sum2(a0, a0, b0, b0));
  put(3, sum2(a1, a1, b1, b1));
  put(4, sum2(a1, a0, b1, b0));
  if (HAS_C) {
    put(5, sum1(ac, bc));
    put(6, sum2(ac, ac, bc, bc));
    put(7, sum2(ac, a0, bc, b0));
  }
  if (HAS_D) {
    put(oD, sum1(ad, bd));
    put(oD + 1, sum2(ad, ad, bd, bd));
    put(oD + 2, sum2(ad, a1, bd, b1));
  }
#pragma unroll
  for (int k = k_lo; k <= k_hi; k++)
#pragma unroll
    for (int m = 0; m < multi_moments<HAS_C, HAS_D>(); m++) asm volatile("" : "+v"(acc[k][m])::"memory");
  __builtin_amdgcn_sched_barrier(0);
}

template <int KM, bool HAS_C, bool HAS_D, int S>
__device__ __forceinline__ void multi_steps(const float* s0, const float* s1, const float* sc, const float* sd,
                                            int base, const double (&wv)[kSsimSym],
                                            double (&acc)[KM][multi_moments<HAS_C, HAS_D>()],
                                            const MultiStep<HAS_C, HAS_D>& cur) {
  if constexpr (S + 1 < MultiTile<KM>::Steps) {
    MultiStep<HAS_C, HAS_D> nxt;
    multi_load<KM, HAS_C, HAS_D, S + 1>(s0, s1, sc, sd, base, nxt);
    multi_fma<KM, HAS_C, HAS_D, S>(s0, s1, sc, sd, base, cur, wv, acc);
    multi_steps<KM, HAS_C, HAS_D, S + 1>(s0, s1, sc, sd, base, wv, acc, nxt);
  } else {
    multi_fma<KM, HAS_C, HAS_D, S>(s0, s1, sc, sd, base, cur, wv, acc);
  }
}

// One output of the set from its (ref, tst) moments, as k_ssim's epilogue computes it.
__device__ __forceinline__ double ssim_value(double aR, double aT, double eRR, double eTT, double eRT, double c1,
                                             double c2, bool calc_l) {
  const double vR = eRR - aR * aR;
  const double vT = eTT - aT * aT;
  const double cov = eRT - aR * aT;
  const double cs = (2 * cov + c2) / (vR + vT + c2);
  if (!calc_l) return cs;
  const double l = (2 * aR * aT + c1) / (aR * aR + aT * aT + c1);
  return l * cs;
}

// The tile fill of k_ssim_multi with the pictures' sample size BYTES known (the shift-compensated
// planes hold 16-bit samples): every load of a thread is issued before its first LDS store, so a
// workgroup waits for one load latency per tile instead of one per element (a load under a branch
// on the sample size is waited for at the branch's end). Loads past the tile's last element read
// that element again and are not stored.
template <int LH, bool HAS_C, bool HAS_D, int BYTES>
__device__ __forceinline__ void multi_fill(const ViewBuf& P0, const ViewBuf& P1, const ViewBuf& SC, const ViewBuf& SD,
                                           int x0, int y0, int w, int h, float* s0, float* s1, float* sc, float* sd) {
  constexpr int N = LH * kSsimLW, IT = (N + 255) / 256;
  uint32_t v[IT][4];
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const int i = min((int)threadIdx.x + 256 * j, N - 1);
    const int ly = i / kSsimLW;
    const int lx = i - ly * kSsimLW;
    const int gy = min(y0 - 5 + ly, h - 1);
    const int gx = min(x0 - 5 + lx, w - 1);
    v[j][0] = P0.at_t<BYTES>(gx, gy);
    v[j][1] = P1.at_t<BYTES>(gx, gy);
    if constexpr (HAS_C) v[j][2] = SC.at_t<2>(gx, gy);
    if constexpr (HAS_D) v[j][3] = SD.at_t<2>(gx, gy);
  }
#pragma unroll
  for (int j = 0; j < IT; j++) {
    const int i = (int)threadIdx.x + 256 * j;
    if (i < N) {
      s0[i] = (float)v[j][0];
      s1[i] = (float)v[j][1];
      if constexpr (HAS_C) sc[i] = (float)v[j][2];
      if constexpr (HAS_D) sd[i] = (float)v[j][3];
    }
  }
}

template <int KM, bool HAS_C, bool HAS_D>
__global__ __launch_bounds__(256) void k_ssim_multi(SsimMultiArgs a) {
  using Tile = MultiTile<KM>;
  constexpr int NM = multi_moments<HAS_C, HAS_D>();
  constexpr int oD = HAS_C ? 8 : 5;
  __shared__ float s0[Tile::LH * kSsimLW];
  __shared__ float s1[Tile::LH * kSsimLW];
  __shared__ float sc[HAS_C ? Tile::LH * kSsimLW : 1];
  __shared__ float sd[HAS_D ? Tile::LH * kSsimLW : 1];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int t = xcd_order(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const int ty = t / gridDim.x;
  const int tx = t - ty * gridDim.x;
  const int x0 = 5 + tx * kSsimTW;
  const int y0 = 5 + ty * Tile::TH;
  const ViewBuf P0(a.p0), P1(a.p1), SC(a.scp1), SD(a.scp0);

  double wv[kSsimSym];
#pragma unroll
  for (int u = 0; u < kSsimSym; u++) {
    wv[u] = a.f[sym_rep(u)];
    asm volatile("" : "+s"(wv[u]));
  }
  if (P0.bytes == 2 && P1.bytes == 2 && (!HAS_C || SC.bytes == 2) && (!HAS_D || SD.bytes == 2))
    multi_fill<Tile::LH, HAS_C, HAS_D, 2>(P0, P1, SC, SD, x0, y0, a.w, a.h, s0, s1, sc, sd);
  else if (P0.bytes == 1 && P1.bytes == 1 && (!HAS_C || SC.bytes == 2) && (!HAS_D || SD.bytes == 2))
    multi_fill<Tile::LH, HAS_C, HAS_D, 1>(P0, P1, SC, SD, x0, y0, a.w, a.h, s0, s1, sc, sd);
  else
    for (int i = threadIdx.x; i < Tile::LH * kSsimLW; i += 256) {
      const int ly = i / kSsimLW;
      const int lx = i - ly * kSsimLW;
      const int gy = min(y0 - 5 + ly, a.h - 1);
      const int gx = min(x0 - 5 + lx, a.w - 1);
      s0[i] = (float)P0.at(gx, gy);
      s1[i] = (float)P1.at(gx, gy);
      if (HAS_C) sc[i] = (float)SC.at(gx, gy);
      if (HAS_D) sd[i] = (float)SD.at(gx, gy);
    }
  __syncthreads();

  double acc[KM][NM];
#pragma unroll
  for (int k = 0; k < KM; k++)
#pragma unroll
    for (int m = 0; m < NM; m++) acc[k][m] = 0.0;
  {
    const int base = (KM * wave) * kSsimLW + lane;
    MultiStep<HAS_C, HAS_D> first;
    multi_load<KM, HAS_C, HAS_D, 0>(s0, s1, sc, sd, base, first);
    multi_steps<KM, HAS_C, HAS_D, 0>(s0, s1, sc, sd, base, wv, acc, first);
  }

  const int x = x0 + lane;
#pragma unroll
  for (int k = 0; k < KM; k++) {
    const int y = y0 + KM * wave + k;
    const bool in = x < a.w - 5 && y < a.h - 5;
    const int64_t o = (int64_t)tx * a.h + y;
    // (ref, tst) = (P1, P0)
    if (a.part_lcs) {
      double v = in ? ssim_value(acc[k][1], acc[k][0], acc[k][3], acc[k][2], acc[k][4], a.c1, a.c2, true) : 0.0;
      v = wave_sum_f64(v);
      if (lane == 0 && y < a.h - 5) a.part_lcs[o] = v;
    }
    if (a.part_cs) {
      double v = in ? ssim_value(acc[k][1], acc[k][0], acc[k][3], acc[k][2], acc[k][4], a.c1, a.c2, false) : 0.0;
      v = wave_sum_f64(v);
      if (lane == 0 && y < a.h - 5) a.part_cs[o] = v;
    }
    if (HAS_C) {  // (ref, tst) = (SCP1, P0)
      double v = in ? ssim_value(acc[k][5], acc[k][0], acc[k][6], acc[k][2], acc[k][7], a.c1, a.c2, true) : 0.0;
      v = wave_sum_f64(v);
      if (lane == 0 && y < a.h - 5) a.part_c[o] = v;
    }
    if (HAS_D) {  // (ref, tst) = (SCP0, P1)
      double v = in ? ssim_value(acc[k][oD], acc[k][1], acc[k][oD + 1], acc[k][3], acc[k][oD + 2], a.c1, a.c2, true)
                    : 0.0;
      v = wave_sum_f64(v);
      if (lane == 0 && y < a.h - 5) a.part_d[o] = v;
    }
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

// All four MS-SSIM pyramid levels of up to six planes in one launch. A workgroup owns a 16 x 16
// block of level 1 (a 32 x 32 block of the source view) and the 8 x 8, 4 x 4 and 2 x 2 blocks of
// levels 2-4 below it, each level from the previous one in LDS with DownsampleHV's rounding
// (libs/pmbb-core/src/xPixelOpsSTD.cpp:50-67), so the cascade is the same as level by level. A level
// pel is written when it lies inside that level's (W >> l) x (H >> l); its inputs then always lie
// inside the previous level. grid.z = plane.
__global__ __launch_bounds__(256) void k_pyramid(PyramidJobs j) {
  __shared__ uint32_t l1[16][17], l2[8][9], l3[4][5];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int p = blockIdx.z;
  const ViewBuf S(j.src[p]);
  const int x1 = blockIdx.x * 16 + tx, y1 = blockIdx.y * 16 + ty;
  const int w1 = j.w >> 1, h1 = j.h >> 1;
  uint32_t v = 0;
  if (x1 < w1 && y1 < h1) {
    v = (S.at(2 * x1, 2 * y1) + S.at(2 * x1 + 1, 2 * y1) + S.at(2 * x1, 2 * y1 + 1) + S.at(2 * x1 + 1, 2 * y1 + 1) + 2) >> 2;
    j.dst[p][0][(int64_t)y1 * w1 + x1] = (uint16_t)v;
  }
  l1[ty][tx] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    const int u = threadIdx.x & 7, t = threadIdx.x >> 3;
    const int x2 = blockIdx.x * 8 + u, y2 = blockIdx.y * 8 + t;
    const int w2 = j.w >> 2, h2 = j.h >> 2;
    const uint32_t v2 = (l1[2 * t][2 * u] + l1[2 * t][2 * u + 1] + l1[2 * t + 1][2 * u] + l1[2 * t + 1][2 * u + 1] + 2) >> 2;
    l2[t][u] = v2;
    if (x2 < w2 && y2 < h2) j.dst[p][1][(int64_t)y2 * w2 + x2] = (uint16_t)v2;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int u = threadIdx.x & 3, t = threadIdx.x >> 2;
    const int x3 = blockIdx.x * 4 + u, y3 = blockIdx.y * 4 + t;
    const int w3 = j.w >> 3, h3 = j.h >> 3;
    const uint32_t v3 = (l2[2 * t][2 * u] + l2[2 * t][2 * u + 1] + l2[2 * t + 1][2 * u] + l2[2 * t + 1][2 * u + 1] + 2) >> 2;
    l3[t][u] = v3;
    if (x3 < w3 && y3 < h3) j.dst[p][2][(int64_t)y3 * w3 + x3] = (uint16_t)v3;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int u = threadIdx.x & 1, t = threadIdx.x >> 1;
    const int x4 = blockIdx.x * 2 + u, y4 = blockIdx.y * 2 + t;
    const int w4 = j.w >> 4, h4 = j.h >> 4;
    const uint32_t v4 = (l3[2 * t][2 * u] + l3[2 * t][2 * u + 1] + l3[2 * t + 1][2 * u] + l3[2 * t + 1][2 * u + 1] + 2) >> 2;
    if (x4 < w4 && y4 < h4) j.dst[p][3][(int64_t)y4 * w4 + x4] = (uint16_t)v4;
  }
}

}  // namespace

hipError_t launch_pyramid(const PyramidJobs& j, hipStream_t stream) {
  if (j.n <= 0 || j.n > 6 || (j.w >> 1) <= 0 || (j.h >> 1) <= 0) return j.n > 6 ? hipErrorInvalidValue : hipSuccess;
  const dim3 grid((unsigned)(((j.w >> 1) + 15) / 16), (unsigned)(((j.h >> 1) + 15) / 16), (unsigned)j.n);
  hipLaunchKernelGGL(k_pyramid, grid, dim3(256), 0, stream, j);
  return hipGetLastError();
}

hipError_t launch_ssim(const SsimArgs& a, hipStream_t stream) {
  for (int i = 0; i < 100; i++)  // the kernel holds only the 21 distinct weights of a symmetric window
    if (a.f[i] != a.f[sym_rep(sym_id(i / 10, i % 10))]) return hipErrorInvalidValue;
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

#ifndef QMIV_SSIM_MULTI_ROWS
#define QMIV_SSIM_MULTI_ROWS 4
#endif
constexpr int kMultiK = QMIV_SSIM_MULTI_ROWS;

hipError_t launch_ssim_multi(const SsimMultiArgs& a, hipStream_t stream) {
  for (int i = 0; i < 100; i++)
    if (a.f[i] != a.f[sym_rep(sym_id(i / 10, i % 10))]) return hipErrorInvalidValue;
  const int tiles_x = a.w > 10 ? (a.w - 10 + kSsimTW - 1) / kSsimTW : 0;
  const int tiles_y = a.h > 10 ? (a.h - 10 + MultiTile<kMultiK>::TH - 1) / MultiTile<kMultiK>::TH : 0;
  if (tiles_x != a.tiles_x || (!a.part_lcs && !a.part_cs)) return hipErrorInvalidValue;
  if (tiles_x == 0 || tiles_y == 0) return hipSuccess;
  const dim3 grid(tiles_x, tiles_y), block(256);
  const bool c = a.part_c != nullptr, d = a.part_d != nullptr;
  if (c && d) hipLaunchKernelGGL((k_ssim_multi<kMultiK, true, true>), grid, block, 0, stream, a);
  else if (c) hipLaunchKernelGGL((k_ssim_multi<kMultiK, true, false>), grid, block, 0, stream, a);
  else if (d) hipLaunchKernelGGL((k_ssim_multi<kMultiK, false, true>), grid, block, 0, stream, a);
  else hipLaunchKernelGGL((k_ssim_multi<kMultiK, false, false>), grid, block, 0, stream, a);
  return hipGetLastError();
}

int32_t ssim_tiles_y(int32_t h) { return h > 10 ? (h - 10 + kSsimTH - 1) / kSsimTH : 0; }

hipError_t launch_ssim_batch(SsimBatchArgs& a, hipStream_t stream) {
  for (int i = 0; i < 100; i++)
    if (a.f[i] != a.f[sym_rep(sym_id(i / 10, i % 10))]) return hipErrorInvalidValue;
  if (a.n <= 0 || a.n > kMaxSsimBatch) return hipErrorInvalidValue;
  int32_t tiles = 0;
  for (int j = 0; j < a.n; j++) {
    SsimBatchJob& J = a.job[j];
    if (J.tiles_x != ssim_tiles_x(J.w)) return hipErrorInvalidValue;
    J.tile0 = tiles;
    tiles += J.tiles_x * ssim_tiles_y(J.h);
  }
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ssim_batch, dim3((unsigned)tiles), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_ssim_rows(const SsimRowsJobs& j, hipStream_t stream) {
  if (j.n <= 0 || j.max_h <= 0) return hipSuccess;
  if (j.n > kMaxSsimEvals) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ssim_rows, dim3((unsigned)((j.max_h + 255) / 256), (unsigned)j.n), dim3(256), 0, stream, j);
  return hipGetLastError();
}

int32_t ssim_tiles_x(int32_t w) { return w > 10 ? (w - 10 + kSsimTW - 1) / kSsimTW : 0; }


}  // namespace qmiv
