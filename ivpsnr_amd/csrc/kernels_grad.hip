// gfx950 kernels of the differentiable SSIM family (float pictures): the forward windowed moments
// with the per-pel partial derivatives of the SSIM map, the gradient correlation, the float 2x2
// pyramid and its adjoint, quantisation for the shift search, and the gather of the
// shift-compensated pictures with its adjoint scatter.
//
// SSIM map of one interior pel p (xStructSim<flt64, CalcL>::CalcPel,
// libs/pmbb-ivqm/src/xStructSim.cpp:15-60, on real-valued samples):
//   muT = sum w T, muR = sum w R, qT = sum w T^2, qR = sum w R^2, x = sum w T R
//   N = 2 (x - muT muR) + C2, D = (qT - muT^2) + (qR - muR^2) + C2, CS = N / D
//   L = (2 muT muR + C1) / (muT^2 + muR^2 + C1),  s = L CS (or CS alone)
// Its derivatives: ds/dmuT = A_T, ds/dmuR = A_R, ds/dqT = ds/dqR = B, ds/dx = Cm, with
//   dCS/dmuT = (2 muT N - 2 muR D) / D^2,  dCS/dq = -N / D^2,  dCS/dx = 2 / D,
//   dL/dmuT = (2 muR Ld - 2 muT Ln) / Ld^2   (and symmetric in T <-> R).
// A sample T(q) enters the windows of pels p = q - o with weight w(o), hence
//   dS/dT(q) = sum_o w(o) [A_T + 2 T(q) B + R(q) Cm](q - o) / (W H),
// a correlation of the maps with the flipped window (k_ssimf_grad).
#include <utility>

#include "device_common.h"
#include "qmiv_internal.h"

namespace qmiv {
namespace {

constexpr int kGW = 64;          // output columns per workgroup (one per lane)
constexpr int kGK = 4;           // output rows per lane
constexpr int kGH = 4 * kGK;     // output rows per workgroup
constexpr int kGLW = kGW + 9;    // window columns
constexpr int kGLH = kGH + 9;
// The window is walked as steps (R, U): loaded window row R = 0 .. K + 8 of the lane's K output
// rows, column group U of the symmetric window (device_common.h): U = 0 the centre column, U = 1 .. 4
// the two columns U either side of it (one weight for both: their values are added first and
// multiplied once), U = 5 the outer column. The centre is window column 5 for the window as stored
// (offsets -5 .. 4 from the output pel), column 4 for the flipped window F'[a][b] = F[9 - a][9 - b]
// of the gradient correlation. Window row R feeds output row k with window row dy = R - k, k over
// max(0, R - 9) .. min(K - 1, R). Each weight is read from the kernel arguments at a constant offset
// (a scalar register). Grouping equal weights changes the sums at the rounding level only.
template <bool FLIP>
__host__ __device__ constexpr int grp_col(int u, int side) {
  return u == 5 ? (FLIP ? 9 : 0) : (FLIP ? 4 : 5) + side * u;
}
template <bool FLIP>
__host__ __device__ constexpr int tap_id(int dy, int c) {
  return FLIP ? sym_rep(sym_id(9 - dy, 9 - c)) : sym_rep(sym_id(dy, c));
}

template <int NV, int R, int U, bool FLIP, typename Load>
__device__ __forceinline__ void win_group(const Load& load, const double* f, double (&acc)[kGK][NV]) {
  constexpr int k_lo = R > 9 ? R - 9 : 0;
  constexpr int k_hi = R < kGK - 1 ? R : kGK - 1;
  constexpr int c0 = grp_col<FLIP>(U, -1);
  double v[NV];
  load(R, c0, v);
  if constexpr (U >= 1 && U <= 4) {
    double v2[NV];
    load(R, grp_col<FLIP>(U, 1), v2);
#pragma unroll
    for (int m = 0; m < NV; m++) v[m] += v2[m];
  }
#pragma unroll
  for (int k = k_lo; k <= k_hi; k++) {
    const double w = f[tap_id<FLIP>(R - k, c0)];
#pragma unroll
    for (int m = 0; m < NV; m++) acc[k][m] = __builtin_fma(v[m], w, acc[k][m]);
  }
  // Keep this step's FMAs ahead of the next step's loads (see kernels_ssim.hip).
#pragma unroll
  for (int k = k_lo; k <= k_hi; k++)
#pragma unroll
    for (int m = 0; m < NV; m++) asm volatile("" : "+v"(acc[k][m])::"memory");
  __builtin_amdgcn_sched_barrier(0);
}

template <int NV, int R, bool FLIP, typename Load, int... U>
__device__ __forceinline__ void win_row(const Load& load, const double* f, double (&acc)[kGK][NV],
                                        std::integer_sequence<int, U...>) {
  (win_group<NV, R, U, FLIP>(load, f, acc), ...);
}

template <int NV, bool FLIP, typename Load, int... R>
__device__ __forceinline__ void win_rows(const Load& load, const double* f, double (&acc)[kGK][NV],
                                         std::integer_sequence<int, R...>) {
  (win_row<NV, R, FLIP>(load, f, acc, std::make_integer_sequence<int, 6>{}), ...);
}

__device__ __forceinline__ double ldsample(const void* p, int32_t f64, int64_t i) {
  return f64 ? ((const double*)p)[i] : (double)((const float*)p)[i];
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Tile fills issue every load of a thread before its first LDS store (kFillIt loads per value and
// thread, the last ones clamped to the tile's final element and not stored), so a workgroup waits for
// one global-load latency per tile rather than one per element.
constexpr int kFillN = kGLH * kGLW;
constexpr int kFillIt = (kFillN + 255) / 256;

// Both pictures' samples over rows [y0 - 5, y0 - 5 + kGLH) x columns [x0 - 5, x0 - 5 + kGLW),
// clamped to the picture, times their scales.
template <typename TT, typename TR>
__device__ __forceinline__ void fill_pictures(const SsimFArgs& a, int x0, int y0, double* s_t, double* s_r) {
  const TT* t = (const TT*)a.tst;
  const TR* r = (const TR*)a.ref;
  TT vt[kFillIt];
  TR vr[kFillIt];
#pragma unroll
  for (int j = 0; j < kFillIt; j++) {
    const int i = min((int)threadIdx.x + 256 * j, kFillN - 1);
    const int ly = i / kGLW;
    const int lx = i - ly * kGLW;
    const int64_t o = (int64_t)min(y0 - 5 + ly, a.h - 1) * a.w + min(x0 - 5 + lx, a.w - 1);
    vt[j] = t[o];
    vr[j] = r[o];
  }
#pragma unroll
  for (int j = 0; j < kFillIt; j++) {
    const int i = (int)threadIdx.x + 256 * j;
    if (i < kFillN) {
      s_t[i] = (double)vt[j] * a.scale_t;
      s_r[i] = (double)vr[j] * a.scale_r;
    }
  }
}

// Maps m0 and m0 + 1 over rows [y0 - 4, y0 - 4 + kGLH) x columns [x0 - 4, x0 - 4 + kGLW), zero outside
// the interior, into registers: the raw loads (a clamped in-range element outside the interior) and
// a bit per element for the interior, applied when the values are stored (a select right after
// the load would wait for it there).
struct MapRegs {
  double v[kFillIt][2];
  uint32_t in = 0;
};
__device__ __forceinline__ void load_maps(const SsimGArgs& a, int x0, int y0, int m0, MapRegs& r) {
  const int64_t plane = (int64_t)a.w * a.h;
  r.in = 0;
#pragma unroll
  for (int j = 0; j < kFillIt; j++) {
    const int i = min((int)threadIdx.x + 256 * j, kFillN - 1);
    const int ly = i / kGLW;
    const int lx = i - ly * kGLW;
    const int py = y0 - 4 + ly;
    const int px = x0 - 4 + lx;
    const bool in = py >= 5 && py < a.h - 5 && px >= 5 && px < a.w - 5;
    const int64_t o = in ? (int64_t)py * a.w + px : 0;
    r.in |= (uint32_t)in << j;
#pragma unroll
    for (int m = 0; m < 2; m++) r.v[j][m] = a.maps[(m0 + m) * plane + o];
  }
}

__device__ __forceinline__ void store_maps(const MapRegs& r, double (*s_m)[kFillN]) {
#pragma unroll
  for (int j = 0; j < kFillIt; j++) {
    const int i = (int)threadIdx.x + 256 * j;
    const bool in = (r.in >> j) & 1u;
    if (i < kFillN) {
      s_m[0][i] = in ? r.v[j][0] : 0.0;
      s_m[1][i] = in ? r.v[j][1] : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Forward: SSIM map row sums of one plane pair, and (GRAD) the maps A_T, A_R, B, Cm of every
// interior pel, divided by W H. Tile 64 x 16 over the interior, like the integer kernel.
// ---------------------------------------------------------------------------------------------
template <int CALC_L, int GRAD>
__global__ __launch_bounds__(256) void k_ssimf(SsimFArgs a) {
  __shared__ double s_t[kGLH * kGLW];
  __shared__ double s_r[kGLH * kGLW];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tx = blockIdx.x;
  const int x0 = 5 + tx * kGW;
  const int y0 = 5 + blockIdx.y * kGH;
  if (!a.tst_f64 && !a.ref_f64) fill_pictures<float, float>(a, x0, y0, s_t, s_r);
  else if (a.tst_f64 && a.ref_f64) fill_pictures<double, double>(a, x0, y0, s_t, s_r);
  else if (a.tst_f64) fill_pictures<double, float>(a, x0, y0, s_t, s_r);
  else fill_pictures<float, double>(a, x0, y0, s_t, s_r);
  __syncthreads();

  double acc[kGK][5];
#pragma unroll
  for (int k = 0; k < kGK; k++)
#pragma unroll
    for (int m = 0; m < 5; m++) acc[k][m] = 0.0;
  const int base = (kGK * wave) * kGLW + lane;
  auto load = [&](int R, int c, double (&v)[5]) {
    const double tv = s_t[base + R * kGLW + c];
    const double rv = s_r[base + R * kGLW + c];
    v[0] = rv;
    v[1] = tv;
    v[2] = rv * rv;
    v[3] = tv * tv;
    v[4] = rv * tv;
  };
  win_rows<5, false>(load, a.f, acc, std::make_integer_sequence<int, kGK + 9>{});

  const int x = x0 + lane;
  const double inv_area = 1.0 / ((double)a.w * (double)a.h);
#pragma unroll
  for (int k = 0; k < kGK; k++) {
    const int y = y0 + kGK * wave + k;
    double sv = 0.0;
    if (x < a.w - 5 && y < a.h - 5) {
      const double muR = acc[k][0], muT = acc[k][1];
      const double vR = acc[k][2] - muR * muR;
      const double vT = acc[k][3] - muT * muT;
      const double cov = acc[k][4] - muR * muT;
      const double N = 2 * cov + a.c2;
      const double D = vR + vT + a.c2;
      const double cs = N / D;
      double L = 1.0, Ln = 0.0, Ld = 1.0;
      if (CALC_L) {
        Ln = 2 * muR * muT + a.c1;
        Ld = muR * muR + muT * muT + a.c1;
        L = Ln / Ld;
      }
      sv = L * cs;
      if (GRAD) {
        const double D2 = D * D;
        const double dcs_dmuT = (2 * muT * N - 2 * muR * D) / D2;
        const double dcs_dmuR = (2 * muR * N - 2 * muT * D) / D2;
        double aT = L * dcs_dmuT, aR = L * dcs_dmuR;
        if (CALC_L) {
          const double Ld2 = Ld * Ld;
          aT += (2 * muR * Ld - 2 * muT * Ln) / Ld2 * cs;
          aR += (2 * muT * Ld - 2 * muR * Ln) / Ld2 * cs;
        }
        const int64_t o = (int64_t)y * a.w + x;
        const int64_t plane = (int64_t)a.w * a.h;
        a.maps[o] = aT * inv_area;
        a.maps[plane + o] = aR * inv_area;
        a.maps[2 * plane + o] = -L * N / D2 * inv_area;
        a.maps[3 * plane + o] = 2 * L / D * inv_area;
      }
    }
    sv = wave_sum_d(sv);
    if (lane == 0 && y < a.h - 5) a.partials[(int64_t)tx * a.h + y] = sv;
  }
}

// ---------------------------------------------------------------------------------------------
// Gradient: gT(q) = coef_t scale_t [cA_T + 2 T(q) cB + R(q) cC](q), gR likewise, with c* the
// correlations of the maps (zero outside the interior) with the flipped window. Written, or
// added (accumulate), to the outputs that are not null. Tile 64 x 16 over the whole plane.
// ---------------------------------------------------------------------------------------------
// The four maps are correlated two at a time through one LDS tile of two maps (29 KB, five
// workgroups per CU; all four at once held 58 KB, two per CU): A_T and A_R, then B and Cm, whose
// loads are in flight while the first two are correlated.
__global__ __launch_bounds__(256) void k_ssimf_grad(SsimGArgs a) {
  __shared__ double s_m[2][kFillN];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int x0 = blockIdx.x * kGW;
  const int y0 = blockIdx.y * kGH;
  // Flipped window F'[a][b] = F[9 - a][9 - b], offsets (a - 4, b - 4) from the output pel.
  const int base = (kGK * wave) * kGLW + lane;
  auto load = [&](int R, int c, double (&v)[2]) {
    v[0] = s_m[0][base + R * kGLW + c];
    v[1] = s_m[1][base + R * kGLW + c];
  };
  double acc[2][kGK][2];  // [half][row][map of the half]: cA_T, cA_R; cB, cCm
  MapRegs v;
  load_maps(a, x0, y0, 0, v);
  store_maps(v, s_m);
  __syncthreads();
  load_maps(a, x0, y0, 2, v);
#pragma unroll
  for (int k = 0; k < kGK; k++) acc[0][k][0] = acc[0][k][1] = 0.0;
  win_rows<2, true>(load, a.f, acc[0], std::make_integer_sequence<int, kGK + 9>{});
  __syncthreads();  // every wave has finished with the first half's tile
  store_maps(v, s_m);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kGK; k++) acc[1][k][0] = acc[1][k][1] = 0.0;
  win_rows<2, true>(load, a.f, acc[1], std::make_integer_sequence<int, kGK + 9>{});

  const int x = x0 + lane;
#pragma unroll
  for (int k = 0; k < kGK; k++) {
    const int y = y0 + kGK * wave + k;
    if (x < a.w && y < a.h) {
      const int64_t o = (int64_t)y * a.w + x;
      const double tv = ldsample(a.tst, a.tst_f64, o) * a.scale_t;
      const double rv = ldsample(a.ref, a.ref_f64, o) * a.scale_r;
      const double cat = acc[0][k][0], car = acc[0][k][1], cb = acc[1][k][0], cc = acc[1][k][1];
      if (a.grad_t) {
        const double g = a.coef_t * a.scale_t * (cat + 2 * tv * cb + rv * cc);
        a.grad_t[o] = (float)(a.accumulate ? (double)a.grad_t[o] + g : g);
      }
      if (a.grad_r) {
        const double g = a.coef_r * a.scale_r * (car + 2 * rv * cb + tv * cc);
        a.grad_r[o] = (float)(a.accumulate ? (double)a.grad_r[o] + g : g);
      }
    }
  }
}

// dst (w/2 x h/2) = mean of the 2 x 2 blocks of src (w x h), exact in fp64.
__global__ __launch_bounds__(256) void k_down2(const void* __restrict__ src, int32_t src_f64, int32_t w, int32_t h,
                                               double* __restrict__ dst) {
  const int dw = w >> 1, dh = h >> 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)dw * dh) return;
  const int y = (int)(i / dw);
  const int x = (int)(i - (int64_t)y * dw);
  const int64_t o = (int64_t)(2 * y) * w + 2 * x;
  dst[i] = 0.25 * ((ldsample(src, src_f64, o) + ldsample(src, src_f64, o + 1)) +
                   (ldsample(src, src_f64, o + w) + ldsample(src, src_f64, o + w + 1)));
}

// Pyramid adjoint: fine = coef * fine + coarse[y >> 1][x >> 1] / 4 where a coarse parent exists.
__global__ __launch_bounds__(256) void k_up2_addf(const float* __restrict__ coarse, double coef, int32_t w, int32_t h,
                                                  float* __restrict__ fine) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)w * h) return;
  const int y = (int)(i / w);
  const int x = (int)(i - (int64_t)y * w);
  double v = coef * (double)fine[i];
  const int dw = w >> 1, dh = h >> 1;
  if (coarse && (y >> 1) < dh && (x >> 1) < dw) v += 0.25 * (double)coarse[(int64_t)(y >> 1) * dw + (x >> 1)];
  fine[i] = (float)v;
}

// dst = rint(clamp(src * scale, 0, maxv)), 8- or 16-bit samples for the shift search.
__global__ __launch_bounds__(256) void k_quantf(const float* __restrict__ src, int64_t n, double scale, int32_t maxv,
                                                void* __restrict__ dst, int32_t bytes) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = fmin(fmax((double)src[i] * scale, 0.0), (double)maxv);
  const uint32_t q = (uint32_t)__builtin_rint(v);
  if (bytes == 1) ((uint8_t*)dst)[i] = (uint8_t)q;
  else ((uint16_t*)dst)[i] = (uint16_t)q;
}

// k_quantf four samples per thread (n4 groups of four; 16-byte aligned source, destination aligned
// to four samples): the same conversion per sample.
template <int BYTES>
__global__ __launch_bounds__(256) void k_quantf4(const float4* __restrict__ src, int64_t n4, double scale, int32_t maxv,
                                                 void* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const float4 s = src[i];
  const float e[4] = {s.x, s.y, s.z, s.w};
  uint32_t q[4];
#pragma unroll
  for (int k = 0; k < 4; k++) q[k] = (uint32_t)__builtin_rint(fmin(fmax((double)e[k] * scale, 0.0), (double)maxv));
  if constexpr (BYTES == 1) ((uint32_t*)dst)[i] = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
  else ((uint2*)dst)[i] = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
}

// The GCD component c of a shift-compensated picture: from device memory when the search's summary
// holds it, else the value given.
__device__ __forceinline__ int32_t scp_gcd(const ScpFArgs& a, int c) { return a.gcd_dev ? a.gcd_dev[c] : a.gcd[c]; }

// Shift-compensated picture in sample units (xShftCompPic::xGenShftCompRow on real samples):
// out_c(p) = clamp(src_c(clamp(p + o(best(p)))) * scale - sgn gcd_c, 0, maxv), 3 planes.
__global__ __launch_bounds__(256) void k_scp_gatherf(ScpFArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t plane = (int64_t)a.w * a.h;
  if (i >= plane) return;
  const int y = (int)(i / a.w);
  const int x = (int)(i - (int64_t)y * a.w);
  const int n = 2 * a.B + 1;
  const int idx = a.best[i];
  const int sy = clampi(y + idx / n - a.B, 0, a.h - 1);
  const int sx = clampi(x + idx % n - a.B, 0, a.w - 1);
  const int64_t s = (int64_t)sy * a.w + sx;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double v = (double)a.src[c * plane + s] * a.scale - a.sgn * scp_gcd(a, c);
    a.out[c * plane + i] = (float)fmin(fmax(v, 0.0), (double)a.maxv);
  }
}

// Adjoint of k_scp_gatherf: grad_src_c(clamp(p + o(p))) += coef * scale * g_c(p) wherever the clamp
// to [0, maxv] is not active. Computed as a gather, one thread per source pel q: every p whose
// winner clamp(p + o(p)) is q lies within B of q (the window clamp only moves a target towards
// the border it crossed), so q visits its (2B + 1)^2 neighbours in raster order and adds the
// contributions of those that chose it. Each q is written once: no atomics, and the same sums in
// the same order on every run. A target clamped into the picture lies on its first or last row or
// column, so a q inside those is chosen by p = q + e exactly when p's window index is
// (B - ey) n + (B - ex): one compare, no division.
struct ScpAdj {
  int32_t y, x, B, n;
  bool border;
  __device__ __forceinline__ bool chose(const ScpFArgs& a, int ey, int ex, int idx) const {
    if (!border) return idx == (B - ey) * n + (B - ex);
    if (idx < 0) return false;
    const int sy = clampi(y + ey + idx / n - B, 0, a.h - 1);
    const int sx = clampi(x + ex + idx % n - B, 0, a.w - 1);
    return sy == y && sx == x;
  }
};

// The clamp test of q's three shifted values; false when none is open (nothing to add).
__device__ __forceinline__ bool scp_open(const ScpFArgs& a, int64_t q, int64_t plane, bool (&open)[3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double v = (double)a.src[c * plane + q] * a.scale - a.sgn * scp_gcd(a, c);
    open[c] = v >= 0.0 && v <= (double)a.maxv;
  }
  return open[0] || open[1] || open[2];
}

// Any search range: the neighbours straight from memory. Every pel is written (0 where nothing applies).
__global__ __launch_bounds__(256) void k_scp_scatterf(ScpFArgs a, const float* __restrict__ g, double coef,
                                                      float* __restrict__ grad_src) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t plane = (int64_t)a.w * a.h;
  if (i >= plane) return;
  ScpAdj q;
  q.y = (int)((uint32_t)i / (uint32_t)a.w);  // planes hold fewer than 2^31 pels
  q.x = (int)((uint32_t)i - (uint32_t)q.y * (uint32_t)a.w);
  q.B = a.B;
  q.n = 2 * a.B + 1;
  q.border = q.y == 0 || q.y == a.h - 1 || q.x == 0 || q.x == a.w - 1;
  bool open[3];
  float acc[3] = {0.f, 0.f, 0.f};
  if (scp_open(a, i, plane, open)) {
    const double k = coef * a.scale;
    for (int py = max(q.y - q.B, 0); py <= min(q.y + q.B, a.h - 1); py++)
      for (int px = max(q.x - q.B, 0); px <= min(q.x + q.B, a.w - 1); px++) {
        const int64_t p = (int64_t)py * a.w + px;
        if (!q.chose(a, py - q.y, px - q.x, a.best[p])) continue;
#pragma unroll
        for (int c = 0; c < 3; c++)
          if (open[c]) acc[c] += (float)(k * (double)g[c * plane + p]);
      }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) grad_src[c * plane + i] = acc[c];
}

// B = 2: a 64 x 16 tile of q per workgroup of 1024 threads (one pel each), its 2-pel neighbourhood of
// winners and incoming gradients staged in LDS (pels outside the picture hold the index -1, chosen by
// nobody). Every load of a thread (its share of the neighbourhood, then its pel's source values for
// the clamp test) is issued before the first LDS store. The 25 neighbours' winners are tested
// without branches into a bit mask, and only the neighbours that chose q are then added, in the
// window's raster order. Every pel of the tile is written, 0 where nothing chose it or its shifted
// values were clamped.
constexpr int kAdjTW = 64, kAdjTH = 16, kAdjLW = kAdjTW + 4, kAdjLH = kAdjTH + 4;
constexpr int kAdjThreads = kAdjTW * kAdjTH;
constexpr int kAdjN = kAdjLH * kAdjLW, kAdjIt = (kAdjN + kAdjThreads - 1) / kAdjThreads;

__global__ __launch_bounds__(kAdjThreads) void k_scp_scatterf_b2(ScpFArgs a, const float* __restrict__ g,
                                                                 double coef, float* __restrict__ grad_src) {
  __shared__ int32_t s_best[kAdjN];
  __shared__ float s_g[3][kAdjN];
  const int64_t plane = (int64_t)a.w * a.h;
  const int x0 = blockIdx.x * kAdjTW, y0 = blockIdx.y * kAdjTH;
  int16_t vb[kAdjIt];
  float vg[kAdjIt][3];
  uint32_t vin = 0;
#pragma unroll
  for (int j = 0; j < kAdjIt; j++) {
    const int k = min((int)threadIdx.x + kAdjThreads * j, kAdjN - 1);
    const int ly = k / kAdjLW, lx = k - ly * kAdjLW;
    const int py = y0 - 2 + ly, px = x0 - 2 + lx;
    const bool in = py >= 0 && py < a.h && px >= 0 && px < a.w;
    const int64_t p = in ? (int64_t)py * a.w + px : 0;
    vin |= (uint32_t)in << j;
    vb[j] = a.best[p];
#pragma unroll
    for (int c = 0; c < 3; c++) vg[j][c] = g[c * plane + p];
  }
  const int tx = threadIdx.x & (kAdjTW - 1), ty = threadIdx.x / kAdjTW;
  ScpAdj q;
  q.y = y0 + ty;
  q.x = x0 + tx;
  const int64_t i = (int64_t)min(q.y, a.h - 1) * a.w + min(q.x, a.w - 1);
  float src[3];
#pragma unroll
  for (int c = 0; c < 3; c++) src[c] = a.src[c * plane + i];
#pragma unroll
  for (int j = 0; j < kAdjIt; j++) {
    const int k = (int)threadIdx.x + kAdjThreads * j;
    const bool in = (vin >> j) & 1u;
    if (k < kAdjN) {
      s_best[k] = in ? (int32_t)vb[j] : -1;
#pragma unroll
      for (int c = 0; c < 3; c++) s_g[c][k] = in ? vg[j][c] : 0.f;
    }
  }
  __syncthreads();
  if (q.y >= a.h || q.x >= a.w) return;
  q.B = 2;
  q.n = 5;
  q.border = q.y == 0 || q.y == a.h - 1 || q.x == 0 || q.x == a.w - 1;
  bool open[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {  // the clamp test of scp_open on the values loaded above
    const double v = (double)src[c] * a.scale - a.sgn * scp_gcd(a, c);
    open[c] = v >= 0.0 && v <= (double)a.maxv;
  }
  float acc[3] = {0.f, 0.f, 0.f};
  if (open[0] || open[1] || open[2]) {
    uint32_t hits = 0;
#pragma unroll
    for (int e = 0; e < 25; e++) {
      const int ey = e / 5 - 2, ex = e % 5 - 2;
      hits |= (uint32_t)q.chose(a, ey, ex, s_best[(ty + 2 + ey) * kAdjLW + tx + 2 + ex]) << e;
    }
    const double kc = coef * a.scale;
    while (hits) {
      const int e = __builtin_ctz(hits);
      hits &= hits - 1;
      const int l = (ty + 2 + e / 5 - 2) * kAdjLW + tx + 2 + e % 5 - 2;
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (open[c]) acc[c] += (float)(kc * (double)s_g[c][l]);
    }
  }
  const int64_t o = (int64_t)q.y * a.w + q.x;
#pragma unroll
  for (int c = 0; c < 3; c++) grad_src[c * plane + o] = acc[c];
}

}  // namespace

int32_t ssimf_tiles_x(int32_t w) { return w > 10 ? (w - 10 + kGW - 1) / kGW : 0; }

hipError_t launch_ssimf(const SsimFArgs& a, int32_t calc_l, int32_t grad, hipStream_t stream) {
  const int tiles_x = ssimf_tiles_x(a.w);
  const int tiles_y = a.h > 10 ? (a.h - 10 + kGH - 1) / kGH : 0;
  if (tiles_x != a.tiles_x || !ssim_window_symmetric(a.f)) return hipErrorInvalidValue;
  if (tiles_x == 0 || tiles_y == 0) return hipSuccess;
  const dim3 grid(tiles_x, tiles_y);
  if (calc_l && grad) k_ssimf<1, 1><<<grid, dim3(256), 0, stream>>>(a);
  else if (calc_l) k_ssimf<1, 0><<<grid, dim3(256), 0, stream>>>(a);
  else if (grad) k_ssimf<0, 1><<<grid, dim3(256), 0, stream>>>(a);
  else k_ssimf<0, 0><<<grid, dim3(256), 0, stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_ssimf_grad(const SsimGArgs& a, hipStream_t stream) {
  if (!ssim_window_symmetric(a.f)) return hipErrorInvalidValue;
  if (a.w <= 0 || a.h <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ssimf_grad, dim3((a.w + kGW - 1) / kGW, (a.h + kGH - 1) / kGH), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_down2(const void* src, int32_t src_f64, int32_t w, int32_t h, double* dst, hipStream_t stream) {
  const int64_t n = (int64_t)(w >> 1) * (h >> 1);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_down2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, src_f64, w, h, dst);
  return hipGetLastError();
}

hipError_t launch_up2_addf(const float* coarse, double coef, int32_t w, int32_t h, float* fine, hipStream_t stream) {
  const int64_t n = (int64_t)w * h;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_up2_addf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, coarse, coef, w, h, fine);
  return hipGetLastError();
}

hipError_t launch_quantf(const float* src, int64_t n, double scale, int32_t maxv, void* dst, int32_t bytes,
                         hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (n % 4 == 0 && ((uintptr_t)src & 15) == 0 && ((uintptr_t)dst & (uintptr_t)(4 * bytes - 1)) == 0) {
    const int64_t n4 = n / 4;
    const dim3 grid((unsigned)((n4 + 255) / 256));
    if (bytes == 1)
      hipLaunchKernelGGL(k_quantf4<1>, grid, dim3(256), 0, stream, (const float4*)src, n4, scale, maxv, dst);
    else
      hipLaunchKernelGGL(k_quantf4<2>, grid, dim3(256), 0, stream, (const float4*)src, n4, scale, maxv, dst);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_quantf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, src, n, scale, maxv, dst, bytes);
  return hipGetLastError();
}

hipError_t launch_scp_gatherf(const ScpFArgs& a, hipStream_t stream) {
  const int64_t n = (int64_t)a.w * a.h;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scp_gatherf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_scp_scatterf(const ScpFArgs& a, const float* g, double coef, float* grad_src, hipStream_t stream) {
  const int64_t n = (int64_t)a.w * a.h;
  if (n <= 0) return hipSuccess;
  if (a.B == 2) {
    const dim3 grid((unsigned)((a.w + kAdjTW - 1) / kAdjTW), (unsigned)((a.h + kAdjTH - 1) / kAdjTH));
    hipLaunchKernelGGL(k_scp_scatterf_b2, grid, dim3(kAdjThreads), 0, stream, a, g, coef, grad_src);
  } else {
    hipLaunchKernelGGL(k_scp_scatterf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, g, coef, grad_src);
  }
  return hipGetLastError();
}

}  // namespace qmiv
