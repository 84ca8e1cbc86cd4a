// gfx950 kernels of masked mode: the masked picture statistics (pel counts, mask-weighted SD and
// SSD per row, ranges, per-component equality) with their plane totals and the masked GCD, and
// the masked corresponding-pixel-shift search of both directions.
#include "device_common.h"
#include "qmiv_internal.h"

#include <climits>

namespace qmiv {
namespace {

template <typename T>
__device__ __forceinline__ uint32_t ldu(const T* __restrict__ p, int64_t i) {
  return (uint32_t)p[i];
}

// Rows of the 4:4:4 domain (xPSNR::xCalcCmpSSDM, libs/pmbb-ivqm/src/xPSNR.cpp:113-136, with
// xDistortionSTD::CalcWeightedSSD / CalcWeightedSD, libs/pmbb-core/src/xDistortionSTD.cpp:63-98):
// per pel of row y with mask value m (the mask's luma plane) and d_c = P1_c - P0_c,
//   sd_c  += (int64)(int32)(d_c * m)                 int32 product, wrapping
//   ssd_c += (uint64)(int64)(int32)(d_c * d_c) * m    xPow2 in int32, sign-extending conversion
//   nnz   += m != 0                                   xPixelOpsSTD::CountNonZero (xPixelOpsSTD.cpp:273-284)
// plus the largest pel of P0, P1 and the mask (all planes; xPicP::check) and whether any d_c is
// non-zero (xPicP::equalCmp). A 4:2:0 chroma sample is read at (y >> 1, x >> 1).
// One wave per row, 4 rows per block; grid = ceil(h / 4).
template <typename T, typename TM>
__global__ __launch_bounds__(256) void k_mstats(const T* __restrict__ p0, const T* __restrict__ p1, Geom g,
                                                const TM* __restrict__ mask, Geom gm, MaskRows r) {
  const int lane = threadIdx.x & 63;
  const int y = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (y >= g.h) return;
  int32_t nnz = 0;
  int64_t sd[3] = {0, 0, 0};
  uint64_t ssd[3] = {0, 0, 0};
  uint32_t mx0 = 0, mx1 = 0, mxm = 0, diff = 0;
  const int64_t row = (int64_t)y * g.w;
  const int64_t crow = (int64_t)(y >> g.cs) * g.cw;
  const int64_t mrow = (int64_t)(y >> gm.cs) * gm.cw;
  for (int x = lane; x < g.w; x += 64) {
    const uint32_t m = ldu(mask, row + x);
    const int64_t co = crow + (x >> g.cs);
    const int64_t mco = mrow + (x >> gm.cs);
    mxm = max(mxm, max(m, max(ldu(mask, gm.cb_off + mco), ldu(mask, gm.cr_off + mco))));
    nnz += m != 0;
    const uint32_t a[3] = {ldu(p0, row + x), ldu(p0, g.cb_off + co), ldu(p0, g.cr_off + co)};
    const uint32_t b[3] = {ldu(p1, row + x), ldu(p1, g.cb_off + co), ldu(p1, g.cr_off + co)};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      mx0 = max(mx0, a[c]);
      mx1 = max(mx1, b[c]);
      const uint32_t d = b[c] - a[c];
      diff |= (d != 0u) << c;
      sd[c] += (int64_t)(int32_t)(d * m);
      ssd[c] += (uint64_t)(int64_t)(int32_t)(d * d) * (uint64_t)m;
    }
  }
  const int64_t n = wave_sum((int64_t)nnz);
  int64_t s[3];
  uint64_t q[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    s[c] = wave_sum(sd[c]);
    q[c] = wave_sum(ssd[c]);
  }
  mx0 = wave_max(mx0);
  mx1 = wave_max(mx1);
  mxm = wave_max(mxm);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) diff |= (uint32_t)__shfl_xor((int)diff, o);
  if (lane == 0) {
    r.nnz[y] = (int32_t)n;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      r.sd[c * g.h + y] = s[c];
      r.ssd[c * g.h + y] = q[c];
    }
    r.max[y] = mx0;
    r.max[g.h + y] = mx1;
    r.max[2 * g.h + y] = mxm;
    r.diff[y] = diff;
  }
}

// _mm_cvtsd_si32 under round-to-nearest-even (libs/pmbb-core/src/xCommonDefCORE.h:130): NaN and
// results outside int32 give INT32_MIN.
__device__ __forceinline__ int32_t cvt_rne_x86(double v) {
  const double rv = __builtin_rint(v);
  if (!(rv >= -2147483648.0 && rv <= 2147483647.0)) return INT_MIN;
  return (int32_t)rv;
}

// Plane totals and the masked GCD (xGlobClrDiff::CalcGlobalColorDiffM,
// libs/pmbb-ivqm/src/xGlobClrDiff.cpp:52-83, as the processor calls it, xGlobClrDiff.h:42):
// g_c = clip(cvt(SD_c / (double)(nnm * maxm)), +-maxdiff_c); the fourth entry has SD 0. With
// gcd_given the caller's values are stored instead; nnm_given >= 0 replaces the counted pels in
// the GCD's denominator (the processor form takes the count as an argument). One block of 1024.
__global__ __launch_bounds__(1024) void k_msummary(Geom g, int32_t maxv, int32_t maxm, MaskRows r, int32_t md0,
                                                   int32_t md1, int32_t md2, int32_t md3, int32_t nnm_given,
                                                   int32_t gcd_given,
                                                   int32_t g0, int32_t g1, int32_t g2, int32_t g3,
                                                   MaskSummary* __restrict__ out) {
  __shared__ int64_t s_sd[3][16];
  __shared__ uint64_t s_ssd[3][16];
  __shared__ int64_t s_n[16];
  __shared__ uint32_t s_max[3][16], s_diff[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t n = 0, sd[3] = {0, 0, 0};
  uint64_t ssd[3] = {0, 0, 0};
  uint32_t mx[3] = {0, 0, 0}, diff = 0;
  for (int y = threadIdx.x; y < g.h; y += blockDim.x) {
    n += r.nnz[y];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      sd[c] += r.sd[c * g.h + y];
      ssd[c] += r.ssd[c * g.h + y];
      mx[c] = max(mx[c], r.max[c * g.h + y]);
    }
    diff |= r.diff[y];
  }
  n = wave_sum(n);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    sd[c] = wave_sum(sd[c]);
    ssd[c] = wave_sum(ssd[c]);
    mx[c] = wave_max(mx[c]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) diff |= (uint32_t)__shfl_xor((int)diff, o);
  if (lane == 0) {
    s_n[wave] = n;
    s_diff[wave] = diff;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      s_sd[c][wave] = sd[c];
      s_ssd[c][wave] = ssd[c];
      s_max[c][wave] = mx[c];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    int64_t tn = 0, tsd[3] = {0, 0, 0};
    uint64_t tssd[3] = {0, 0, 0};
    uint32_t tmx[3] = {0, 0, 0}, tdiff = 0;
    for (int w = 0; w < nw; w++) {
      tn += s_n[w];
      tdiff |= s_diff[w];
      for (int c = 0; c < 3; c++) {
        tsd[c] += s_sd[c][w];
        tssd[c] += s_ssd[c][w];
        tmx[c] = max(tmx[c], s_max[c][w]);
      }
    }
    out->nnm = (int32_t)tn;
    out->in_range[0] = tmx[0] <= (uint32_t)maxv;
    out->in_range[1] = tmx[1] <= (uint32_t)maxv;
    out->in_range[2] = tmx[2] <= (uint32_t)maxm;
    out->all_in_range = out->in_range[0] && out->in_range[1] && out->in_range[2];
    for (int c = 0; c < 3; c++) {
      out->sd[c] = tsd[c];
      out->ssd[c] = tssd[c];
      out->exact[c] = ((tdiff >> c) & 1u) == 0u;
    }
    if (gcd_given) {
      out->gcd[0] = g0;
      out->gcd[1] = g1;
      out->gcd[2] = g2;
      out->gcd[3] = g3;
    } else {
      const int64_t nnm = nnm_given >= 0 ? (int64_t)nnm_given : tn;
      const double denom = (double)(nnm * (int64_t)maxm);
      const int32_t md[4] = {md0, md1, md2, md3};
      for (int c = 0; c < 4; c++) {
        const int32_t v = cvt_rne_x86((double)(c < 3 ? tsd[c] : 0) / denom);
        out->gcd[c] = max(-md[c], min(v, md[c]));
      }
    }
  }
}

// Masked CPS search, both directions (xCorrespPixelShiftSTD::CalcDistAsymmetricRowM and
// FindBestPixelWithinBlockM, libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp:80-144): a pel with
// mask 0 contributes nothing; a candidate whose mask (read at the clamped position, as the
// replicated margin holds it) is 0 is skipped; the error is the int32 weighted sum of the four
// lanes' squared differences (the fourth lane: the GCD's fourth entry against 0), the first
// strict minimum in raster order wins, and the winner's squared differences, sign-extended to
// 64 bits, are multiplied by the pel's mask value and summed per row in uint64.
// Direction 0: X = P1 - g searched in Y = P0 (the masked processor's R2T); direction 1: X = P0 + g
// searched in P1 (its T2R).
// Tile: 64 columns x kMTileH rows; lane = column, wave w takes rows [w * kMRows, (w + 1) * kMRows).
// LDS holds both pictures' three planes and the mask over the tile plus a B-pel halo, read at
// clamped coordinates. grid = (ceil(w / 64), ceil(h / kMTileH)).
constexpr int kMTileW = 64;
constexpr int kMRows = 4;
constexpr int kMTileH = 4 * kMRows;

template <typename T, typename TM>
__global__ __launch_bounds__(256) void k_cps_masked(const T* __restrict__ p0, const T* __restrict__ p1, Geom g,
                                                    const TM* __restrict__ mask, const MaskSummary* __restrict__ sum,
                                                    int32_t B, int32_t w0, int32_t w1, int32_t w2, int32_t w3,
                                                    uint64_t* __restrict__ partials) {
  extern __shared__ uint16_t s_lds[];
  if (!sum->all_in_range) return;
  const int lw = kMTileW + 2 * B, lh = kMTileH + 2 * B;
  const int lsz = lw * lh;
  uint16_t* s_pic[2][3] = {{s_lds, s_lds + lsz, s_lds + 2 * lsz}, {s_lds + 3 * lsz, s_lds + 4 * lsz, s_lds + 5 * lsz}};
  uint16_t* s_msk = s_lds + 6 * lsz;
  const int x0 = blockIdx.x * kMTileW, y0 = blockIdx.y * kMTileH;
  for (int i = threadIdx.x; i < lsz; i += blockDim.x) {
    const int ly = i / lw, lx = i - ly * lw;
    const int gy = clampi(y0 - B + ly, 0, g.h - 1), gx = clampi(x0 - B + lx, 0, g.w - 1);
    const int64_t o = (int64_t)gy * g.w + gx;
    const int64_t co = (int64_t)(gy >> g.cs) * g.cw + (gx >> g.cs);
    s_pic[0][0][i] = (uint16_t)ldu(p0, o);
    s_pic[0][1][i] = (uint16_t)ldu(p0, g.cb_off + co);
    s_pic[0][2][i] = (uint16_t)ldu(p0, g.cr_off + co);
    s_pic[1][0][i] = (uint16_t)ldu(p1, o);
    s_pic[1][1][i] = (uint16_t)ldu(p1, g.cb_off + co);
    s_pic[1][2][i] = (uint16_t)ldu(p1, g.cr_off + co);
    s_msk[i] = (uint16_t)ldu(mask, o);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = x0 + lane;
  const int32_t wts[4] = {w0, w1, w2, w3};
  for (int k = 0; k < kMRows; k++) {
    const int ty = wave * kMRows + k;
    const int y = y0 + ty;
    uint64_t acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
    const int ci = (ty + B) * lw + lane + B;
    const uint32_t m = (x < g.w && y < g.h) ? s_msk[ci] : 0u;
    if (m != 0u) {
#pragma unroll
      for (int dir = 0; dir < 2; dir++) {
        // dir 0: X = P1, Y = P0, -g; dir 1: X = P0, Y = P1, +g
        const int xs = dir == 0 ? 1 : 0;
        const int32_t sg = dir == 0 ? -1 : 1;
        const int32_t t[4] = {(int32_t)s_pic[xs][0][ci] + sg * sum->gcd[0], (int32_t)s_pic[xs][1][ci] + sg * sum->gcd[1],
                              (int32_t)s_pic[xs][2][ci] + sg * sum->gcd[2], sg * sum->gcd[3]};
        const uint32_t d3 = (uint32_t)t[3];
        const uint32_t dist3 = d3 * d3;
        const uint16_t* Y0 = s_pic[1 - xs][0];
        const uint16_t* Y1 = s_pic[1 - xs][1];
        const uint16_t* Y2 = s_pic[1 - xs][2];
        int32_t best = INT_MAX;
        uint32_t bd0 = 0, bd1 = 0, bd2 = 0;
        for (int dy = -B; dy <= B; dy++) {
          const int rb = (ty + B + dy) * lw + lane + B;
          for (int dx = -B; dx <= B; dx++) {
            const int q = rb + dx;
            if (s_msk[q] == 0) continue;
            const uint32_t e0 = (uint32_t)(t[0] - (int32_t)Y0[q]);
            const uint32_t e1 = (uint32_t)(t[1] - (int32_t)Y1[q]);
            const uint32_t e2 = (uint32_t)(t[2] - (int32_t)Y2[q]);
            const uint32_t s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
            const int32_t err = (int32_t)(s0 * (uint32_t)wts[0] + s1 * (uint32_t)wts[1] + s2 * (uint32_t)wts[2] +
                                          dist3 * (uint32_t)wts[3]);
            if (err < best) {
              best = err;
              bd0 = s0;
              bd1 = s1;
              bd2 = s2;
            }
          }
        }
        acc[dir][0] = (uint64_t)(int64_t)(int32_t)bd0 * (uint64_t)m;
        acc[dir][1] = (uint64_t)(int64_t)(int32_t)bd1 * (uint64_t)m;
        acc[dir][2] = (uint64_t)(int64_t)(int32_t)bd2 * (uint64_t)m;
      }
    }
#pragma unroll
    for (int dir = 0; dir < 2; dir++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const uint64_t v = wave_sum(acc[dir][c]);
        if (lane == 0 && y < g.h) partials[(((int64_t)blockIdx.x * 2 + dir) * 3 + c) * g.h + y] = v;
      }
  }
}

template <typename T, typename TM>
hipError_t mstats_t(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm, const MaskRows& r,
                    hipStream_t st) {
  k_mstats<T, TM><<<dim3((unsigned)((g.h + 3) / 4)), dim3(256), 0, st>>>((const T*)p0, (const T*)p1, g,
                                                                         (const TM*)mask, gm, r);
  return hipGetLastError();
}

template <typename T, typename TM>
hipError_t cps_t(const void* p0, const void* p1, const Geom& g, const void* mask, const MaskSummary* sum, int32_t B,
                 const int32_t w[4], uint64_t* partials, hipStream_t st) {
  const dim3 grid((unsigned)((g.w + kMTileW - 1) / kMTileW), (unsigned)((g.h + kMTileH - 1) / kMTileH));
  k_cps_masked<T, TM><<<grid, dim3(256), masked_cps_lds_bytes(B), st>>>(
      (const T*)p0, (const T*)p1, g, (const TM*)mask, sum, B, w[0], w[1], w[2], w[3], partials);
  return hipGetLastError();
}

}  // namespace

size_t masked_cps_lds_bytes(int32_t B) {
  return (size_t)7 * (kMTileW + 2 * B) * (kMTileH + 2 * B) * sizeof(uint16_t);
}

int32_t masked_cps_segments(const Geom& g) { return (g.w + kMTileW - 1) / kMTileW; }

hipError_t launch_mstats(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm,
                         const MaskRows& r, hipStream_t st) {
  if (g.bytes_per_sample == 1)
    return gm.bytes_per_sample == 1 ? mstats_t<uint8_t, uint8_t>(p0, p1, g, mask, gm, r, st)
                                    : mstats_t<uint8_t, uint16_t>(p0, p1, g, mask, gm, r, st);
  return gm.bytes_per_sample == 1 ? mstats_t<uint16_t, uint8_t>(p0, p1, g, mask, gm, r, st)
                                  : mstats_t<uint16_t, uint16_t>(p0, p1, g, mask, gm, r, st);
}

hipError_t launch_msummary(const Geom& g, int32_t maxm, const MaskRows& r, const int32_t maxdiff[4], int32_t nnm_given,
                           const int32_t* gcd_given, MaskSummary* out, hipStream_t st) {
  const int32_t z[4] = {0, 0, 0, 0};
  const int32_t* gg = gcd_given ? gcd_given : z;
  k_msummary<<<dim3(1), dim3(1024), 0, st>>>(g, (1 << g.bd) - 1, maxm, r, maxdiff[0], maxdiff[1], maxdiff[2],
                                             maxdiff[3], nnm_given, gcd_given ? 1 : 0, gg[0], gg[1], gg[2], gg[3],
                                             out);
  return hipGetLastError();
}

hipError_t launch_cps_masked(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm,
                             const MaskSummary* sum, int32_t B, const int32_t w[4], uint64_t* partials,
                             hipStream_t st) {
  if (g.bytes_per_sample == 1)
    return gm.bytes_per_sample == 1 ? cps_t<uint8_t, uint8_t>(p0, p1, g, mask, sum, B, w, partials, st)
                                    : cps_t<uint8_t, uint16_t>(p0, p1, g, mask, sum, B, w, partials, st);
  return gm.bytes_per_sample == 1 ? cps_t<uint16_t, uint8_t>(p0, p1, g, mask, sum, B, w, partials, st)
                                  : cps_t<uint16_t, uint16_t>(p0, p1, g, mask, sum, B, w, partials, st);
}

}  // namespace qmiv
