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
// gate (nullable): a first pass's summary; when it found every input in range this pass has nothing
// to redo (the concealment flow re-runs the statistics only after clipping).
template <typename T, typename TM>
__global__ __launch_bounds__(256) void k_mstats(const T* __restrict__ p0, const T* __restrict__ p1, Geom g,
                                                const TM* __restrict__ mask, Geom gm, MaskRows r,
                                                const MaskSummary* __restrict__ gate) {
  if (gate && gate->all_in_range) return;
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

// Buffer offset past every frame (frames are below 2 GiB): a raw buffer load there returns 0.
constexpr uint32_t kOob = 0x80000000u;

// Four adjacent samples from sample index i (i * sizeof(T) aligned to 4 * sizeof(T)); zeros when !ok.
template <typename T>
__device__ __forceinline__ void ld4(const Buf<T>& b, uint32_t i, bool ok, uint32_t (&v)[4]) {
  const uint32_t off = ok ? i * (uint32_t)sizeof(T) : kOob;
  if constexpr (sizeof(T) == 2) {
    const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(b.r, off, 0, 0));
    v[0] = w.x & 0xffffu; v[1] = w.x >> 16; v[2] = w.y & 0xffffu; v[3] = w.y >> 16;
  } else {
    const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(b.r, off, 0, 0);
    v[0] = w & 0xffu; v[1] = (w >> 8) & 0xffu; v[2] = (w >> 16) & 0xffu; v[3] = w >> 24;
  }
}

// Two adjacent samples from sample index i (aligned to 2 * sizeof(T)); zeros when !ok.
template <typename T>
__device__ __forceinline__ void ld2(const Buf<T>& b, uint32_t i, bool ok, uint32_t (&v)[4]) {
  const uint32_t off = ok ? i * (uint32_t)sizeof(T) : kOob;
  if constexpr (sizeof(T) == 2) {
    const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(b.r, off, 0, 0);
    v[0] = w & 0xffffu; v[1] = w >> 16;
  } else {
    const uint32_t w = __builtin_amdgcn_raw_buffer_load_b16(b.r, off, 0, 0);
    v[0] = w & 0xffu; v[1] = w >> 8;
  }
  v[2] = v[3] = 0u;
}

// Chroma samples of four adjacent pels from x (x divisible by 4): two for 4:2:0, four for 4:4:4.
template <typename T, int CS>
__device__ __forceinline__ void ldc(const Buf<T>& b, uint32_t i, bool ok, uint32_t (&v)[4]) {
  if constexpr (CS == 1) ld2<T>(b, i, ok, v);
  else ld4<T>(b, i, ok, v);
}

// Block totals of the eleven per-pel statistics over a 256-lane workgroup: q 0 counted pels,
// 1..3 SD, 4..6 SSD (64-bit sums; two's complement, so one adder serves the signed SD), 7..9 maxima
// of P0, P1 and the mask, 10 the OR of the difference flags. The lanes' values are transposed
// through LDS and each quantity is reduced by 16 lanes (16 values each, conflict-free reads), then
// across those 16. tot[q] is valid in every lane on return.
constexpr int kStatQ = 11;
__device__ __forceinline__ uint64_t stat_comb(int q, uint64_t a, uint64_t b) {
  return q < 7 ? a + b : (q < 10 ? max(a, b) : (a | b));
}
__device__ __forceinline__ void block_stats(const uint64_t (&v)[kStatQ], uint64_t (*s)[256], uint64_t* tot) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kStatQ; q++) s[q][t] = v[q];
  __syncthreads();
  if (t < kStatQ * 16) {
    const int q = t >> 4, p = t & 15;
    uint64_t acc = s[q][p];
#pragma unroll
    for (int i = 1; i < 16; i++) acc = stat_comb(q, acc, s[q][16 * i + p]);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) acc = stat_comb(q, acc, __shfl_xor(acc, o));
    if (p == 0) tot[q] = acc;
  }
  __syncthreads();
}

// k_mstats for widths divisible by 4, the same sums in the same integer types: one workgroup per
// row, each lane four adjacent pels per step (one load per plane and picture, 1024 pels per step),
// two steps in flight per pass. Lanes past the row read zeros (the buffer's bounds check), which
// add nothing: mask 0, equal pels, no larger maximum. grid = h.
template <typename T, typename TM, int CS, int CSM>
__global__ __launch_bounds__(256) void k_mstats_v(const T* __restrict__ p0, const T* __restrict__ p1, Geom g,
                                                  const TM* __restrict__ mask, Geom gm, MaskRows r,
                                                  const MaskSummary* __restrict__ gate) {
  __shared__ uint64_t s_red[kStatQ][256];
  __shared__ uint64_t s_tot[kStatQ];
  if (gate && gate->all_in_range) return;  // uniform over the workgroup
  const int y = blockIdx.x;
  const Buf<T> b0(p0, g.frame_bytes), b1(p1, g.frame_bytes);
  const Buf<TM> bm(mask, gm.frame_bytes);
  const uint32_t row = (uint32_t)y * (uint32_t)g.w;
  const uint32_t crow = (uint32_t)(y >> CS) * (uint32_t)g.cw;
  const uint32_t mrow = (uint32_t)(y >> CSM) * (uint32_t)gm.cw;
  const uint32_t cb = (uint32_t)g.cb_off, cr = (uint32_t)g.cr_off;
  const uint32_t mcb = (uint32_t)gm.cb_off, mcr = (uint32_t)gm.cr_off;
  int32_t nnz = 0;
  int64_t sd[3] = {0, 0, 0};
  uint64_t ssd[3] = {0, 0, 0};
  uint32_t mx0 = 0, mx1 = 0, mxm = 0, diff = 0;
  for (int base = 0; base < g.w; base += 2048) {
    uint32_t m[2][4], mc[2][2][4], a[2][3][4], b[2][3][4];
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const uint32_t x = (uint32_t)(base + 1024 * s + 4 * (int)threadIdx.x);
      const bool ok = x < (uint32_t)g.w;
      const uint32_t cx = x >> CS, mcx = x >> CSM;
      ld4<TM>(bm, row + x, ok, m[s]);
      ldc<TM, CSM>(bm, mcb + mrow + mcx, ok, mc[s][0]);
      ldc<TM, CSM>(bm, mcr + mrow + mcx, ok, mc[s][1]);
      ld4<T>(b0, row + x, ok, a[s][0]);
      ldc<T, CS>(b0, cb + crow + cx, ok, a[s][1]);
      ldc<T, CS>(b0, cr + crow + cx, ok, a[s][2]);
      ld4<T>(b1, row + x, ok, b[s][0]);
      ldc<T, CS>(b1, cb + crow + cx, ok, b[s][1]);
      ldc<T, CS>(b1, cr + crow + cx, ok, b[s][2]);
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t mk = m[s][k];
        mxm = max(mxm, max(mk, max(mc[s][0][k >> CSM], mc[s][1][k >> CSM])));
        nnz += mk != 0u;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const int kc = c == 0 ? k : (k >> CS);
          const uint32_t av = a[s][c][kc], bv = b[s][c][kc];
          mx0 = max(mx0, av);
          mx1 = max(mx1, bv);
          const uint32_t d = bv - av;
          diff |= (d != 0u) << c;
          sd[c] += (int64_t)(int32_t)(d * mk);
          ssd[c] += (uint64_t)(int64_t)(int32_t)(d * d) * (uint64_t)mk;
        }
      }
  }
  const uint64_t v[kStatQ] = {(uint64_t)nnz, (uint64_t)sd[0], (uint64_t)sd[1], (uint64_t)sd[2], ssd[0], ssd[1], ssd[2],
                               mx0, mx1, mxm, diff};
  block_stats(v, s_red, s_tot);
  if (threadIdx.x < kStatQ) {
    const int q = threadIdx.x;
    const uint64_t t = s_tot[q];
    if (q == 0) r.nnz[y] = (int32_t)t;
    else if (q < 4) r.sd[(q - 1) * g.h + y] = (int64_t)t;
    else if (q < 7) r.ssd[(q - 4) * g.h + y] = t;
    else if (q < 10) r.max[(q - 7) * g.h + y] = (uint32_t)t;
    else r.diff[y] = (uint32_t)t;
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
// the GCD's denominator (the processor form takes the count as an argument). One block of 256;
// each lane sums every 256th row, four rows' loads in flight.
__global__ __launch_bounds__(256) void k_msummary(Geom g, int32_t maxv, int32_t maxm, MaskRows r, int32_t md0,
                                                  int32_t md1, int32_t md2, int32_t md3, int32_t nnm_given,
                                                  int32_t gcd_given,
                                                  int32_t g0, int32_t g1, int32_t g2, int32_t g3,
                                                  MaskSummary* __restrict__ out, const MaskSummary* __restrict__ gate) {
  __shared__ uint64_t s_red[kStatQ][256];
  __shared__ uint64_t s_tot[kStatQ];
  // Concealment flow: a first pass that found every input in range is this pass's result.
  if (gate && gate->all_in_range) {
    if (threadIdx.x == 0) *out = *gate;
    return;
  }
  uint64_t v[kStatQ];
#pragma unroll
  for (int q = 0; q < kStatQ; q++) v[q] = 0;
  // Raw buffer loads, the offset of a row past h replaced by one past every array: such loads
  // return 0 and need no branch, so all four rows' loads issue together.
  const uint32_t h = (uint32_t)g.h;
  const __amdgpu_buffer_rsrc_t rn = __builtin_amdgcn_make_buffer_rsrc(r.nnz, 0, (int)(4 * h), kBufDword3);
  const __amdgpu_buffer_rsrc_t rsd = __builtin_amdgcn_make_buffer_rsrc(r.sd, 0, (int)(24 * h), kBufDword3);
  const __amdgpu_buffer_rsrc_t rss = __builtin_amdgcn_make_buffer_rsrc(r.ssd, 0, (int)(24 * h), kBufDword3);
  const __amdgpu_buffer_rsrc_t rmx = __builtin_amdgcn_make_buffer_rsrc(r.max, 0, (int)(12 * h), kBufDword3);
  const __amdgpu_buffer_rsrc_t rdf = __builtin_amdgcn_make_buffer_rsrc(r.diff, 0, (int)(4 * h), kBufDword3);
  for (uint32_t y0 = 0; y0 < h; y0 += 4 * 256) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t y = y0 + 256u * k + threadIdx.x;
      const bool ok = y < h;
      const uint32_t o4 = ok ? 4u * y : kOob, o8 = ok ? 8u * y : kOob;
      v[0] += (uint64_t)(int64_t)(int32_t)__builtin_amdgcn_raw_buffer_load_b32(rn, o4, 0, 0);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        v[1 + c] += __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rsd, o8 + (ok ? 8u * c * h : 0u), 0, 0));
        v[4 + c] += __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rss, o8 + (ok ? 8u * c * h : 0u), 0, 0));
        v[7 + c] = max(v[7 + c], (uint64_t)__builtin_amdgcn_raw_buffer_load_b32(rmx, o4 + (ok ? 4u * c * h : 0u), 0, 0));
      }
      v[10] |= __builtin_amdgcn_raw_buffer_load_b32(rdf, o4, 0, 0);
    }
  }
  block_stats(v, s_red, s_tot);
  const int64_t tn = (int64_t)s_tot[0];
  const int64_t tsd[3] = {(int64_t)s_tot[1], (int64_t)s_tot[2], (int64_t)s_tot[3]};
  const uint64_t tssd[3] = {s_tot[4], s_tot[5], s_tot[6]};
  const uint32_t tmx[3] = {(uint32_t)s_tot[7], (uint32_t)s_tot[8], (uint32_t)s_tot[9]};
  const uint32_t tdiff = (uint32_t)s_tot[10];
  if (threadIdx.x == 0) {
    out->nnm = (int32_t)tn;
    out->in_range[0] = tmx[0] <= (uint32_t)maxv;
    out->in_range[1] = tmx[1] <= (uint32_t)maxv;
    out->in_range[2] = tmx[2] <= (uint32_t)maxm;
    out->all_in_range = out->in_range[0] && out->in_range[1] && out->in_range[2];
    for (int i = 0; i < 3; i++) out->max_pel[i] = (int32_t)tmx[i];
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
// clamped coordinates. grid = (ceil(w / 64), ceil(h / kMTileH)). Each XCD takes whole tile rows
// (the halos of horizontally neighbouring tiles are read through one L2), the rows 8 apart, and
// the workgroups it receives one after the other go down a column of its rows: the work follows
// the mask, and consecutive workgroups of an XCD tend to share a CU, so a CU's tiles are spread
// over the picture's height rather than bunched along a row.
//
// Position t (contiguous per XCD, xcd_order) lists the tile rows by their index mod 8: class m holds
// the n_m rows m, m + 8, ..., its positions column by column, each column top to bottom.
__device__ __forceinline__ void masked_tile(int t, int gx, int gy, int& bx, int& by) {
  int m = 0, start = 0, n = 0;
  for (; m < 8; m++) {
    n = m < gy ? (gy - m + 7) >> 3 : 0;  // tile rows of class m
    if (t < start + n * gx) break;
    start += n * gx;
  }
  const int l = t - start;
  bx = l / n;
  by = m + 8 * (l - bx * n);
}
constexpr int kMTileW = 64;
constexpr int kMRows = 4;
constexpr int kMTileH = 4 * kMRows;

template <typename T, typename TM>
__global__ __launch_bounds__(256) void k_cps_masked(const T* __restrict__ p0, const T* __restrict__ p1, Geom g,
                                                    const TM* __restrict__ mask, const MaskSummary* __restrict__ sum,
                                                    int32_t B, int32_t w0, int32_t w1, int32_t w2, int32_t w3,
                                                    uint64_t* __restrict__ partials, int32_t gate) {
  extern __shared__ uint16_t s_lds[];
  if (!cps_gate_open(gate, sum->all_in_range)) return;
  const int lw = kMTileW + 2 * B, lh = kMTileH + 2 * B;
  const int lsz = lw * lh;
  uint16_t* s_pic[2][3] = {{s_lds, s_lds + lsz, s_lds + 2 * lsz}, {s_lds + 3 * lsz, s_lds + 4 * lsz, s_lds + 5 * lsz}};
  uint16_t* s_msk = s_lds + 6 * lsz;
  int bx, by;
  masked_tile(xcd_order(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y), gridDim.x, gridDim.y, bx, by);
  const int x0 = bx * kMTileW, y0 = by * kMTileH;
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
        if (lane == 0 && y < g.h) partials[(((int64_t)bx * 2 + dir) * 3 + c) * g.h + y] = v;
      }
  }
}

template <typename T, typename TM>
hipError_t mstats_t(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm, const MaskRows& r,
                    hipStream_t st, const MaskSummary* gate) {
#ifndef QMIV_MSTATS_VECTOR
#define QMIV_MSTATS_VECTOR 1
#endif
  if (QMIV_MSTATS_VECTOR && g.w % 4 == 0) {
    const dim3 grid((unsigned)g.h), block(256);
    const T* a = (const T*)p0;
    const T* b = (const T*)p1;
    const TM* m = (const TM*)mask;
    if (g.cs == 1)
      gm.cs == 1 ? k_mstats_v<T, TM, 1, 1><<<grid, block, 0, st>>>(a, b, g, m, gm, r, gate)
                 : k_mstats_v<T, TM, 1, 0><<<grid, block, 0, st>>>(a, b, g, m, gm, r, gate);
    else
      gm.cs == 1 ? k_mstats_v<T, TM, 0, 1><<<grid, block, 0, st>>>(a, b, g, m, gm, r, gate)
                 : k_mstats_v<T, TM, 0, 0><<<grid, block, 0, st>>>(a, b, g, m, gm, r, gate);
    return hipGetLastError();
  }
  k_mstats<T, TM><<<dim3((unsigned)((g.h + 3) / 4)), dim3(256), 0, st>>>((const T*)p0, (const T*)p1, g,
                                                                         (const TM*)mask, gm, r, gate);
  return hipGetLastError();
}

template <typename T, typename TM>
hipError_t cps_t(const void* p0, const void* p1, const Geom& g, const void* mask, const MaskSummary* sum, int32_t B,
                 const int32_t w[4], uint64_t* partials, hipStream_t st, int32_t gate) {
  const dim3 grid((unsigned)((g.w + kMTileW - 1) / kMTileW), (unsigned)((g.h + kMTileH - 1) / kMTileH));
  k_cps_masked<T, TM><<<grid, dim3(256), masked_cps_lds_bytes(B), st>>>(
      (const T*)p0, (const T*)p1, g, (const TM*)mask, sum, B, w[0], w[1], w[2], w[3], partials, gate);
  return hipGetLastError();
}

}  // namespace

size_t masked_cps_lds_bytes(int32_t B) {
  return (size_t)7 * (kMTileW + 2 * B) * (kMTileH + 2 * B) * sizeof(uint16_t);
}

int32_t masked_cps_segments(const Geom& g) { return (g.w + kMTileW - 1) / kMTileW; }

hipError_t launch_mstats(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm,
                         const MaskRows& r, hipStream_t st, const MaskSummary* gate) {
  if (g.bytes_per_sample == 1)
    return gm.bytes_per_sample == 1 ? mstats_t<uint8_t, uint8_t>(p0, p1, g, mask, gm, r, st, gate)
                                    : mstats_t<uint8_t, uint16_t>(p0, p1, g, mask, gm, r, st, gate);
  return gm.bytes_per_sample == 1 ? mstats_t<uint16_t, uint8_t>(p0, p1, g, mask, gm, r, st, gate)
                                  : mstats_t<uint16_t, uint16_t>(p0, p1, g, mask, gm, r, st, gate);
}

hipError_t launch_msummary(const Geom& g, int32_t maxm, const MaskRows& r, const int32_t maxdiff[4], int32_t nnm_given,
                           const int32_t* gcd_given, MaskSummary* out, hipStream_t st, const MaskSummary* gate) {
  if (gate == out) return hipErrorInvalidValue;
  const int32_t z[4] = {0, 0, 0, 0};
  const int32_t* gg = gcd_given ? gcd_given : z;
  k_msummary<<<dim3(1), dim3(256), 0, st>>>(g, (1 << g.bd) - 1, maxm, r, maxdiff[0], maxdiff[1], maxdiff[2],
                                             maxdiff[3], nnm_given, gcd_given ? 1 : 0, gg[0], gg[1], gg[2], gg[3],
                                             out, gate);
  return hipGetLastError();
}

hipError_t launch_cps_masked(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm,
                             const MaskSummary* sum, int32_t B, const int32_t w[4], uint64_t* partials,
                             hipStream_t st, int32_t gate) {
  if (g.bytes_per_sample == 1)
    return gm.bytes_per_sample == 1 ? cps_t<uint8_t, uint8_t>(p0, p1, g, mask, sum, B, w, partials, st, gate)
                                    : cps_t<uint8_t, uint16_t>(p0, p1, g, mask, sum, B, w, partials, st, gate);
  return gm.bytes_per_sample == 1 ? cps_t<uint16_t, uint8_t>(p0, p1, g, mask, sum, B, w, partials, st, gate)
                                  : cps_t<uint16_t, uint16_t>(p0, p1, g, mask, sum, B, w, partials, st, gate);
}

}  // namespace qmiv
