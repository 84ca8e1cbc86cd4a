// gfx950 corresponding-pixel-shift (CPS) search for 4:2:0 pictures, B = 2 (the fast path's
// parameter range, checked on the host).
//
// The key of a tested pel p against a candidate q of the searched picture Y is the one of the
// generic fast kernel (kernels_psnr.hip):
//   key(p, q) = C(q) + a0(p) Y0(q) + a1(p) U(q) + a2(p) V(q),
//   C(q) = 32 (w0 Y0^2 + w1 U^2 + w2 V^2)(q) + h(q),  a_c = -64 w_c X_c,  h = 5 ly + lx,
// whose signed minimum over the 5 x 5 window is the reference's first strict minimum of the
// weighted squared error (libs/pmbb-ivqm/src/xCorrespPixelShiftSSE.cpp:40-71).
//
// With 4:2:0 pictures the chroma of both pictures is constant over 2 x 2 cells (the reference
// upsamples by nearest neighbour, libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49): a1, a2 are shared
// by the four pels of a tested cell and U, V by the four candidates of a searched cell. So
//   key(p, q) = [C(q) + a0(p) Y0(q)] + T(P, Q),   T(P, Q) = a1(P) U(Q) + a2(P) V(Q),
// and min over the window = min over the 3 x 3 cells Q it touches of
//   T(P, Q) + min over its candidates q in Q of [C(q) + a0(p) Y0(q)].
// Every key keeps its exact value (only the grouping of the minimum changes), so the winner and
// its tie-break are identical. Per candidate: one 24-bit multiply-add and a minimum.
//
// Workgroup: 4 waves, tile 128 x 8 pels; each lane owns one 2 x 2 block (one tested chroma cell;
// QMIV_CPS420_KB stacks more blocks per lane, which measured slower: fewer waves fit a SIMD). LDS
// holds the searched picture's luma records {Y0, C} for rows [y0 - 2, y0 + 10) x columns
// [x0 - 4, x0 + 132) and its chroma cells {U | V << 16}.
#include <climits>
#include <type_traits>

#include "device_common.h"
#include "qmiv_internal.h"

namespace qmiv {
namespace {

constexpr int kTW = 128;          // tile columns: 64 lanes x 2
#ifndef QMIV_CPS420_KB
#define QMIV_CPS420_KB 1
#endif
constexpr int kKB = QMIV_CPS420_KB;  // 2 x 2 blocks per lane, stacked vertically (1: 104 VGPRs, 4 waves per SIMD)
constexpr int kTH = 4 * 2 * kKB;  // tile rows: 4 waves x 2 kKB
constexpr int kLW = kTW + 8;      // luma record columns [x0 - 4, x0 + kTW + 4)
constexpr int kLH = kTH + 4;      // luma record rows [y0 - 2, y0 + kTH + 2)
constexpr int kCW = kLW / 2;      // chroma cells: luma record (ly, lx) lies in cell (ly >> 1, lx >> 1)
constexpr int kCH = kLH / 2;

// Records of the searched picture Yp for the tile at (x0, y0). Coordinates are clamped into the
// picture; with even picture sizes and even tile origins, the two luma columns (rows) of an LDS
// cell always clamp into one chroma sample, so cells stay uniform at the edges too.
// Each thread and step builds four luma records and two chroma cells of one LDS row: inside the
// picture from one aligned 8-byte (u16) or 4-byte (u8) luma load and two chroma samples per
// plane, at the edges from clamped single-sample loads; all steps' loads are issued together.
constexpr int kQW = kLW / 4;
constexpr int kFillItems = kLH * kQW;
constexpr int kFillSteps = (kFillItems + 255) / 256;

// The pels one thread loads for one fill, kept as the loaded words (a 4:2:0 row quad of luma,
// two samples of each chroma plane) from the load, issued a whole search phase ahead, to the store
// into LDS; nothing reads them in between, so the loads stay in flight across the search.
//   16-bit samples: l0 = Y0 | Y1 << 16, l1 = Y2 | Y3 << 16, u = U0 | U1 << 16, v = V0 | V1 << 16
//   8-bit samples:  l0 = Y0 | Y1 << 8 | Y2 << 16 | Y3 << 24, u = U0 | U1 << 8, v = V0 | V1 << 8
struct FillRegs {
  uint32_t l0[kFillSteps], l1[kFillSteps], u[kFillSteps], v[kFillSteps];
  uint32_t m0[kFillSteps], m1[kFillSteps];  // masked search: the quad's mask samples, as loaded
};

// Masked search: a searched pel whose mask is 0 gets the record {Y0 = 0, C = kMaskedC + (h mod 32)}
// (h: its candidate index, so a winner still decodes to a window position). Its keys are then
// kMaskedC + (h mod 32) + T(P, Q) > 2^29 while every real key lies in (-2^29, 2^29) and
// |T| < 2^29 (the host checks both bounds), so it never beats an unmasked candidate. A tested pel
// whose mask is 0 contributes nothing; any other has itself as an unmasked candidate, so its
// winner is the first strict minimum over the unmasked ones, as in
// xCorrespPixelShiftSTD::FindBestPixelWithinBlockM (libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp).
constexpr int32_t kMaskedC = 1 << 30;

template <typename T, typename TM, bool MASKED>
__device__ __forceinline__ void fill420_load(const Buf<T>& Yp, const Buf<TM>& Mp, const Geom& g, int x0, int y0,
                                             FillRegs& fr) {
  const uint32_t cb_off = (uint32_t)g.cb_off, cr_off = (uint32_t)g.cr_off;
  const bool interior = x0 >= 4 && x0 + kTW + 4 <= g.w && y0 >= 2 && y0 + kTH + 2 <= g.h && (g.w & 3) == 0;
  constexpr uint32_t sh = 8 * sizeof(T);
#pragma unroll
  for (int k = 0; k < kFillSteps; k++) {
    const int i = threadIdx.x + 256 * k;
    if (kFillItems % 256 != 0 && k == kFillSteps - 1 && i >= kFillItems) break;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    if (interior) {
      const int gy = y0 - 2 + ly;
      const int gx = x0 - 4 + lx;
      const uint32_t lo = __umul24((uint32_t)gy, (uint32_t)g.w) + (uint32_t)gx;
      const uint32_t co = __umul24((uint32_t)(gy >> 1), (uint32_t)g.cw) + (uint32_t)(gx >> 1);
      if constexpr (sizeof(T) == 2) {
        const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(Yp.r, lo * 2u, 0, 0));
        fr.l0[k] = w.x;
        fr.l1[k] = w.y;
        fr.u[k] = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, (cb_off + co) * 2u, 0, 0);
        fr.v[k] = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, (cr_off + co) * 2u, 0, 0);
      } else {
        fr.l0[k] = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, lo, 0, 0);
        fr.l1[k] = 0u;
        fr.u[k] = __builtin_amdgcn_raw_buffer_load_b16(Yp.r, cb_off + co, 0, 0);
        fr.v[k] = __builtin_amdgcn_raw_buffer_load_b16(Yp.r, cr_off + co, 0, 0);
      }
      if constexpr (MASKED) {
        if constexpr (sizeof(TM) == 2) {
          const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(Mp.r, lo * 2u, 0, 0));
          fr.m0[k] = w.x;
          fr.m1[k] = w.y;
        } else {
          fr.m0[k] = __builtin_amdgcn_raw_buffer_load_b32(Mp.r, lo, 0, 0);
        }
      }
    } else {
      const int gy = clampi(y0 - 2 + ly, 0, g.h - 1);
      const uint32_t rowo = __umul24((uint32_t)gy, (uint32_t)g.w);
      const uint32_t crow = __umul24((uint32_t)(gy >> 1), (uint32_t)g.cw);
      uint32_t yq[4], uq[2], vq[2];
#pragma unroll
      for (int m = 0; m < 4; m++) yq[m] = Yp.ld(rowo + (uint32_t)clampi(x0 - 4 + lx + m, 0, g.w - 1));
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const uint32_t co = crow + (uint32_t)(clampi(x0 - 4 + lx + 2 * c, 0, g.w - 1) >> 1);
        uq[c] = Yp.ld(cb_off + co);
        vq[c] = Yp.ld(cr_off + co);
      }
      if constexpr (sizeof(T) == 2) {
        fr.l0[k] = yq[0] | (yq[1] << 16);
        fr.l1[k] = yq[2] | (yq[3] << 16);
      } else {
        fr.l0[k] = yq[0] | (yq[1] << 8) | (yq[2] << 16) | (yq[3] << 24);
        fr.l1[k] = 0u;
      }
      fr.u[k] = uq[0] | (uq[1] << sh);
      fr.v[k] = vq[0] | (vq[1] << sh);
      if constexpr (MASKED) {
        uint32_t mq[4];
#pragma unroll
        for (int m = 0; m < 4; m++) mq[m] = Mp.ld(rowo + (uint32_t)clampi(x0 - 4 + lx + m, 0, g.w - 1));
        if constexpr (sizeof(TM) == 2) {
          fr.m0[k] = mq[0] | (mq[1] << 16);
          fr.m1[k] = mq[2] | (mq[3] << 16);
        } else {
          fr.m0[k] = mq[0] | (mq[1] << 8) | (mq[2] << 16) | (mq[3] << 24);
        }
      }
    }
  }
}

// Builds the LDS records {Y0, C} and chroma cells of one fill from the loaded words.
template <typename T, typename TM, bool MASKED>
__device__ __forceinline__ void fill420_store(const FillRegs& fr, int32_t w0, int32_t w1, int32_t w2, uint2* s_luma,
                                              uint32_t* s_cell) {
  constexpr uint32_t sh = 8 * sizeof(T);
  constexpr uint32_t mask = sizeof(T) == 2 ? 0xffffu : 0xffu;
#pragma unroll
  for (int k = 0; k < kFillSteps; k++) {
    const int i = threadIdx.x + 256 * k;
    if (kFillItems % 256 != 0 && k == kFillSteps - 1 && i >= kFillItems) break;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    uint32_t yv[4];
    if constexpr (sizeof(T) == 2) {
      yv[0] = fr.l0[k] & 0xffffu; yv[1] = fr.l0[k] >> 16; yv[2] = fr.l1[k] & 0xffffu; yv[3] = fr.l1[k] >> 16;
    } else {
      yv[0] = fr.l0[k] & 0xffu; yv[1] = (fr.l0[k] >> 8) & 0xffu; yv[2] = (fr.l0[k] >> 16) & 0xffu; yv[3] = fr.l0[k] >> 24;
    }
    const uint32_t cu[2] = {fr.u[k] & mask, (fr.u[k] >> sh) & mask};
    const uint32_t cv[2] = {fr.v[k] & mask, (fr.v[k] >> sh) & mask};
    // Squares and weights in unsigned 24-bit multiplies (full rate): the host's eligibility check
    // (non-negative weights, keys below 2^31) keeps samples <= 4095, so every operand, a square
    // included, is below 2^24 and every product below 2^31.
    const int32_t sc0 = (int32_t)__umul24((uint32_t)w1, __umul24(cu[0], cu[0])) + (int32_t)__umul24((uint32_t)w2, __umul24(cv[0], cv[0]));
    const int32_t sc1 = (int32_t)__umul24((uint32_t)w1, __umul24(cu[1], cu[1])) + (int32_t)__umul24((uint32_t)w2, __umul24(cv[1], cv[1]));
    const int32_t h = 5 * ly + lx;
    uint4 r01, r23;
    r01.x = yv[0]; r01.y = (uint32_t)(32 * ((int32_t)__umul24((uint32_t)w0, __umul24(yv[0], yv[0])) + sc0) + h);
    r01.z = yv[1]; r01.w = (uint32_t)(32 * ((int32_t)__umul24((uint32_t)w0, __umul24(yv[1], yv[1])) + sc0) + h + 1);
    r23.x = yv[2]; r23.y = (uint32_t)(32 * ((int32_t)__umul24((uint32_t)w0, __umul24(yv[2], yv[2])) + sc1) + h + 2);
    r23.z = yv[3]; r23.w = (uint32_t)(32 * ((int32_t)__umul24((uint32_t)w0, __umul24(yv[3], yv[3])) + sc1) + h + 3);
    if constexpr (MASKED) {
      bool on[4];
      if constexpr (sizeof(TM) == 2) {
        on[0] = (fr.m0[k] & 0xffffu) != 0; on[1] = (fr.m0[k] >> 16) != 0;
        on[2] = (fr.m1[k] & 0xffffu) != 0; on[3] = (fr.m1[k] >> 16) != 0;
      } else {
        on[0] = (fr.m0[k] & 0xffu) != 0; on[1] = (fr.m0[k] & 0xff00u) != 0;
        on[2] = (fr.m0[k] & 0xff0000u) != 0; on[3] = (fr.m0[k] >> 24) != 0;
      }
      if (!on[0]) { r01.x = 0u; r01.y = (uint32_t)kMaskedC | (r01.y & 31u); }
      if (!on[1]) { r01.z = 0u; r01.w = (uint32_t)kMaskedC | (r01.w & 31u); }
      if (!on[2]) { r23.x = 0u; r23.y = (uint32_t)kMaskedC | (r23.y & 31u); }
      if (!on[3]) { r23.z = 0u; r23.w = (uint32_t)kMaskedC | (r23.w & 31u); }
    }
    uint4* dst = reinterpret_cast<uint4*>(s_luma + ly * kLW + lx);
    dst[0] = r01;
    dst[1] = r23;
    if ((ly & 1) == 0)
      *reinterpret_cast<uint2*>(s_cell + (ly >> 1) * kCW + (lx >> 1)) =
          make_uint2(cu[0] | (cv[0] << 16), cu[1] | (cv[1] << 16));
  }
}

// One direction over one tile. acc[r][c]: the lane's per-row, per-component sums (both of its
// columns) kept across the tile sweep.
// The tested pels of one lane for one phase, as loaded words: per block b, the two luma rows of
// its 2 x 2 pels (one 32-bit load of two 16-bit samples or one 16-bit load of two 8-bit samples)
// and its chroma cell's Cb and Cr samples. Loaded a phase ahead, like FillRegs.
struct XRegs {
  uint32_t row[kKB][2], cb[kKB], cr[kKB];
  uint32_t mrow[kKB][2];  // masked search: the block's mask samples, two per row
};

template <typename T, typename TM, bool MASKED>
__device__ __forceinline__ void x420_load(const Buf<T>& Xp, const Buf<TM>& Mp, const Geom& g, int x0, int y0, int lane,
                                          int wave, XRegs& xr) {
  const uint32_t cb_off = (uint32_t)g.cb_off, cr_off = (uint32_t)g.cr_off;
  const int bx = min(x0 + 2 * lane, g.w - 2);  // block origin (even), clamped for the tile's tail
#pragma unroll
  for (int b = 0; b < kKB; b++) {
    const int by = min(y0 + 2 * (kKB * wave + b), g.h - 2);
    const uint32_t co = __umul24((uint32_t)(by >> 1), (uint32_t)g.cw) + (uint32_t)(bx >> 1);
    xr.cb[b] = Xp.ld(cb_off + co);
    xr.cr[b] = Xp.ld(cr_off + co);
#pragma unroll
    for (int py = 0; py < 2; py++) {
      // Even width and even bx: the pair is aligned to its own size.
      const uint32_t o = __umul24((uint32_t)(by + py), (uint32_t)g.w) + (uint32_t)bx;
      if constexpr (sizeof(T) == 2) xr.row[b][py] = __builtin_amdgcn_raw_buffer_load_b32(Xp.r, o * 2u, 0, 0);
      else xr.row[b][py] = __builtin_amdgcn_raw_buffer_load_b16(Xp.r, o, 0, 0);
      if constexpr (MASKED) {
        if constexpr (sizeof(TM) == 2) xr.mrow[b][py] = __builtin_amdgcn_raw_buffer_load_b32(Mp.r, o * 2u, 0, 0);
        else xr.mrow[b][py] = __builtin_amdgcn_raw_buffer_load_b16(Mp.r, o, 0, 0);
      }
    }
  }
}

// OUT: bit 0 writes the shift-compensated pels, bit 1 the winning offsets (both optional outputs
// of the search, chosen at launch so that the plain IV-PSNR search carries neither).
// fr and xr hold this phase's fill and tested pels, loaded during the previous phase; once they are
// consumed the loads of the next phase's (fill of picture next_fill, tested pels of next_x, tile
// column next_x0, when has_next) are issued into them, so their latency hides behind this phase's
// search.
template <bool MASKED>
using AccT = typename std::conditional<MASKED, uint64_t, uint32_t>::type;

template <typename T, typename TM, bool MASKED, int DIR, int OUT>
__device__ __forceinline__ void tile_dir420(const CpsArgs& a, const Geom& g, int x0, int y0, int lane, int wave,
                                            const int32_t (&gcd)[3], uint2* s_luma, uint32_t* s_cell,
                                            AccT<MASKED> (&acc)[2 * kKB][3], FillRegs& fr, XRegs& xr,
                                            const void* next_fill, const void* next_x, int next_x0, bool has_next) {
  // Opaque copies: keeps per-row address arithmetic from being hoisted out of the tile sweep and
  // held in registers across it.
  asm volatile("" : "+v"(lane));
  asm volatile("" : "+v"(wave));
  // R2T (DIR 0): P0 + gcd searched in P1; T2R (DIR 1): P1 - gcd searched in P0.
  uint16_t* scp_ptr = DIR == 0 ? a.scp1 : a.scp0;
  int16_t* best_out = DIR == 0 ? a.best0 : a.best1;
  constexpr int32_t sgn = DIR == 0 ? 1 : -1;
  const int32_t w0 = a.ws[0], w1 = a.ws[1], w2 = a.ws[2];

  // Tested values X = pel + sgn * gcd (not clipped: xc_CLIP_CURR_TST_RANGE = false,
  // libs/pmbb-ivqm/src/xCommonDefIVQM.h:36).
  constexpr uint32_t smask = sizeof(T) == 2 ? 0xffffu : 0xffu;
  constexpr uint32_t ssh = 8 * sizeof(T);
  int32_t x0v[kKB][2][2], a0[kKB][2][2], x1v[kKB], x2v[kKB], a1[kKB], a2[kKB];
  uint32_t mv[kKB][2][2];  // masked search: the tested pels' mask values
#pragma unroll
  for (int b = 0; b < kKB; b++) {
    x1v[b] = (int32_t)xr.cb[b];
    x2v[b] = (int32_t)xr.cr[b];
#pragma unroll
    for (int py = 0; py < 2; py++) {
      x0v[b][py][0] = (int32_t)(xr.row[b][py] & smask);
      x0v[b][py][1] = (int32_t)((xr.row[b][py] >> ssh) & smask);
      if constexpr (MASKED) {
        constexpr uint32_t mmask = sizeof(TM) == 2 ? 0xffffu : 0xffu;
        constexpr uint32_t msh = 8 * sizeof(TM);
        mv[b][py][0] = xr.mrow[b][py] & mmask;
        mv[b][py][1] = (xr.mrow[b][py] >> msh) & mmask;
      }
    }
  }

  __syncthreads();
  fill420_store<T, TM, MASKED>(fr, w0, w1, w2, s_luma, s_cell);
  __syncthreads();
  if (has_next) {
    const Buf<TM> mbuf(a.mask, a.mask_bytes);
    fill420_load<T, TM, MASKED>(Buf<T>(next_fill, g.frame_bytes), mbuf, g, next_x0, y0, fr);
    x420_load<T, TM, MASKED>(Buf<T>(next_x, g.frame_bytes), mbuf, g, next_x0, y0, lane, wave, xr);
  }

  // The key coefficients a_c = -64 w_c X_c.
  const int32_t m0 = -64 * w0, m1 = -64 * w1, m2 = -64 * w2;
#pragma unroll
  for (int b = 0; b < kKB; b++) {
    x1v[b] += sgn * gcd[1];
    x2v[b] += sgn * gcd[2];
    a1[b] = __mul24(m1, x1v[b]);
    a2[b] = __mul24(m2, x2v[b]);
#pragma unroll
    for (int py = 0; py < 2; py++)
#pragma unroll
      for (int px = 0; px < 2; px++) {
        x0v[b][py][px] += sgn * gcd[0];
        a0[b][py][px] = __mul24(m0, x0v[b][py][px]);
      }
  }

  // T(P, Q) for the 3 x 3 searched cells around each block's cell. Block b's cell is LDS cell
  // (1 + kKB wave + b, 2 + lane); its window cells are rows kKB wave + b + cr, columns lane + 1 + cc.
  int32_t Tq[kKB][3][3];
#pragma unroll
  for (int b = 0; b < kKB; b++)
#pragma unroll
    for (int cr = 0; cr < 3; cr++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        const uint32_t uv = s_cell[(kKB * wave + b + cr) * kCW + lane + 1 + cc];
        Tq[b][cr][cc] = mad24(a2[b], (int32_t)(uv >> 16), __mul24(a1[b], (int32_t)(uv & 0xffffu)));
      }

  int32_t best[kKB][2][2];
#pragma unroll
  for (int b = 0; b < kKB; b++)
#pragma unroll
    for (int py = 0; py < 2; py++)
#pragma unroll
      for (int px = 0; px < 2; px++) best[b][py][px] = INT_MAX;

  // Candidate luma rows of the lane: LDS rows 2 kKB wave + r, r = 0 .. 2 kKB + 3, taken as cell-row
  // pairs j (rows 2j, 2j + 1); columns 2 + 2 lane + cx, cx = 0 .. 5, as three 16-byte pairs.
  // Block b's pel (py, px) sees rows r in [2b + py, 2b + py + 4] and columns cx in [px, px + 4].
#pragma unroll
  for (int j = 0; j < kKB + 2; j++) {
    uint2 q[2][6];
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
      const uint4* row = reinterpret_cast<const uint4*>(s_luma + (2 * kKB * wave + 2 * j + rr) * kLW + 2 + 2 * lane);
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const uint4 v = row[m];
        q[rr][2 * m] = make_uint2(v.x, v.y);
        q[rr][2 * m + 1] = make_uint2(v.z, v.w);
      }
    }
#pragma unroll
    for (int b = 0; b < kKB; b++) {
      const int cr = j - b;
      if (cr < 0 || cr > 2) continue;
#pragma unroll
      for (int py = 0; py < 2; py++)
#pragma unroll
        for (int px = 0; px < 2; px++) {
          int32_t bp = best[b][py][px];
#pragma unroll
          for (int cc = 0; cc < 3; cc++) {
            int32_t m = INT_MAX;
#pragma unroll
            for (int rr = 0; rr < 2; rr++) {
              const int dr = 2 * cr + rr - py;  // window row index 0..4
              if (dr < 0 || dr > 4) continue;
#pragma unroll
              for (int c2 = 0; c2 < 2; c2++) {
                const int dc = 2 * cc + c2 - px;  // window column index 0..4
                if (dc < 0 || dc > 4) continue;
                m = min(m, mad24(a0[b][py][px], (int32_t)q[rr][2 * cc + c2].x, (int32_t)q[rr][2 * cc + c2].y));
              }
            }
            bp = min(bp, m + Tq[b][cr][cc]);  // every window cell holds at least one candidate
          }
          best[b][py][px] = bp;
          // Materialise the running minimum: keeps the min chain from being re-associated across
          // cell rows (which would keep every key of the window live).
          asm volatile("" : "+v"(best[b][py][px]));
        }
    }
    __builtin_amdgcn_sched_barrier(0);
  }

  // Winning candidates: per-component squared errors (and the optional outputs). Pels beyond the
  // picture (a partial last tile or band) add nothing; interior tiles skip the test.
  const bool full = x0 + kTW <= g.w && y0 + kTH <= g.h;
#pragma unroll
  for (int b = 0; b < kKB; b++)
#pragma unroll
    for (int py = 0; py < 2; py++)
#pragma unroll
      for (int px = 0; px < 2; px++) {
        const int ly = 2 + 2 * (kKB * wave + b) + py;
        const int lx = 4 + 2 * lane + px;
        const int idx = (best[b][py][px] - (5 * ly + lx) + 12) & 31;  // 5 (dy + 2) + (dx + 2)
        const int dy = (int)(__umul24((uint32_t)idx, 13u) >> 6);     // idx / 5 for idx <= 24
        const int dx = idx - 5 * dy;
        const int qy = ly - 2 + dy, qx = lx - 2 + dx;
        const uint32_t yv = s_luma[qy * kLW + qx].x;
        const uint32_t uv = s_cell[(qy >> 1) * kCW + (qx >> 1)];
        // |d| < 2^23: the low 32 bits of the 24-bit products are the reference's int32 squares.
        const int32_t d0 = x0v[b][py][px] - (int32_t)yv;
        const int32_t d1 = x1v[b] - (int32_t)(uv & 0xffffu);
        const int32_t d2 = x2v[b] - (int32_t)(uv >> 16);
        const int y = y0 + 2 * (kKB * wave + b) + py;
        const int x = x0 + 2 * lane + px;
        const bool inside = full || (x < g.w && y < g.h);
        if constexpr (MASKED) {
          // xCorrespPixelShiftSTD::CalcDistAsymmetricRowM: the winner's squared errors (int32,
          // sign-extended) times the tested pel's mask value; masked pels add nothing.
          const uint64_t m = inside ? (uint64_t)mv[b][py][px] : 0u;
          acc[2 * b + py][0] += (uint64_t)(uint32_t)__mul24(d0, d0) * m;
          acc[2 * b + py][1] += (uint64_t)(uint32_t)__mul24(d1, d1) * m;
          acc[2 * b + py][2] += (uint64_t)(uint32_t)__mul24(d2, d2) * m;
        } else {
          const uint32_t keep = inside ? 0xffffffffu : 0u;
          acc[2 * b + py][0] += (uint32_t)__mul24(d0, d0) & keep;
          acc[2 * b + py][1] += (uint32_t)__mul24(d1, d1) & keep;
          acc[2 * b + py][2] += (uint32_t)__mul24(d2, d2) & keep;
        }
        if constexpr ((OUT & 3) != 0) {
          if (inside) {
            const uint32_t o = __umul24((uint32_t)y, (uint32_t)g.w) + (uint32_t)x;
            if constexpr ((OUT & 2) != 0) best_out[o] = (int16_t)idx;
            if constexpr ((OUT & 1) != 0) {
              const Buf<uint16_t> scp(scp_ptr, 6u * (uint32_t)g.w * (uint32_t)g.h);
              const uint32_t plane = (uint32_t)(g.w * g.h);
              scp.st(o, (uint32_t)clampi((int32_t)yv - sgn * gcd[0], 0, a.maxv));
              scp.st(plane + o, (uint32_t)clampi((int32_t)(uv & 0xffffu) - sgn * gcd[1], 0, a.maxv));
              scp.st(2 * plane + o, (uint32_t)clampi((int32_t)(uv >> 16) - sgn * gcd[2], 0, a.maxv));
            }
          }
        }
      }
}

template <typename T, typename TM, bool MASKED, int OUT>
__global__ __launch_bounds__(256) void k_cps_fast420(CpsArgs a) {
  __shared__ __attribute__((aligned(16))) uint2 s_luma[kLH * kLW];
  __shared__ __attribute__((aligned(16))) uint32_t s_cell[kCH * kCW];

  // Nothing to do when a pel exceeds the bit depth: the host reports that error.
  const int32_t* gsrc;
  if constexpr (MASKED) {
    if (!a.msum->all_in_range) return;
    gsrc = a.msum->gcd;
  } else {
    if (!a.sum->in_range) return;
    gsrc = a.sum->gcd;
  }
  const Geom g = a.g;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // Workgroups are dealt to the XCDs round-robin; renumber them so that each XCD takes a contiguous
  // run of (band, segment) work and vertically adjacent bands share their halo rows in one L2.
#ifndef QMIV_CPS420_XCD
#define QMIV_CPS420_XCD 1
#endif
  int bid = blockIdx.x;
  if (QMIV_CPS420_XCD && (gridDim.x & 7) == 0) bid = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const int band = bid / a.nseg;
  const int seg = bid - band * a.nseg;
  const int y0 = band * kTH;
  const int tx_begin = seg * a.sweep;
  const int tx_end = min(a.tiles_x, tx_begin + a.sweep);
  const int32_t gcd[3] = {gsrc[0], gsrc[1], gsrc[2]};

  // Per-lane row sums over the sweep (the host bounds sweep * 2 * max squared error below 2^32;
  // masked sums, each square times a mask value of up to 16 bits, are kept in 64 bits).
  AccT<MASKED> acc0[2 * kKB][3], acc1[2 * kKB][3];
#pragma unroll
  for (int r = 0; r < 2 * kKB; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc0[r][c] = acc1[r][c] = 0;

  // R2T searches P1, T2R searches P0: the fills alternate P1, P0 tile after tile, each loaded one
  // phase ahead.
  FillRegs fr;
  XRegs xr;
  if (tx_begin < tx_end) {
    const Buf<TM> mbuf(a.mask, a.mask_bytes);
    fill420_load<T, TM, MASKED>(Buf<T>(a.p1, g.frame_bytes), mbuf, g, tx_begin * kTW, y0, fr);
    x420_load<T, TM, MASKED>(Buf<T>(a.p0, g.frame_bytes), mbuf, g, tx_begin * kTW, y0, lane, wave, xr);
  }
  for (int tx = tx_begin; tx < tx_end; tx++) {
    const int x0 = tx * kTW;
    tile_dir420<T, TM, MASKED, 0, OUT>(a, g, x0, y0, lane, wave, gcd, s_luma, s_cell, acc0, fr, xr, a.p0, a.p1, x0,
                                       true);
    tile_dir420<T, TM, MASKED, 1, OUT>(a, g, x0, y0, lane, wave, gcd, s_luma, s_cell, acc1, fr, xr, a.p1, a.p0,
                                       x0 + kTW, tx + 1 < tx_end);
  }

  // One partial per (segment, direction, component, row): reduce over the wave's 64 lanes.
#pragma unroll
  for (int r = 0; r < 2 * kKB; r++) {
    const int y = y0 + 2 * kKB * wave + r;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint64_t v0 = wave_sum((uint64_t)acc0[r][c]);
      const uint64_t v1 = wave_sum((uint64_t)acc1[r][c]);
      // The masked processor's R2T is the search of P1 - g in P0 (xIVPSNR.cpp:152-165): the
      // second direction here.
      const int d0 = MASKED ? 1 : 0, d1 = MASKED ? 0 : 1;
      if (lane == 0 && y < g.h) {
        a.partials[(((int64_t)seg * 2 + d0) * 3 + c) * g.h + y] = v0;
        a.partials[(((int64_t)seg * 2 + d1) * 3 + c) * g.h + y] = v1;
      }
    }
  }
}

}  // namespace

hipError_t launch_cps_fast420(const CpsArgs& a, hipStream_t stream) {
  if (a.g.cs != 1 || (a.g.w & 1) || (a.g.h & 1) || a.g.w < 2 || a.g.h < 2) return hipErrorInvalidValue;
  const int bands = (a.g.h + kTH - 1) / kTH;
  if (a.tiles_x != (a.g.w + kTW - 1) / kTW) return hipErrorInvalidValue;
  dim3 grid(bands * a.nseg), block(256);
  // Both directions' outputs are requested together (SCP for IV-SSIM, offsets for inspection).
  if ((a.scp0 == nullptr) != (a.scp1 == nullptr) || (a.best0 == nullptr) != (a.best1 == nullptr))
    return hipErrorInvalidValue;
  const int out = (a.scp0 ? 1 : 0) | (a.best0 ? 2 : 0);
  if (a.mask) {  // masked search: no optional outputs
    if (out != 0 || !a.msum || (a.mask_bytes_per_sample != 1 && a.mask_bytes_per_sample != 2))
      return hipErrorInvalidValue;
    if (a.g.bytes_per_sample == 1) {
      if (a.mask_bytes_per_sample == 1) k_cps_fast420<uint8_t, uint8_t, true, 0><<<grid, block, 0, stream>>>(a);
      else k_cps_fast420<uint8_t, uint16_t, true, 0><<<grid, block, 0, stream>>>(a);
    } else {
      if (a.mask_bytes_per_sample == 1) k_cps_fast420<uint16_t, uint8_t, true, 0><<<grid, block, 0, stream>>>(a);
      else k_cps_fast420<uint16_t, uint16_t, true, 0><<<grid, block, 0, stream>>>(a);
    }
    return hipGetLastError();
  }
  if (a.g.bytes_per_sample == 1) {
    switch (out) {
      case 0: k_cps_fast420<uint8_t, uint8_t, false, 0><<<grid, block, 0, stream>>>(a); break;
      case 1: k_cps_fast420<uint8_t, uint8_t, false, 1><<<grid, block, 0, stream>>>(a); break;
      case 2: k_cps_fast420<uint8_t, uint8_t, false, 2><<<grid, block, 0, stream>>>(a); break;
      default: k_cps_fast420<uint8_t, uint8_t, false, 3><<<grid, block, 0, stream>>>(a); break;
    }
  } else {
    switch (out) {
      case 0: k_cps_fast420<uint16_t, uint8_t, false, 0><<<grid, block, 0, stream>>>(a); break;
      case 1: k_cps_fast420<uint16_t, uint8_t, false, 1><<<grid, block, 0, stream>>>(a); break;
      case 2: k_cps_fast420<uint16_t, uint8_t, false, 2><<<grid, block, 0, stream>>>(a); break;
      default: k_cps_fast420<uint16_t, uint8_t, false, 3><<<grid, block, 0, stream>>>(a); break;
    }
  }
  return hipGetLastError();
}

void cps_fast420_tile(int32_t* tile_w, int32_t* tile_h) {
  *tile_w = kTW;
  *tile_h = kTH;
}

int cps_fast420_blocks_per_cu(int bytes_per_sample) {
  int n = 0;
  hipError_t e = bytes_per_sample == 1
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps_fast420<uint8_t, uint8_t, false, 0>, 256, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps_fast420<uint16_t, uint8_t, false, 0>, 256, 0);
  return e == hipSuccess && n > 0 ? n : 1;
}

}  // namespace qmiv
