// gfx950 corresponding-pixel-shift (CPS) search for 4:2:0 pictures, B = 2 (the fast path's
// parameter range, checked on the host).
//
// The key of a tested pel p against a candidate q of the searched picture Y is the one of the
// generic fast kernel (kernels_psnr.hip):
//   key(p, q) = 32 sum_c w_c Y_c(q) (Y_c(q) - 2 X_c(p)) + h(q),   h = 5 ly + lx,
// whose signed minimum over the 5 x 5 window is the reference's first strict minimum of the
// weighted squared error (libs/pmbb-ivqm/src/xCorrespPixelShiftSSE.cpp:40-71): the weighted error
// minus a term of p alone, times 32, plus the candidate's raster position in the window.
//
// With 4:2:0 pictures the chroma of both pictures is constant over 2 x 2 cells (the reference
// upsamples by nearest neighbour, libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49), so the key splits into
// a luma part of the candidate pel and a chroma part of the tested cell P and the searched cell Q:
//   key(p, q) = [L(q) + a0(p) Y0(q)] + T(P, Q),
//   L(q) = 32 w0 Y0(q)^2 + h(q),   a_c = -64 w_c X_c,
//   T(P, Q) = a1(P) U(Q) + a2(P) V(Q) + K(Q),   K(Q) = 32 (w1 U(Q)^2 + w2 V(Q)^2).
// The searched picture's tile is held in LDS as luma records {Y0, L} and cell records {U | V << 16,
// K}. A lane computes the nine T of its tested cell, adds them to its 36 candidates' L, and each of
// its four pels then needs one 24-bit multiply-add per candidate and a three-way minimum per two
// candidates. Every key keeps its exact value (only the grouping of the sum changes), so the
// winner and its tie-break are identical.
//
// Workgroup: 4 waves, tile 128 x 8 pels; each lane owns one 2 x 2 block (one tested chroma cell).
// Both pictures' tiles stay in LDS at once: the records of rows [y0 - 2, y0 + 10) x columns
// [x0 - 4, x0 + 132) and their cells. The tested pels of one direction are read from the other
// direction's searched tile, so a tile costs one pair of barriers and one fill per picture for both
// directions. The next tile's fill is loaded into registers while the current tile is searched.
#include <climits>
#include <type_traits>

#include "device_common.h"
#include "qmiv_internal.h"

namespace qmiv {
namespace {

constexpr int kTW = 128;          // tile columns: 64 lanes x 2
constexpr int kTH = 8;            // tile rows: 4 waves x 2
constexpr int kLW = kTW + 8;      // record columns [x0 - 4, x0 + kTW + 4)
constexpr int kLH = kTH + 4;      // record rows [y0 - 2, y0 + kTH + 2)
constexpr int kCW = kLW / 2;      // cells: record (ly, lx) lies in cell (ly >> 1, lx >> 1)
constexpr int kCH = kLH / 2;
constexpr int kQW = kLW / 4;      // record quads per row
constexpr int kLumaItems = kLH * kQW;  // one item: four records of one row
constexpr int kLumaSteps = (kLumaItems + 255) / 256;
constexpr int kCellItems = kCH * kQW;  // one item: two cells of one row
static_assert(kCellItems <= 256, "one cell step per fill");

// One picture's tile. Luma records {Y0, L}; cells {U | V << 16, K}.
struct alignas(16) PicTile {
  uint2 luma[kLH * kLW];
  uint2 cell[kCH * kCW];
};

// Masked search: a searched pel whose mask is 0 gets the record {Y0 = 0, L = kMaskedC + (h mod 32)}
// (h: its candidate index, so a winner still decodes to a window position). Its keys are then
// kMaskedC + (h mod 32) + T(P, Q) > 2^29 while every real key lies in (-2^29, 2^29) and
// |T| < 2^29 (the host checks both bounds), so it never beats an unmasked candidate. A tested pel
// whose mask is 0 contributes nothing; any other has itself as an unmasked candidate, so its
// winner is the first strict minimum over the unmasked ones, as in
// xCorrespPixelShiftSTD::FindBestPixelWithinBlockM (libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp).
constexpr int32_t kMaskedC = 1 << 30;

// One picture's fill as loaded, kept as the loaded words from the load (issued a tile ahead) to the
// store into LDS; nothing reads them in between, so the loads stay in flight across the search.
//   16-bit samples: l0 = Y0 | Y1 << 16, l1 = Y2 | Y3 << 16, u = U0 | U1 << 16, v = V0 | V1 << 16
//   8-bit samples:  l0 = Y0 | Y1 << 8 | Y2 << 16 | Y3 << 24, u = U0 | U1 << 8, v = V0 | V1 << 8
struct PicRegs {
  uint32_t l0[kLumaSteps], l1[kLumaSteps], u, v;
};
struct MaskRegs {  // the mask samples of the luma items, as loaded
  uint32_t m0[kLumaSteps], m1[kLumaSteps];
};

// The thread's fill items, fixed for the kernel: luma item i = tid + 256 k covers records
// (ly, lx .. lx + 3) with ly = i / kQW, lx = 4 (i mod kQW); cell item tid covers cells (cy, 2 cq)
// and (cy, 2 cq + 1) with cy = tid / kQW, cq = tid mod kQW.
struct Items {
  int32_t lofs[kLumaSteps];  // luma sample offset from the tile's record origin (y0 - 2, x0 - 4)
  int32_t cofs;              // chroma sample offset from the tile's cell origin
};

__device__ __forceinline__ bool luma_item(int k) {
  return kLumaItems % 256 == 0 || k < kLumaSteps - 1 || (int)threadIdx.x + 256 * k < kLumaItems;
}
__device__ __forceinline__ bool cell_item() { return (int)threadIdx.x < kCellItems; }

__device__ __forceinline__ bool tile_interior(const Geom& g, int x0, int y0) {
  // Aligned vector loads need w a multiple of 4 (the tile origin and every item column are).
  return x0 >= 4 && x0 + kTW + 4 <= g.w && y0 >= 2 && y0 + kTH + 2 <= g.h && (g.w & 3) == 0;
}

// Loads one picture's fill for the tile at (x0, y0). Interior tiles: one aligned 8-byte (u16) or
// 4-byte (u8) load per luma item and one 4-byte (u16) or 2-byte (u8) load per chroma plane and cell
// item, addressed as the item's fixed offset plus the tile origin (a scalar). Edge tiles: clamped
// single-sample loads; with even picture sizes and even tile origins, the two luma columns (rows)
// of a cell always clamp into one chroma sample, so cells stay uniform at the edges too.
template <typename T>
__device__ __forceinline__ void load_pic(const void* base, const Geom& g, const Items& it, int x0, int y0,
                                         bool interior, PicRegs& f) {
  const Buf<T> P(base, g.frame_bytes);
  constexpr uint32_t S = sizeof(T);
  constexpr uint32_t sh = 8 * sizeof(T);
  if (interior) {
    const uint32_t tl = (uint32_t)((y0 - 2) * g.w + x0 - 4) * S;
    const uint32_t tc = (uint32_t)(((y0 - 2) >> 1) * g.cw + ((x0 - 4) >> 1));
    const uint32_t tu = ((uint32_t)g.cb_off + tc) * S, tv = ((uint32_t)g.cr_off + tc) * S;
#pragma unroll
    for (int k = 0; k < kLumaSteps; k++) {
      if (!luma_item(k)) break;
      if constexpr (S == 2) {
        const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(P.r, (uint32_t)it.lofs[k] * S, tl, 0));
        f.l0[k] = w.x;
        f.l1[k] = w.y;
      } else {
        f.l0[k] = __builtin_amdgcn_raw_buffer_load_b32(P.r, (uint32_t)it.lofs[k], tl, 0);
        f.l1[k] = 0u;
      }
    }
    if (cell_item()) {
      if constexpr (S == 2) {
        f.u = __builtin_amdgcn_raw_buffer_load_b32(P.r, (uint32_t)it.cofs * S, tu, 0);
        f.v = __builtin_amdgcn_raw_buffer_load_b32(P.r, (uint32_t)it.cofs * S, tv, 0);
      } else {
        f.u = __builtin_amdgcn_raw_buffer_load_b16(P.r, (uint32_t)it.cofs, tu, 0);
        f.v = __builtin_amdgcn_raw_buffer_load_b16(P.r, (uint32_t)it.cofs, tv, 0);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    if (!luma_item(k)) break;
    const int i = threadIdx.x + 256 * k;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    const uint32_t row = __umul24((uint32_t)clampi(y0 - 2 + ly, 0, g.h - 1), (uint32_t)g.w);
    uint32_t q[4];
#pragma unroll
    for (int m = 0; m < 4; m++) q[m] = P.ld(row + (uint32_t)clampi(x0 - 4 + lx + m, 0, g.w - 1));
    if constexpr (S == 2) {
      f.l0[k] = q[0] | (q[1] << 16);
      f.l1[k] = q[2] | (q[3] << 16);
    } else {
      f.l0[k] = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
      f.l1[k] = 0u;
    }
  }
  if (cell_item()) {
    const int cy = threadIdx.x / kQW;
    const int cq = threadIdx.x - cy * kQW;
    const uint32_t crow = __umul24((uint32_t)(clampi(y0 - 2 + 2 * cy, 0, g.h - 1) >> 1), (uint32_t)g.cw);
    uint32_t uq[2], vq[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const uint32_t co = crow + (uint32_t)(clampi(x0 - 4 + 4 * cq + 2 * c, 0, g.w - 1) >> 1);
      uq[c] = P.ld((uint32_t)g.cb_off + co);
      vq[c] = P.ld((uint32_t)g.cr_off + co);
    }
    f.u = uq[0] | (uq[1] << sh);
    f.v = vq[0] | (vq[1] << sh);
  }
}

template <typename TM>
__device__ __forceinline__ void load_mask(const void* base, uint32_t bytes, const Geom& g, const Items& it, int x0,
                                          int y0, bool interior, MaskRegs& f) {
  const Buf<TM> M(base, bytes);
  constexpr uint32_t S = sizeof(TM);
  if (interior) {
    const uint32_t tl = (uint32_t)((y0 - 2) * g.w + x0 - 4) * S;
#pragma unroll
    for (int k = 0; k < kLumaSteps; k++) {
      if (!luma_item(k)) break;
      if constexpr (S == 2) {
        const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(M.r, (uint32_t)it.lofs[k] * S, tl, 0));
        f.m0[k] = w.x;
        f.m1[k] = w.y;
      } else {
        f.m0[k] = __builtin_amdgcn_raw_buffer_load_b32(M.r, (uint32_t)it.lofs[k], tl, 0);
        f.m1[k] = 0u;
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    if (!luma_item(k)) break;
    const int i = threadIdx.x + 256 * k;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    const uint32_t row = __umul24((uint32_t)clampi(y0 - 2 + ly, 0, g.h - 1), (uint32_t)g.w);
    uint32_t q[4];
#pragma unroll
    for (int m = 0; m < 4; m++) q[m] = M.ld(row + (uint32_t)clampi(x0 - 4 + lx + m, 0, g.w - 1));
    if constexpr (S == 2) {
      f.m0[k] = q[0] | (q[1] << 16);
      f.m1[k] = q[2] | (q[3] << 16);
    } else {
      f.m0[k] = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
      f.m1[k] = 0u;
    }
  }
}

// Builds one picture's records and cells from its loaded words. w0s, w1s, w2s: 32 w_c.
// Squares and weights in unsigned 24-bit multiplies (full rate): the host's eligibility check
// (non-negative weights, keys below 2^31) keeps samples <= 4095, so every operand, a square
// included, is below 2^24 and every product below 2^31.
template <typename T, typename TM, bool MASKED>
__device__ __forceinline__ void store_pic(const PicRegs& f, const MaskRegs& fm, uint32_t w0s, uint32_t w1s,
                                          uint32_t w2s, PicTile& s) {
  constexpr uint32_t sh = 8 * sizeof(T);
  constexpr uint32_t mask = sizeof(T) == 2 ? 0xffffu : 0xffu;
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    if (!luma_item(k)) break;
    const int i = threadIdx.x + 256 * k;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    uint32_t y[4];
    if constexpr (sizeof(T) == 2) {
      y[0] = f.l0[k] & 0xffffu; y[1] = f.l0[k] >> 16; y[2] = f.l1[k] & 0xffffu; y[3] = f.l1[k] >> 16;
    } else {
      y[0] = f.l0[k] & 0xffu; y[1] = (f.l0[k] >> 8) & 0xffu; y[2] = (f.l0[k] >> 16) & 0xffu; y[3] = f.l0[k] >> 24;
    }
    const uint32_t h = (uint32_t)(5 * ly + lx);
    uint32_t L[4];
#pragma unroll
    for (int m = 0; m < 4; m++) L[m] = __umul24(__umul24(y[m], y[m]), w0s) + h + (uint32_t)m;
    if constexpr (MASKED) {
      bool on[4];
      if constexpr (sizeof(TM) == 2) {
        on[0] = (fm.m0[k] & 0xffffu) != 0; on[1] = (fm.m0[k] >> 16) != 0;
        on[2] = (fm.m1[k] & 0xffffu) != 0; on[3] = (fm.m1[k] >> 16) != 0;
      } else {
        on[0] = (fm.m0[k] & 0xffu) != 0; on[1] = (fm.m0[k] & 0xff00u) != 0;
        on[2] = (fm.m0[k] & 0xff0000u) != 0; on[3] = (fm.m0[k] >> 24) != 0;
      }
#pragma unroll
      for (int m = 0; m < 4; m++)
        if (!on[m]) {
          y[m] = 0u;
          L[m] = (uint32_t)kMaskedC | ((h + (uint32_t)m) & 31u);
        }
    }
    uint4* dst = reinterpret_cast<uint4*>(&s.luma[ly * kLW + lx]);
    dst[0] = make_uint4(y[0], L[0], y[1], L[1]);
    dst[1] = make_uint4(y[2], L[2], y[3], L[3]);
  }
  if (cell_item()) {
    const int cy = threadIdx.x / kQW;
    const int cq = threadIdx.x - cy * kQW;
    const uint32_t u0 = f.u & mask, u1 = (f.u >> sh) & mask;
    const uint32_t v0 = f.v & mask, v1 = (f.v >> sh) & mask;
    const uint32_t k0 = __umul24(__umul24(u0, u0), w1s) + __umul24(__umul24(v0, v0), w2s);
    const uint32_t k1 = __umul24(__umul24(u1, u1), w1s) + __umul24(__umul24(v1, v1), w2s);
    *reinterpret_cast<uint4*>(&s.cell[cy * kCW + 2 * cq]) = make_uint4(u0 | (v0 << 16), k0, u1 | (v1 << 16), k1);
  }
}

// The tested pels' mask values (masked search), from the loaded mask words of the interior items:
// s_mask[(ly - 2) * kTW + lx - 4] for records ly in [2, 10), lx in [4, 132).
template <typename TM>
__device__ __forceinline__ void store_mask(const MaskRegs& fm, uint16_t* s_mask) {
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    if (!luma_item(k)) break;
    const int i = threadIdx.x + 256 * k;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    if (ly < 2 || ly >= kTH + 2 || lx < 4 || lx >= kTW + 4) continue;
    uint2 v;
    if constexpr (sizeof(TM) == 2) {
      v = make_uint2(fm.m0[k], fm.m1[k]);
    } else {
      const uint32_t w = fm.m0[k];
      v = make_uint2((w & 0xffu) | ((w & 0xff00u) << 8), ((w >> 16) & 0xffu) | ((w >> 24) << 16));
    }
    *reinterpret_cast<uint2*>(&s_mask[(ly - 2) * kTW + lx - 4]) = v;
  }
}

template <bool MASKED>
using AccT = typename std::conditional<MASKED, uint64_t, uint32_t>::type;

// Winner decode table: entry [2 py + px][idx], idx = 5 dy + dx the window raster index of the
// winner of block pel (py, px), holds the winner's byte offsets from the lane's window base:
// low 16 bits in the luma records, high 16 bits in the cells.
__device__ __forceinline__ void init_decode_table(uint32_t* s_tbl) {
  if (threadIdx.x < 128) {
    const int par = threadIdx.x >> 5, idx = threadIdx.x & 31;
    uint32_t e = 0;
    if (idx < 25) {
      const int dy = idx / 5, dx = idx - 5 * (idx / 5);
      const int ry = (par >> 1) + dy, rx = (par & 1) + dx;
      e = (uint32_t)((ry * kLW + rx) * 8) | ((uint32_t)(((ry >> 1) * kCW + (rx >> 1)) * 8) << 16);
    }
    s_tbl[threadIdx.x] = e;
  }
}

// Folds keys into a running minimum two at a time (one three-way minimum each).
template <int N>
__device__ __forceinline__ int32_t fold_min(int32_t best, const int32_t (&k)[N]) {
#pragma unroll
  for (int i = 0; i + 1 < N; i += 2) best = min(min(best, k[i]), k[i + 1]);
  if constexpr ((N & 1) != 0) best = min(best, k[N - 1]);
  return best;
}

// One direction over the tile: DIR 0 (R2T) tests P0 + gcd against P1, DIR 1 (T2R) tests P1 - gcd
// against P0. acc[py][c]: the lane's row sums (both of its columns) kept across the tile sweep.
// OUT: bit 0 writes the shift-compensated pels, bit 1 the winning offsets (both optional outputs
// of the search, chosen at launch so that the plain IV-PSNR search carries neither).
template <typename T, typename TM, bool MASKED, int DIR, int OUT>
__device__ __forceinline__ void search_dir(const CpsArgs& a, const Geom& g, int x0, int y0, int lane, int wave,
                                           const int32_t (&gcd)[3], const PicTile* s_pic, const uint32_t* s_tbl,
                                           const uint16_t* s_mask, AccT<MASKED> (&acc)[2][3]) {
  const PicTile& X = s_pic[DIR];      // tested picture's tile
  const PicTile& S = s_pic[1 - DIR];  // searched picture's tile
  constexpr int32_t sgn = DIR == 0 ? 1 : -1;
  const int32_t w0 = a.ws[0], w1 = a.ws[1], w2 = a.ws[2];

  // Tested values X = pel + sgn * gcd (not clipped: xc_CLIP_CURR_TST_RANGE = false,
  // libs/pmbb-ivqm/src/xCommonDefIVQM.h:36), from the tested picture's records and cell. A masked-out
  // tested pel's record holds 0: its value only feeds a square that its mask value of 0 cancels.
  const uint4 r0 = *reinterpret_cast<const uint4*>(&X.luma[(2 + 2 * wave) * kLW + 4 + 2 * lane]);
  const uint4 r1 = *reinterpret_cast<const uint4*>(&X.luma[(3 + 2 * wave) * kLW + 4 + 2 * lane]);
  const uint32_t tuv = X.cell[(1 + wave) * kCW + 2 + lane].x;
  int32_t xv[2][2], a0[2][2];
  xv[0][0] = (int32_t)r0.x + sgn * gcd[0];
  xv[0][1] = (int32_t)r0.z + sgn * gcd[0];
  xv[1][0] = (int32_t)r1.x + sgn * gcd[0];
  xv[1][1] = (int32_t)r1.z + sgn * gcd[0];
  const int32_t x1 = (int32_t)(tuv & 0xffffu) + sgn * gcd[1];
  const int32_t x2 = (int32_t)(tuv >> 16) + sgn * gcd[2];
  const int32_t m0 = -64 * w0;
#pragma unroll
  for (int py = 0; py < 2; py++)
#pragma unroll
    for (int px = 0; px < 2; px++) a0[py][px] = __mul24(m0, xv[py][px]);
  const int32_t a1 = __mul24(-64 * w1, x1), a2 = __mul24(-64 * w2, x2);

  // T(P, Q) for the 3 x 3 searched cells around the lane's cell: cells (wave + cr, lane + 1 + cc).
  int32_t Tq[3][3];
#pragma unroll
  for (int cr = 0; cr < 3; cr++)
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      const uint2 c = S.cell[(wave + cr) * kCW + lane + 1 + cc];
      Tq[cr][cc] = mad24(a2, (int32_t)(c.x >> 16), mad24(a1, (int32_t)(c.x & 0xffffu), (int32_t)c.y));
    }

  // Candidates: record rows 2 wave + 2 cr + rr (cell row cr, rr = 0, 1), columns 2 + 2 lane + cx,
  // cx = 0 .. 5 (cell column cx >> 1). Block pel (py, px) sees rows 2 cr + rr - py in [0, 4] of its
  // window and columns cx - px in [0, 4].
  int32_t best[2][2];
#pragma unroll
  for (int py = 0; py < 2; py++)
#pragma unroll
    for (int px = 0; px < 2; px++) best[py][px] = INT_MAX;
#pragma unroll
  for (int cr = 0; cr < 3; cr++) {
    int32_t Y[2][6], C[2][6];
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
      const uint4* row = reinterpret_cast<const uint4*>(&S.luma[(2 * wave + 2 * cr + rr) * kLW + 2 + 2 * lane]);
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const uint4 v = row[m];
        Y[rr][2 * m] = (int32_t)v.x;
        C[rr][2 * m] = (int32_t)v.y + Tq[cr][m];
        Y[rr][2 * m + 1] = (int32_t)v.z;
        C[rr][2 * m + 1] = (int32_t)v.w + Tq[cr][m];
      }
    }
#pragma unroll
    for (int py = 0; py < 2; py++)
#pragma unroll
      for (int px = 0; px < 2; px++) {
        // Keys of this cell row pair in the pel's window: rows 2 cr + rr - py in [0, 4].
        constexpr int kMax = 10;
        int32_t k[kMax];
        int n = 0;
#pragma unroll
        for (int rr = 0; rr < 2; rr++) {
          const int dr = 2 * cr + rr - py;
          if (dr < 0 || dr > 4) continue;
#pragma unroll
          for (int dc = 0; dc < 5; dc++) k[n++] = mad24(a0[py][px], Y[rr][dc + px], C[rr][dc + px]);
        }
        int32_t b = best[py][px];
        if (n == 10) {
          const int32_t (&kk)[10] = k;
          b = fold_min<10>(b, kk);
        } else {
          int32_t k5[5] = {k[0], k[1], k[2], k[3], k[4]};
          b = fold_min<5>(b, k5);
        }
        best[py][px] = b;
        // Materialise the running minimum: keeps the chain from being re-associated across cell
        // rows (which would keep every key of the window live).
        asm volatile("" : "+v"(best[py][px]));
      }
    __builtin_amdgcn_sched_barrier(0);
  }

  // Winners: key = h(q) (mod 32) and h(q) - h(p) = 5 (dy - 2) + (dx - 2), so the window index is
  // (best - h(p) + 12) mod 32; the decode table turns it into the winner's LDS offsets. Pels beyond
  // the picture (a partial last tile or band) add nothing; interior tiles skip the test.
  const bool full = x0 + kTW <= g.w && y0 + kTH <= g.h;
  const char* sl = reinterpret_cast<const char*>(S.luma) + ((2 * wave) * kLW + 2 + 2 * lane) * 8;
  const char* sc = reinterpret_cast<const char*>(S.cell) + (wave * kCW + 1 + lane) * 8;
  const int32_t nh = -2 - 10 * wave - 2 * lane;  // 12 - h(p) for the block's pel (0, 0)
#pragma unroll
  for (int py = 0; py < 2; py++)
#pragma unroll
    for (int px = 0; px < 2; px++) {
      const uint32_t idx = (uint32_t)(best[py][px] + nh - 5 * py - px) & 31u;
      const uint32_t t = s_tbl[(2 * py + px) * 32 + idx];
      const uint32_t yv = *reinterpret_cast<const uint32_t*>(sl + (t & 0xffffu));
      const uint32_t uv = *reinterpret_cast<const uint32_t*>(sc + (t >> 16));
      // |d| < 2^23: the low 32 bits of the 24-bit products are the reference's int32 squares.
      const int32_t d0 = xv[py][px] - (int32_t)yv;
      const int32_t d1 = x1 - (int32_t)(uv & 0xffffu);
      const int32_t d2 = x2 - (int32_t)(uv >> 16);
      const int y = y0 + 2 * wave + py;
      const int x = x0 + 2 * lane + px;
      const bool inside = full || (x < g.w && y < g.h);
      if constexpr (MASKED) {
        // xCorrespPixelShiftSTD::CalcDistAsymmetricRowM: the winner's squared errors (int32,
        // sign-extended) times the tested pel's mask value; masked pels add nothing.
        const uint64_t mv = inside ? (uint64_t)s_mask[(2 * wave + py) * kTW + 2 * lane + px] : 0u;
        acc[py][0] += (uint64_t)(uint32_t)__mul24(d0, d0) * mv;
        acc[py][1] += (uint64_t)(uint32_t)__mul24(d1, d1) * mv;
        acc[py][2] += (uint64_t)(uint32_t)__mul24(d2, d2) * mv;
      } else if (full) {
        acc[py][0] = (uint32_t)mad24(d0, d0, (int32_t)acc[py][0]);
        acc[py][1] = (uint32_t)mad24(d1, d1, (int32_t)acc[py][1]);
        acc[py][2] = (uint32_t)mad24(d2, d2, (int32_t)acc[py][2]);
      } else {
        const uint32_t keep = inside ? 0xffffffffu : 0u;
        acc[py][0] += (uint32_t)__mul24(d0, d0) & keep;
        acc[py][1] += (uint32_t)__mul24(d1, d1) & keep;
        acc[py][2] += (uint32_t)__mul24(d2, d2) & keep;
      }
      if constexpr ((OUT & 3) != 0) {
        if (inside) {
          const uint32_t o = __umul24((uint32_t)y, (uint32_t)g.w) + (uint32_t)x;
          if constexpr ((OUT & 2) != 0) (DIR == 0 ? a.best0 : a.best1)[o] = (int16_t)idx;
          if constexpr ((OUT & 1) != 0) {
            const Buf<uint16_t> scp(DIR == 0 ? a.scp1 : a.scp0, 6u * (uint32_t)g.w * (uint32_t)g.h);
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
  __shared__ PicTile s_pic[2];  // [0]: P0, [1]: P1
  __shared__ uint32_t s_tbl[128];
  __shared__ __attribute__((aligned(16))) uint16_t s_mask[MASKED ? kTH * kTW : 8];

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
  int bid = blockIdx.x;
  if ((gridDim.x & 7) == 0) bid = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const int band = bid / a.nseg;
  const int seg = bid - band * a.nseg;
  const int y0 = band * kTH;
  const int tx_begin = seg * a.sweep;
  const int tx_end = min(a.tiles_x, tx_begin + a.sweep);
  const int32_t gcd[3] = {gsrc[0], gsrc[1], gsrc[2]};
  const uint32_t w0s = 32u * (uint32_t)a.ws[0], w1s = 32u * (uint32_t)a.ws[1], w2s = 32u * (uint32_t)a.ws[2];

  Items it;
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    const int i = threadIdx.x + 256 * k;
    const int ly = i / kQW;
    it.lofs[k] = ly * g.w + (i - ly * kQW) * 4;
  }
  {
    const int cy = threadIdx.x / kQW;
    it.cofs = cy * g.cw + (threadIdx.x - cy * kQW) * 2;
  }
  init_decode_table(s_tbl);

  // Per-lane row sums over the sweep (the host bounds sweep * 2 * max squared error below 2^32;
  // masked sums, each square times a mask value of up to 16 bits, are kept in 64 bits).
  AccT<MASKED> acc0[2][3], acc1[2][3];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc0[r][c] = acc1[r][c] = 0;

  PicRegs f0, f1;
  MaskRegs fm;
  if (tx_begin < tx_end) {
    const int x0 = tx_begin * kTW;
    const bool in = tile_interior(g, x0, y0);
    load_pic<T>(a.p0, g, it, x0, y0, in, f0);
    load_pic<T>(a.p1, g, it, x0, y0, in, f1);
    if constexpr (MASKED) load_mask<TM>(a.mask, a.mask_bytes, g, it, x0, y0, in, fm);
  }
  for (int tx = tx_begin; tx < tx_end; tx++) {
    const int x0 = tx * kTW;
    __syncthreads();  // the previous tile's searches are done with LDS
    store_pic<T, TM, MASKED>(f0, fm, w0s, w1s, w2s, s_pic[0]);
    store_pic<T, TM, MASKED>(f1, fm, w0s, w1s, w2s, s_pic[1]);
    if constexpr (MASKED) store_mask<TM>(fm, s_mask);
    __syncthreads();
    if (tx + 1 < tx_end) {
      const bool in = tile_interior(g, x0 + kTW, y0);
      load_pic<T>(a.p0, g, it, x0 + kTW, y0, in, f0);
      load_pic<T>(a.p1, g, it, x0 + kTW, y0, in, f1);
      if constexpr (MASKED) load_mask<TM>(a.mask, a.mask_bytes, g, it, x0 + kTW, y0, in, fm);
    }
    search_dir<T, TM, MASKED, 0, OUT>(a, g, x0, y0, lane, wave, gcd, s_pic, s_tbl, s_mask, acc0);
    search_dir<T, TM, MASKED, 1, OUT>(a, g, x0, y0, lane, wave, gcd, s_pic, s_tbl, s_mask, acc1);
  }

  // One sum per (segment, direction, component, row): reduce over the wave's 64 lanes, then either
  // write the segment's partial or add it into the row (integer sums: any order gives the same).
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int y = y0 + 2 * wave + r;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint64_t v0 = wave_sum((uint64_t)acc0[r][c]);
      const uint64_t v1 = wave_sum((uint64_t)acc1[r][c]);
      // The masked processor's R2T is the search of P1 - g in P0 (xIVPSNR.cpp:152-165): the
      // second direction here.
      const int d0 = MASKED ? 1 : 0, d1 = MASKED ? 0 : 1;
      if (lane == 0 && y < g.h) {
        if (a.rows) {
          atomicAdd(reinterpret_cast<unsigned long long*>(&a.rows[((int64_t)d0 * 3 + c) * g.h + y]), (unsigned long long)v0);
          atomicAdd(reinterpret_cast<unsigned long long*>(&a.rows[((int64_t)d1 * 3 + c) * g.h + y]), (unsigned long long)v1);
        } else {
          a.partials[(((int64_t)seg * 2 + d0) * 3 + c) * g.h + y] = v0;
          a.partials[(((int64_t)seg * 2 + d1) * 3 + c) * g.h + y] = v1;
        }
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
