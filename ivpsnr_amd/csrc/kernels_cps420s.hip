// gfx950 corresponding-pixel-shift (CPS) search for 4:2:0 pictures of at most 10 bits, B = 2,
// with 16-bit key operands and two tested chroma cells per lane.
//
// The key is the one of the 4:2:0 kernel (kernels_cps420.hip): for a tested pel p with value
// X = P(p) + g and a candidate q of the searched picture Y,
//   key(p, q) = 32 sum_c w_c Y_c(q) (Y_c(q) - 2 X_c) + h(q),   h = 5 ly + lx,
//             = L(q) + a0(p) Y0(q) + T(P, Q),
//   L(q) = 32 w0 Y0(q)^2 + h(q),   a0 = -64 w0 X0,
//   T(P, Q) = a1(P) U(Q) + a2(P) V(Q) + K(Q),   K = 32 (w1 U^2 + w2 V^2),
// whose signed minimum over the 5 x 5 window is the first strict minimum of the weighted squared
// error of libs/pmbb-ivqm/src/xCorrespPixelShiftSSE.cpp:40-71 (raster order breaks ties).
//
// 16-bit operands. With samples below 2^10 the factor 64 w_c splits into 32 on the searched side and
// 2 w_c on the tested side: Y' = 32 Y fits a signed 16-bit half, and so does -2 w_c X_c for the
// weights the host admits. Each product a0 Y0 = (-2 w0 X0)(32 Y0) is then one v_mad_i32_i16 whose
// operands are read straight from packed sample pairs (op_sel picks the half), and each cell term is
// one v_dot2_i32_i16: T = (-2 w1 X1, -2 w2 X2) . (32 U, 32 V) + K. All sums are exact in int32, so
// every key keeps its value and the winner and its tie-break are those of the reference.
//
// Workgroup: one direction of one tile of 128 x 16 pels, 4 waves of 4 rows, each lane a 2 x 4 block
// (two tested chroma cells stacked), so one candidate row pair serves up to both cells. The R2T and
// T2R workgroups of a tile are adjacent in dispatch order (one XCD, one L2). One direction per
// workgroup keeps a lane's registers within four waves per SIMD and its LDS within four workgroups
// per CU, so each SIMD holds waves of four independent workgroups. LDS holds the searched picture's
// tile and the tested picture's samples of the tile itself:
//   L  int32 32 w0 Y^2, rows [y0 - 2, y0 + 18) x columns [x0 - 4, x0 + 132)
//   Y  int16 32 Y over the same rows and columns
//   cell {32 U | 32 V << 16, K + 5 ly + lx of the cell's top-left record}.
// The candidate's raster term h moves out of the luma record: the cell constant carries it for the
// cell's top-left record, and the other three records add 1, 5 or 6 in the same add that forms
// L + T (an add3 with an inline constant), so a key keeps exactly its value.
//   TY 16-bit tested samples, rows [y0, y0 + 16) x columns [x0, x0 + 128); TUV U | V << 16 per cell.
// The next tile's fill is loaded into registers while the current tile is searched.
#include <climits>

#include "device_common.h"
#include "qmiv_internal.h"

namespace qmiv {
namespace {

constexpr int kTW = 128;                 // tile columns: 64 lanes x 2
constexpr int kTH = 16;                  // tile rows: 4 waves x 4
constexpr int kLW = kTW + 8;             // record columns [x0 - 4, x0 + kTW + 4)
constexpr int kLH = kTH + 4;             // record rows [y0 - 2, y0 + kTH + 2)
constexpr int kCW = kLW / 2;             // cells: record (ly, lx) lies in cell (ly >> 1, lx >> 1)
constexpr int kCH = kLH / 2;
constexpr int kQW = kLW / 4;             // record quads per row
constexpr int kThreads = 256;           // 4 waves, 4 rows each
// Fill items: thread t < kFillThreads covers quad column q = t mod kQW of the rows r + 14 k,
// r = t / kQW (one luma item: four records of a row; one cell item: two cells of a cell row).
constexpr int kFillRows = kThreads / kQW;
constexpr int kFillThreads = kFillRows * kQW;
constexpr int kLumaSteps = (kLH + kFillRows - 1) / kFillRows;
constexpr int kCellSteps = (kCH + kFillRows - 1) / kFillRows;
static_assert(kFillThreads <= kThreads, "one fill item per thread and step");

struct alignas(16) Tile {
  int32_t L[kLH * kLW];
  uint2 cell[kCH * kCW];
  int16_t Y[kLH * kLW];
};
// The workgroup's LDS: the decode table, the searched picture's tile, the tested picture's samples.
// Every search access is a per-lane byte base plus a compile-time offset (the DS instructions'
// offset field).
struct alignas(16) Lds {
  uint32_t tbl[128];
  Tile pic;
  uint16_t ty[kTH * kTW];
  uint32_t tuv[(kTH / 2) * (kTW / 2)];
};
constexpr uint32_t kOffTbl = 0;
constexpr uint32_t kOffPic = 512;
constexpr uint32_t kTileBytes = sizeof(Tile);
constexpr uint32_t kOffL = 0, kOffCell = kLH * kLW * 4, kOffY = kOffCell + kCH * kCW * 8;
constexpr uint32_t kOffTY = kOffPic + kTileBytes;
constexpr uint32_t kOffTUV = kOffTY + kTH * kTW * 2;
static_assert(kOffTUV + (kTH / 2) * (kTW / 2) * 4 == sizeof(Lds), "LDS layout");
static_assert(sizeof(Lds) <= 65536, "LDS offsets fit the DS offset field");
// Tested-sample fill: thread t covers quad column t mod 32 of the rows t / 32 + 8 k (luma) and the
// cell pair t mod 32 of cell row t / 32.
constexpr int kTestLumaSteps = kTH * (kTW / 4) / kThreads;
static_assert(kTH * (kTW / 4) % kThreads == 0 && (kTH / 2) * (kTW / 4) == kThreads, "tested fill items");
// Per-lane byte bases of the lane's search accesses within a tile: the L record of (4 wave, 2 lane),
// the Y record of (4 wave, 2 lane) and the cell (2 wave, lane).
struct Bases {
  uint32_t l, y, c;
  uint32_t t, tc;  // tested sample (4 wave, 2 lane) and tested cell (2 wave, lane) in ty / tuv
};

// The tested picture's samples of the tile as loaded: 16-bit samples ty0 = Y0 | Y1 << 16,
// ty1 = Y2 | Y3 << 16 (8-bit: ty0 = the four bytes), u = U0 | U1 << 16, v = V0 | V1 << 16 (8-bit: bytes).
struct TestRegs {
  uint32_t ty0[kTestLumaSteps], ty1[kTestLumaSteps], u, v;
};

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_mul(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) * __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pack16(int32_t lo, int32_t hi) {
  return ((uint32_t)lo & 0xffffu) | ((uint32_t)hi << 16);
}

// d = a.half(S0) * b.half(S1) + c, signed 16-bit halves, 32-bit sum (one instruction; ALU only).
template <int S0, int S1>
__device__ __forceinline__ int32_t mad16(uint32_t a, uint32_t b, int32_t c) {
  int32_t d;
  if constexpr (S0 == 0 && S1 == 0) asm("v_mad_i32_i16 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  else if constexpr (S0 == 0 && S1 == 1) asm("v_mad_i32_i16 %0, %1, %2, %3 op_sel:[0,1,0,0]" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  else if constexpr (S0 == 1 && S1 == 0) asm("v_mad_i32_i16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  else asm("v_mad_i32_i16 %0, %1, %2, %3 op_sel:[1,1,0,0]" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
// d = a.lo * b.lo + a.hi * b.hi + c, signed 16-bit halves (one instruction; ALU only).
__device__ __forceinline__ int32_t dot2i(uint32_t a, uint32_t b, int32_t c) {
  int32_t d;
  asm("v_dot2_i32_i16 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
// d = a.lo * b.lo + a.hi * b.hi + c, unsigned 16-bit halves (one instruction; ALU only).
__device__ __forceinline__ uint32_t dot2u(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t d;
  asm("v_dot2_u32_u16 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
// a + b + c (one v_add3_u32; ALU only), for the raster offsets of the records after the first.
template <int C>
__device__ __forceinline__ int32_t add3c(int32_t a, int32_t b) {
  int32_t d;
  asm("v_add3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "i"(C));
  return d;
}

// One picture's fill as loaded, kept as the loaded words from the load (issued a tile ahead) to the
// store into LDS.
//   16-bit samples: l0 = Y0 | Y1 << 16, l1 = Y2 | Y3 << 16, u = U0 | U1 << 16, v = V0 | V1 << 16
//   8-bit samples:  l0 = Y0 | Y1 << 8 | Y2 << 16 | Y3 << 24, u = U0 | U1 << 8, v = V0 | V1 << 8
struct PicRegs {
  uint32_t l0[kLumaSteps], l1[kLumaSteps], u[kCellSteps], v[kCellSteps];
};

// Masked search: a searched pel whose mask is 0 gets the records {32 Y = 0, L = kMaskedC}, so its
// keys are kMaskedC + T(P, Q) + h > 2^29 while every real key lies in (-2^29, 2^29) and |T| < 2^29
// (host check), and it never beats an unmasked candidate; kMaskedC = 0 (mod 32) keeps the raster
// term, so a winner still decodes to a window position. A tested pel whose mask is 0 contributes
// nothing; any other has itself as an unmasked candidate, so its winner is the first strict minimum
// over the unmasked ones, as in xCorrespPixelShiftSTD::FindBestPixelWithinBlockM
// (libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp).
constexpr int32_t kMaskedC = 1 << 30;
// The mask samples of a luma item (searched tile) as loaded: 16-bit masks m0 = M0 | M1 << 16,
// m1 = M2 | M3 << 16; 8-bit masks m0 = the four bytes.
struct MaskRegs {
  uint32_t m0[kLumaSteps], m1[kLumaSteps];
};
// The tested tile's mask samples as loaded (the items of load_test).
struct TestMaskRegs {
  uint32_t m0[kTestLumaSteps], m1[kTestLumaSteps];
};

// The thread's fill position: fill row r and quad column q (r >= kFillRows: no fill items).
struct Items {
  int32_t r, q;
};

__device__ __forceinline__ bool luma_item(const Items& it, int k) { return it.r + kFillRows * k < kLH; }
__device__ __forceinline__ bool cell_item(const Items& it, int k) { return it.r + kFillRows * k < kCH; }

__device__ __forceinline__ bool tile_interior(const Geom& g, int x0, int y0) {
  // Aligned vector loads need w a multiple of 4 (the tile origin and every item column are).
  return x0 >= 4 && x0 + kTW + 4 <= g.w && y0 >= 2 && y0 + kTH + 2 <= g.h && (g.w & 3) == 0;
}

// Loads one picture's fill for the tile at (x0, y0): aligned vector loads on interior tiles,
// clamped single-sample loads at the picture edges (with even picture sizes and even tile origins
// the two luma columns (rows) of a cell always clamp into one chroma sample).
template <typename T>
__device__ __forceinline__ void load_pic(const void* base, const Geom& g, const Items& it, int x0, int y0,
                                         bool interior, PicRegs& f) {
  const Buf<T> P(base, g.frame_bytes);
  constexpr uint32_t S = sizeof(T);
  constexpr uint32_t sh = 8 * sizeof(T);
  if (interior) {
    // Every load is issued by every thread: an item past the tile reads at an offset beyond the
    // buffer, which the hardware range check answers with 0. Loads under per-thread branches would
    // make the compiler wait for them at the branch's end instead of a tile later at their use.
    const uint32_t lofs = (uint32_t)(it.r * g.w + 4 * it.q) * S;
    const uint32_t cofs = (uint32_t)(it.r * g.cw + 2 * it.q) * S;
    const uint32_t tl = (uint32_t)((y0 - 2) * g.w + x0 - 4) * S;
    const uint32_t tc = (uint32_t)(((y0 - 2) >> 1) * g.cw + ((x0 - 4) >> 1));
    const uint32_t tu = ((uint32_t)g.cb_off + tc) * S, tv = ((uint32_t)g.cr_off + tc) * S;
    constexpr uint32_t kNone = 0x80000000u;  // beyond any frame (frames are below 2^31 bytes)
#pragma unroll
    for (int k = 0; k < kLumaSteps; k++) {
      const uint32_t vo = luma_item(it, k) ? lofs : kNone;
      const uint32_t so = tl + (uint32_t)(kFillRows * k * g.w) * S;
      if constexpr (S == 2) {
        const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(P.r, vo, so, 0));
        f.l0[k] = w.x;
        f.l1[k] = w.y;
      } else {
        f.l0[k] = __builtin_amdgcn_raw_buffer_load_b32(P.r, vo, so, 0);
        f.l1[k] = 0u;
      }
    }
#pragma unroll
    for (int k = 0; k < kCellSteps; k++) {
      const uint32_t vo = cell_item(it, k) ? cofs : kNone;
      const uint32_t co = (uint32_t)(kFillRows * k * g.cw) * S;
      if constexpr (S == 2) {
        f.u[k] = __builtin_amdgcn_raw_buffer_load_b32(P.r, vo, tu + co, 0);
        f.v[k] = __builtin_amdgcn_raw_buffer_load_b32(P.r, vo, tv + co, 0);
      } else {
        f.u[k] = __builtin_amdgcn_raw_buffer_load_b16(P.r, vo, tu + co, 0);
        f.v[k] = __builtin_amdgcn_raw_buffer_load_b16(P.r, vo, tv + co, 0);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    if (!luma_item(it, k)) break;
    const int ly = it.r + kFillRows * k;
    const int lx = 4 * it.q;
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
#pragma unroll
  for (int k = 0; k < kCellSteps; k++) {
    if (!cell_item(it, k)) break;
    const int cy = it.r + kFillRows * k;
    const uint32_t crow = __umul24((uint32_t)(clampi(y0 - 2 + 2 * cy, 0, g.h - 1) >> 1), (uint32_t)g.cw);
    uint32_t uq[2], vq[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const uint32_t co = crow + (uint32_t)(clampi(x0 - 4 + 4 * it.q + 2 * c, 0, g.w - 1) >> 1);
      uq[c] = P.ld((uint32_t)g.cb_off + co);
      vq[c] = P.ld((uint32_t)g.cr_off + co);
    }
    f.u[k] = uq[0] | (uq[1] << sh);
    f.v[k] = vq[0] | (vq[1] << sh);
  }
}

// Loads the mask samples of the searched tile's luma items for the tile at (x0, y0), as load_pic
// loads the luma: vector loads on interior tiles (an item past the tile reads 0 through the range
// check), clamped single-sample loads at the edges. The mask's luma plane starts at offset 0.
template <typename TM>
__device__ __forceinline__ void load_mask(const void* base, uint32_t bytes, const Geom& g, const Items& it, int x0,
                                          int y0, bool interior, MaskRegs& f) {
  const Buf<TM> M(base, bytes);
  constexpr uint32_t S = sizeof(TM);
  if (interior) {
    const uint32_t lofs = (uint32_t)(it.r * g.w + 4 * it.q) * S;
    const uint32_t tl = (uint32_t)((y0 - 2) * g.w + x0 - 4) * S;
    constexpr uint32_t kNone = 0x80000000u;
#pragma unroll
    for (int k = 0; k < kLumaSteps; k++) {
      const uint32_t vo = luma_item(it, k) ? lofs : kNone;
      const uint32_t so = tl + (uint32_t)(kFillRows * k * g.w) * S;
      if constexpr (S == 2) {
        const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(M.r, vo, so, 0));
        f.m0[k] = w.x;
        f.m1[k] = w.y;
      } else {
        f.m0[k] = __builtin_amdgcn_raw_buffer_load_b32(M.r, vo, so, 0);
        f.m1[k] = 0u;
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    if (!luma_item(it, k)) break;
    const int ly = it.r + kFillRows * k;
    const int lx = 4 * it.q;
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

// Builds one picture's tile from its loaded words. lb: the thread's LDS byte offset of its first L
// record (r, 4 q), which is also that of its first cell (r, 2 q) and twice that of its first Y record;
// hc = 10 r + 4 q, the raster term 5 ly + lx of the first cell's top-left record. w0s = 32 w0;
// w12 = w1 | w2 << 16. Samples are below 2^10 (host check), so every shift by 5 stays inside its
// 16-bit half.
template <typename T, typename TM, bool MASKED>
__device__ __forceinline__ void store_pic(const PicRegs& f, const MaskRegs& fm, const Items& it, uint32_t lb,
                                          uint32_t hc, uint32_t w0s, uint32_t w12, char* lds) {
  constexpr uint32_t tile = kOffPic;
#pragma unroll
  for (int k = 0; k < kLumaSteps; k++) {
    if (!luma_item(it, k)) break;
    uint32_t p0, p1;  // Y0 | Y1 << 16, Y2 | Y3 << 16
    if constexpr (sizeof(T) == 2) {
      p0 = f.l0[k];
      p1 = f.l1[k];
    } else {
      p0 = __builtin_amdgcn_perm(0u, f.l0[k], 0x0c010c00u);
      p1 = __builtin_amdgcn_perm(0u, f.l0[k], 0x0c030c02u);
    }
    const uint32_t y0 = p0 & 0xffffu, y1 = p0 >> 16, y2 = p1 & 0xffffu, y3 = p1 >> 16;
    int4 L;
    L.x = (int32_t)__umul24(__umul24(y0, y0), w0s);
    L.y = (int32_t)__umul24(__umul24(y1, y1), w0s);
    L.z = (int32_t)__umul24(__umul24(y2, y2), w0s);
    L.w = (int32_t)__umul24(__umul24(y3, y3), w0s);
    uint32_t s0 = p0 << 5, s1 = p1 << 5;
    if constexpr (MASKED) {
      // masked records: 32 Y = 0 and L = kMaskedC
      uint32_t k0, k1, k2, k3;  // mask samples of the four records
      if constexpr (sizeof(TM) == 2) {
        k0 = fm.m0[k] & 0xffffu;
        k1 = fm.m0[k] >> 16;
        k2 = fm.m1[k] & 0xffffu;
        k3 = fm.m1[k] >> 16;
      } else {
        k0 = fm.m0[k] & 0xffu;
        k1 = (fm.m0[k] >> 8) & 0xffu;
        k2 = (fm.m0[k] >> 16) & 0xffu;
        k3 = fm.m0[k] >> 24;
      }
      L.x = k0 ? L.x : kMaskedC;
      L.y = k1 ? L.y : kMaskedC;
      L.z = k2 ? L.z : kMaskedC;
      L.w = k3 ? L.w : kMaskedC;
      s0 &= (k0 ? 0xffffu : 0u) | (k1 ? 0xffff0000u : 0u);
      s1 &= (k2 ? 0xffffu : 0u) | (k3 ? 0xffff0000u : 0u);
    }
    const uint32_t ko = (uint32_t)(kFillRows * k * kLW);
    *reinterpret_cast<int4*>(lds + lb + tile + kOffL + ko * 4) = L;
    *reinterpret_cast<uint2*>(lds + (lb >> 1) + tile + kOffY + ko * 2) = make_uint2(s0, s1);
  }
#pragma unroll
  for (int k = 0; k < kCellSteps; k++) {
    if (!cell_item(it, k)) break;
    uint32_t uva, uvb;  // U | V << 16 of the item's two cells
    if constexpr (sizeof(T) == 2) {
      uva = __builtin_amdgcn_perm(f.v[k], f.u[k], 0x05040100u);
      uvb = __builtin_amdgcn_perm(f.v[k], f.u[k], 0x07060302u);
    } else {
      uva = __builtin_amdgcn_perm(f.v[k], f.u[k], 0x0c040c00u);
      uvb = __builtin_amdgcn_perm(f.v[k], f.u[k], 0x0c050c01u);
    }
    const uint32_t sa = uva << 5, sb = uvb << 5;
    // 32 (w1 U^2 + w2 V^2) = (32 U, 32 V) . (w1 U, w2 V), plus the cell's raster term.
    const uint32_t h = hc + (uint32_t)(10 * kFillRows * k);
    const uint32_t ka = dot2u(sa, pk_mul(uva, w12), h);
    const uint32_t kb = dot2u(sb, pk_mul(uvb, w12), h + 2u);
    const uint32_t ko = (uint32_t)(kFillRows * k * kCW);
    *reinterpret_cast<uint4*>(lds + lb + tile + kOffCell + ko * 8) = make_uint4(sa, ka, sb, kb);
  }
}

// Loads the tested picture's samples of the tile at (x0, y0): rows [y0, y0 + 16) and columns
// [x0, x0 + 128), one vector load per item. The width is a multiple of the tile width, so the columns
// lie inside the picture; rows past it are clamped (they only feed pels whose sums are dropped).
template <typename T>
__device__ __forceinline__ void load_test(const void* base, const Geom& g, int x0, int y0, TestRegs& f) {
  const Buf<T> P(base, g.frame_bytes);
  constexpr uint32_t S = sizeof(T);
  const int t = threadIdx.x;
  const int r = t >> 5, q = t & 31;
#pragma unroll
  for (int k = 0; k < kTestLumaSteps; k++) {
    const uint32_t o = ((uint32_t)min(y0 + r + 8 * k, g.h - 1) * (uint32_t)g.w + (uint32_t)(x0 + 4 * q)) * S;
    if constexpr (S == 2) {
      const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(P.r, o, 0, 0));
      f.ty0[k] = w.x;
      f.ty1[k] = w.y;
    } else {
      f.ty0[k] = __builtin_amdgcn_raw_buffer_load_b32(P.r, o, 0, 0);
      f.ty1[k] = 0u;
    }
  }
  const uint32_t co = ((uint32_t)(min(y0 + 2 * r, g.h - 1) >> 1) * (uint32_t)g.cw + (uint32_t)((x0 >> 1) + 2 * q)) * S;
  if constexpr (S == 2) {
    f.u = __builtin_amdgcn_raw_buffer_load_b32(P.r, co, (uint32_t)g.cb_off * S, 0);
    f.v = __builtin_amdgcn_raw_buffer_load_b32(P.r, co, (uint32_t)g.cr_off * S, 0);
  } else {
    f.u = __builtin_amdgcn_raw_buffer_load_b16(P.r, co, (uint32_t)g.cb_off * S, 0);
    f.v = __builtin_amdgcn_raw_buffer_load_b16(P.r, co, (uint32_t)g.cr_off * S, 0);
  }
}

// Stores the tested samples: ty[row][col] as 16-bit samples, tuv[cell row][cell col] = U | V << 16.
template <typename T>
__device__ __forceinline__ void store_test(const TestRegs& f, uint32_t tb, char* lds) {
#pragma unroll
  for (int k = 0; k < kTestLumaSteps; k++) {
    uint32_t p0, p1;
    if constexpr (sizeof(T) == 2) {
      p0 = f.ty0[k];
      p1 = f.ty1[k];
    } else {
      p0 = __builtin_amdgcn_perm(0u, f.ty0[k], 0x0c010c00u);
      p1 = __builtin_amdgcn_perm(0u, f.ty0[k], 0x0c030c02u);
    }
    *reinterpret_cast<uint2*>(lds + tb * 2 + kOffTY + (uint32_t)(8 * k * kTW * 2)) = make_uint2(p0, p1);
  }
  uint32_t uva, uvb;
  if constexpr (sizeof(T) == 2) {
    uva = __builtin_amdgcn_perm(f.v, f.u, 0x05040100u);
    uvb = __builtin_amdgcn_perm(f.v, f.u, 0x07060302u);
  } else {
    uva = __builtin_amdgcn_perm(f.v, f.u, 0x0c040c00u);
    uvb = __builtin_amdgcn_perm(f.v, f.u, 0x0c050c01u);
  }
  // cell pair (t / 32, 2 (t mod 32)): byte offset (64 (t / 32) + 2 (t mod 32)) * 4 = 2 tb
  *reinterpret_cast<uint2*>(lds + tb * 2 + kOffTUV) = make_uint2(uva, uvb);
}

// Loads the mask samples of the tested tile (the items of load_test): rows [y0, y0 + 16), columns
// [x0, x0 + 128), rows past the picture clamped (their sums are dropped).
template <typename TM>
__device__ __forceinline__ void load_test_mask(const void* base, uint32_t bytes, const Geom& g, int x0, int y0,
                                               TestMaskRegs& f) {
  const Buf<TM> M(base, bytes);
  constexpr uint32_t S = sizeof(TM);
  const int t = threadIdx.x;
  const int r = t >> 5, q = t & 31;
#pragma unroll
  for (int k = 0; k < kTestLumaSteps; k++) {
    const uint32_t o = ((uint32_t)min(y0 + r + 8 * k, g.h - 1) * (uint32_t)g.w + (uint32_t)(x0 + 4 * q)) * S;
    if constexpr (S == 2) {
      const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(M.r, o, 0, 0));
      f.m0[k] = w.x;
      f.m1[k] = w.y;
    } else {
      f.m0[k] = __builtin_amdgcn_raw_buffer_load_b32(M.r, o, 0, 0);
      f.m1[k] = 0u;
    }
  }
}

// Stores the tested tile's mask samples as 16-bit values: s_tm[row][col].
template <typename TM>
__device__ __forceinline__ void store_test_mask(const TestMaskRegs& f, uint32_t tb, uint16_t* s_tm) {
#pragma unroll
  for (int k = 0; k < kTestLumaSteps; k++) {
    uint32_t p0, p1;
    if constexpr (sizeof(TM) == 2) {
      p0 = f.m0[k];
      p1 = f.m1[k];
    } else {
      p0 = __builtin_amdgcn_perm(0u, f.m0[k], 0x0c010c00u);
      p1 = __builtin_amdgcn_perm(0u, f.m0[k], 0x0c030c02u);
    }
    *reinterpret_cast<uint2*>(&s_tm[tb + (uint32_t)(8 * k * kTW)]) = make_uint2(p0, p1);
  }
}

// Winner decode table: entry [2 (py & 1) + px][idx], idx = 5 dy + dx the window raster index of
// the winner of block pel (py, px), holds the winner's byte offsets from the pel row's window base:
// low 16 bits in the Y records (from record (4 wave + py, 2 lane + 2)), high 16 bits in the cells
// (from cell (2 wave + (py >> 1), lane + 1)).
__device__ __forceinline__ void init_decode_table(uint32_t* s_tbl) {
  if (threadIdx.x < 128) {
    const int par = threadIdx.x >> 5, idx = threadIdx.x & 31;
    const int pyp = par >> 1, px = par & 1;
    uint32_t e = 0;
    if (idx < 25) {
      const int dy = idx / 5, dx = idx - 5 * (idx / 5);
      const uint32_t yoff = (uint32_t)((dy * kLW + px + dx) * 2);
      const uint32_t coff = (uint32_t)((((pyp + dy) >> 1) * kCW + ((px + dx) >> 1)) * 8);
      e = yoff | (coff << 16);
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

// Keys of block pel (py, px) against one candidate row pair (cell row cr) folded into its running
// minimum. Yc: packed 32 Y of the row pair's six candidate columns; C: L + T.
template <int CR, int PY, int PX>
__device__ __forceinline__ void pel_keys(uint32_t A, const uint32_t (&Yc)[2][3], const int32_t (&C)[2][6],
                                         int32_t& best) {
  constexpr int r0 = 2 * CR - PY;  // window row of the pair's first record row
  constexpr bool ra = r0 >= 0 && r0 <= 4, rb = r0 + 1 >= 0 && r0 + 1 <= 4;
  constexpr int n = (ra ? 5 : 0) + (rb ? 5 : 0);
  if constexpr (n > 0) {
    int32_t k[n];
    int j = 0;
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
      if ((rr == 0 && !ra) || (rr == 1 && !rb)) continue;
#pragma unroll
      for (int dx = 0; dx < 5; dx++) {
        const int kc = dx + PX;
        if ((kc & 1) == 0) k[j++] = mad16<PX, 0>(A, Yc[rr][kc >> 1], C[rr][kc]);
        else k[j++] = mad16<PX, 1>(A, Yc[rr][kc >> 1], C[rr][kc]);
      }
    }
    best = fold_min<n>(best, k);
    // Materialise the running minimum: keeps the chain from being re-associated across row pairs
    // (not volatile, so the LDS loads of later row pairs may still be scheduled across it).
    asm("" : "+v"(best));
  }
}

// One candidate row pair for tested cell P (rows 2P, 2P + 1 of the block), which sees the cell rows
// P .. P + 2 of the lane's four.
template <int CR, int P>
__device__ __forceinline__ void cell_pass(const uint32_t (&A)[4], uint32_t B, const int32_t (&Lc)[2][6],
                                          const uint32_t (&Yc)[2][3], const uint2 (&cq)[3], int32_t (&best)[4][2]) {
  if constexpr (CR >= P && CR <= P + 2) {
    int32_t T[3];
#pragma unroll
    for (int c = 0; c < 3; c++) T[c] = dot2i(B, cq[c].x, (int32_t)cq[c].y);
    // C = L + T + the record's raster offset from its cell's top-left record (5 rr + (kc & 1)).
    int32_t C[2][6];
#pragma unroll
    for (int kc = 0; kc < 6; kc += 2) {
      C[0][kc] = Lc[0][kc] + T[kc >> 1];
      C[0][kc + 1] = add3c<1>(Lc[0][kc + 1], T[kc >> 1]);
      C[1][kc] = add3c<5>(Lc[1][kc], T[kc >> 1]);
      C[1][kc + 1] = add3c<6>(Lc[1][kc + 1], T[kc >> 1]);
    }
    pel_keys<CR, 2 * P, 0>(A[2 * P], Yc, C, best[2 * P][0]);
    pel_keys<CR, 2 * P, 1>(A[2 * P], Yc, C, best[2 * P][1]);
    pel_keys<CR, 2 * P + 1, 0>(A[2 * P + 1], Yc, C, best[2 * P + 1][0]);
    pel_keys<CR, 2 * P + 1, 1>(A[2 * P + 1], Yc, C, best[2 * P + 1][1]);
  }
}

// Sum over the wave of a 32-bit value, complete in lane 63: an inclusive scan within each row of 16
// lanes (DPP row shifts, zero shifted in), then the row totals carried by row broadcasts.
__device__ __forceinline__ uint32_t wave_total_u32(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return v;
}

// One candidate row pair (cell row CR of the lane's four) as read from LDS: the L records and packed
// 32 Y of record rows 4 wave + 2 CR + rr, columns 2 lane + 2 .. 2 lane + 7, and the cells of cell
// row 2 wave + CR, columns lane + 1 .. lane + 3.
struct RowPair {
  int32_t L[2][6];
  uint32_t Y[2][3];
  uint2 cq[3];
};

template <int CR>
__device__ __forceinline__ void load_row_pair(const char* lds, const Bases& b, RowPair& d) {
  constexpr uint32_t tile = kOffPic;
#pragma unroll
  for (int rr = 0; rr < 2; rr++) {
    constexpr uint32_t lo0 = tile + kOffL + (uint32_t)((2 * CR) * kLW + 2) * 4;
    constexpr uint32_t yo0 = tile + kOffY + (uint32_t)((2 * CR) * kLW + 2) * 2;
    const uint32_t lo = lo0 + (uint32_t)(rr * kLW * 4);
    const uint32_t yo = yo0 + (uint32_t)(rr * kLW * 2);
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const uint2 v = *reinterpret_cast<const uint2*>(lds + b.l + lo + 8 * m);
      d.L[rr][2 * m] = (int32_t)v.x;
      d.L[rr][2 * m + 1] = (int32_t)v.y;
      d.Y[rr][m] = *reinterpret_cast<const uint32_t*>(lds + b.y + yo + 4 * m);
    }
  }
  constexpr uint32_t co = tile + kOffCell + (uint32_t)(CR * kCW + 1) * 8;
#pragma unroll
  for (int c = 0; c < 3; c++) d.cq[c] = *reinterpret_cast<const uint2*>(lds + b.c + co + 8 * c);
}

template <int CR>
__device__ __forceinline__ void row_pair(const RowPair& d, const uint32_t (&A)[4], const uint32_t (&B)[2],
                                         int32_t (&best)[4][2]) {
  cell_pass<CR, 0>(A, B[0], d.L, d.Y, d.cq, best);
  cell_pass<CR, 1>(A, B[1], d.L, d.Y, d.cq, best);
}

// One direction over the tile: DIR 0 (R2T) tests P0 + gcd against P1, DIR 1 (T2R) tests P1 - gcd
// against P0. acc[c][py]: the lane's row sums (both of its columns) kept across the tile sweep. Every
// column of the tile lies inside the picture (the width is a multiple of the tile width); rows past it
// are dropped with their sums. OUT: bit 0 writes the shift-compensated pels, bit 1 the winning offsets.
template <int DIR, int OUT, bool MASKED>
__device__ __forceinline__ void search_dir(const CpsArgs& a, const Geom& g, int x0, int y0, int lane, int wave,
                                           const int32_t (&gcd)[3], const char* lds, const Bases& b,
                                           const uint16_t* s_tm, uint32_t (&acc)[3][4]) {
  constexpr uint32_t st = kOffPic;  // searched picture's tile
  constexpr int32_t sgn = DIR == 0 ? 1 : -1;
  const int32_t w0 = a.ws[0], w1 = a.ws[1], w2 = a.ws[2];

  // Tested values X = pel + sgn * gcd (not clipped: xc_CLIP_CURR_TST_RANGE = false,
  // libs/pmbb-ivqm/src/xCommonDefIVQM.h:36), two pels per packed register: xs[py] for the block's
  // row py, x12[P] = X1 | X2 << 16 for its cell P; A = -2 w0 X0 and B = (-2 w1 X1, -2 w2 X2) packed.
  const uint32_t g0 = pack16(sgn * gcd[0], sgn * gcd[0]);
  const uint32_t g12 = pack16(sgn * gcd[1], sgn * gcd[2]);
  const uint32_t m0 = pack16(-2 * w0, -2 * w0);
  const uint32_t m12 = pack16(-2 * w1, -2 * w2);
  uint32_t xs[4], A[4], x12[2], B[2];
#pragma unroll
  for (int py = 0; py < 4; py++) {
    // tested samples (4 wave + py, 2 lane .. 2 lane + 1)
    const uint32_t ty = *reinterpret_cast<const uint32_t*>(lds + b.t + kOffTY + (uint32_t)(py * kTW * 2));
    xs[py] = pk_add(ty, g0);
    A[py] = pk_mul(xs[py], m0);
  }
#pragma unroll
  for (int P = 0; P < 2; P++) {
    // tested cell (2 wave + P, lane)
    const uint32_t tc = *reinterpret_cast<const uint32_t*>(lds + b.tc + kOffTUV + (uint32_t)(P * (kTW / 2) * 4));
    x12[P] = pk_add(tc, g12);
    B[P] = pk_mul(x12[P], m12);
  }

  int32_t best[4][2];
#pragma unroll
  for (int py = 0; py < 4; py++) best[py][0] = best[py][1] = INT_MAX;
  {
    RowPair d;
    load_row_pair<0>(lds, b, d);
    row_pair<0>(d, A, B, best);
  }
  {
    RowPair d;
    load_row_pair<1>(lds, b, d);
    row_pair<1>(d, A, B, best);
  }
  {
    RowPair d;
    load_row_pair<2>(lds, b, d);
    row_pair<2>(d, A, B, best);
  }
  {
    RowPair d;
    load_row_pair<3>(lds, b, d);
    row_pair<3>(d, A, B, best);
  }

  // Winners: key = h(q) (mod 32) and h(q) - h(p) = 5 (dy - 2) + (dx - 2), so the window index is
  // (best - h(p) + 12) mod 32; the decode table turns it into the winner's LDS offsets from the pel
  // row's window base, record (4 wave + py, 2 lane + 2) and cell (2 wave + (py >> 1), lane + 1).
  // 12 - h(p) for the block's pel (0, 0): h(p) = 5 (2 + 4 wave) + 4 + 2 lane.
  const int32_t nh = -2 - 20 * wave - 2 * lane;
#pragma unroll
  for (int py = 0; py < 4; py++) {
    const int P = py >> 1;
    // masked search: the mask values of the pel pair (4 wave + py, 2 lane .. 2 lane + 1)
    uint32_t mp = 0u;
    if constexpr (MASKED) mp = *reinterpret_cast<const uint32_t*>(&s_tm[(4 * wave + py) * kTW + 2 * lane]);
#pragma unroll
    for (int px = 0; px < 2; px++) {
      const uint32_t ti = ((uint32_t)(best[py][px] + nh - 5 * py - px) << 2) & 124u;
      const uint32_t t = *reinterpret_cast<const uint32_t*>(lds + kOffTbl + ((py & 1) * 2 + px) * 128 + ti);
      const uint32_t yw = (uint32_t)*reinterpret_cast<const uint16_t*>(
                              lds + b.y + (t & 0xffffu) + st + kOffY + (uint32_t)(py * kLW + 2) * 2) >> 5;
      const uint32_t uvw = *reinterpret_cast<const uint32_t*>(
                               lds + b.c + (t >> 16) + st + kOffCell + (uint32_t)(P * kCW + 1) * 8) >> 5;
      const int32_t xv = (int32_t)(int16_t)(px == 0 ? (xs[py] & 0xffffu) : (xs[py] >> 16));
      const int32_t d0 = xv - (int32_t)yw;
      const uint32_t d12 = pk_sub(x12[P], uvw);  // X1 - U | X2 - V << 16, |d| < 2^15
      const int y = y0 + 4 * wave + py;
      const int x = x0 + 2 * lane + px;
      if constexpr (MASKED) {
        // xCorrespPixelShiftSTD::CalcDistAsymmetricRowM: the winner's squared errors times the tested
        // pel's mask value (squares below 2^24; the host bounds the lane sums below 2^32)
        const uint32_t mv = px == 0 ? (mp & 0xffffu) : (mp >> 16);
        acc[0][py] += __umul24((uint32_t)__mul24(d0, d0), mv);
        acc[1][py] += __umul24((uint32_t)mad16<0, 0>(d12, d12, 0), mv);
        acc[2][py] += __umul24((uint32_t)mad16<1, 1>(d12, d12, 0), mv);
      } else {
        acc[0][py] = (uint32_t)mad24(d0, d0, (int32_t)acc[0][py]);
        acc[1][py] = (uint32_t)mad16<0, 0>(d12, d12, (int32_t)acc[1][py]);
        acc[2][py] = (uint32_t)mad16<1, 1>(d12, d12, (int32_t)acc[2][py]);
      }
      if constexpr ((OUT & 3) != 0) {
        if (x < g.w && y < g.h) {
          const uint32_t o = __umul24((uint32_t)y, (uint32_t)g.w) + (uint32_t)x;
          if constexpr ((OUT & 2) != 0) {
            const uint32_t idx = (uint32_t)(best[py][px] + nh - 5 * py - px) & 31u;
            (DIR == 0 ? a.best0 : a.best1)[o] = (int16_t)idx;
          }
          if constexpr ((OUT & 1) != 0) {
            const Buf<uint16_t> scp(DIR == 0 ? a.scp1 : a.scp0, 6u * (uint32_t)g.w * (uint32_t)g.h);
            const uint32_t plane = (uint32_t)(g.w * g.h);
            scp.st(o, (uint32_t)clampi((int32_t)yw - sgn * gcd[0], 0, a.maxv));
            scp.st(plane + o, (uint32_t)clampi((int32_t)(uvw & 0xffffu) - sgn * gcd[1], 0, a.maxv));
            scp.st(2 * plane + o, (uint32_t)clampi((int32_t)(uvw >> 16) - sgn * gcd[2], 0, a.maxv));
          }
        }
      }
    }
  }
  // Keeps the next direction's set-up from overlapping this decode (register pressure).
  __builtin_amdgcn_sched_barrier(0);
}

// The tile sweep of one direction: DIR 0 searches P1 for P0 + gcd (R2T), DIR 1 P0 for P1 - gcd (T2R).
template <typename T, typename TM, bool MASKED, int OUT, int DIR>
__device__ __forceinline__ void sweep_dir(const CpsArgs& a, const Geom& g, char* lds, uint16_t* s_tm, int y0,
                                          int tx_begin, int tx_end, int lane, int wave, const int32_t (&gcd)[3],
                                          uint32_t w0s, uint32_t w12, uint32_t (&acc)[3][4]) {
  const void* sp = DIR == 0 ? a.p1 : a.p0;  // searched picture
  const void* tp = DIR == 0 ? a.p0 : a.p1;  // tested picture
  Items it0;
  it0.r = (int)threadIdx.x < kFillThreads ? (int)threadIdx.x / kQW : kLH;  // kLH: no items
  it0.q = (int)threadIdx.x - ((int)threadIdx.x / kQW) * kQW;
  PicRegs f;
  TestRegs ft;
  MaskRegs fm;
  TestMaskRegs ftm;
  if (tx_begin < tx_end) {
    const int x0 = tx_begin * kTW;
    load_pic<T>(sp, g, it0, x0, y0, tile_interior(g, x0, y0), f);
    load_test<T>(tp, g, x0, y0, ft);
    if constexpr (MASKED) {
      load_mask<TM>(a.mask, a.mask_bytes, g, it0, x0, y0, tile_interior(g, x0, y0), fm);
      load_test_mask<TM>(a.mask, a.mask_bytes, g, x0, y0, ftm);
    }
  }
  for (int tx = tx_begin; tx < tx_end; tx++) {
    const int x0 = tx * kTW;
    // The fill position, formed in the loop so that what derives from it is not kept in registers
    // across the sweep.
    Items it = it0;
    asm volatile("" : "+v"(it.r), "+v"(it.q));
    it.r &= 31;
    it.q &= 63;
    const uint32_t lb = (uint32_t)(it.r * kLW + 4 * it.q) * 4u;
    const uint32_t hc = (uint32_t)(10 * it.r + 4 * it.q);
    uint32_t tt = threadIdx.x;
    asm volatile("" : "+v"(tt));
    tt &= 255u;
    const uint32_t tb = (tt >> 5) * (uint32_t)kTW + 4u * (tt & 31u);  // tested item: sample (t / 32, 4 (t mod 32))
    __syncthreads();  // the previous tile's searches are done with LDS
    store_pic<T, TM, MASKED>(f, fm, it, lb, hc, w0s, w12, lds);
    store_test<T>(ft, tb, lds);
    if constexpr (MASKED) store_test_mask<TM>(ftm, tb, s_tm);
    __syncthreads();
    if (tx + 1 < tx_end) {
      load_pic<T>(sp, g, it, x0 + kTW, y0, tile_interior(g, x0 + kTW, y0), f);
      load_test<T>(tp, g, x0 + kTW, y0, ft);
      if constexpr (MASKED) {
        load_mask<TM>(a.mask, a.mask_bytes, g, it, x0 + kTW, y0, tile_interior(g, x0 + kTW, y0), fm);
        load_test_mask<TM>(a.mask, a.mask_bytes, g, x0 + kTW, y0, ftm);
      }
    }
    // The lane's bases, formed in the loop so that the compiler keeps a few registers rather than
    // one per access.
    uint32_t ln = (uint32_t)lane;
    asm volatile("" : "+v"(ln));
    ln &= 63u;  // known small and non-negative: the compiler folds the constant offsets into the DS instructions
    Bases b;
    b.l = ln * 8u + (uint32_t)wave * (4 * kLW * 4);
    b.y = ln * 4u + (uint32_t)wave * (4 * kLW * 2);
    b.c = ln * 8u + (uint32_t)wave * (2 * kCW * 8);
    b.t = ln * 4u + (uint32_t)wave * (4 * kTW * 2);
    b.tc = ln * 4u + (uint32_t)wave * (2 * (kTW / 2) * 4);
    search_dir<DIR, OUT, MASKED>(a, g, x0, y0, (int)ln, wave, gcd, lds, b, s_tm, acc);
  }
}

template <typename T, typename TM, bool MASKED, int OUT>
__global__ __launch_bounds__(kThreads, 4) void k_cps420s(CpsArgs a) {
  __shared__ Lds smem;
  __shared__ __attribute__((aligned(16))) uint16_t s_tm[MASKED ? kTH * kTW : 8];  // tested tile's mask
  char* lds = reinterpret_cast<char*>(&smem);

  // Nothing to do when a pel exceeds the bit depth: the host reports that error.
  if constexpr (MASKED) {
    if (!a.msum->all_in_range) return;
  } else {
    if (!a.sum->in_range) return;
  }
  const Geom g = a.g;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) & 3;  // rows 4 wave .. 4 wave + 3
  // Workgroups are dealt to the XCDs round-robin; renumber them so that each XCD takes a contiguous
  // run of (band, segment, direction) work: the two directions of a tile and vertically adjacent
  // bands share their rows in one L2.
  int bid = blockIdx.x;
  if ((gridDim.x & 7) == 0) bid = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const int t2r = bid & 1;
  const int work = bid >> 1;
  const int band = work / a.nseg;
  const int seg = work - band * a.nseg;
  const int y0 = band * kTH;
  const int tx_begin = seg * a.sweep;
  const int tx_end = min(a.tiles_x, tx_begin + a.sweep);
  const int32_t* gsrc = MASKED ? a.msum->gcd : a.sum->gcd;
  const int32_t gcd[3] = {gsrc[0], gsrc[1], gsrc[2]};
  const uint32_t w0s = 32u * (uint32_t)a.ws[0];
  const uint32_t w12 = (uint32_t)a.ws[1] | ((uint32_t)a.ws[2] << 16);
  init_decode_table(smem.tbl);

  // Per-lane row sums of the workgroup's direction over the sweep; the host bounds
  // sweep * 128 * max squared error below 2^32, so the lane sums and their wave totals fit 32 bits.
  uint32_t acc[3][4];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[c][r] = 0u;
  if (t2r == 0) sweep_dir<T, TM, MASKED, OUT, 0>(a, g, lds, s_tm, y0, tx_begin, tx_end, lane, wave, gcd, w0s, w12, acc);
  else sweep_dir<T, TM, MASKED, OUT, 1>(a, g, lds, s_tm, y0, tx_begin, tx_end, lane, wave, gcd, w0s, w12, acc);

  if constexpr (MASKED) {
    // The segment's masked sums per (component, row): 64-bit wave totals from the lane sums' 16-bit
    // halves, written to the segment's partials (one workgroup per segment, direction and band). The
    // masked processor's R2T is the search of P1 - g in P0 (xIVPSNR.cpp:152-165): the second
    // direction here.
    const int d = t2r ? 0 : 1;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int y = y0 + 4 * wave + r;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const uint64_t lo = wave_total_u32(acc[c][r] & 0xffffu);
        const uint64_t hi = wave_total_u32(acc[c][r] >> 16);
        if (lane == 63 && y < g.h) a.partials[(((int64_t)seg * 2 + d) * 3 + c) * g.h + y] = lo + (hi << 16);
      }
    }
    return;
  }

  // One sum per (component, row) of the workgroup's direction: reduce over the wave's 64 lanes in
  // 32 bits, then add it into the row (integer sums: any order gives the same).
  uint64_t* rows = a.rows + (t2r ? 3 * (int64_t)g.h : 0);
  uint32_t tot[3][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) tot[c][r] = wave_total_u32(acc[c][r]);
  if (lane == 63) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int y = y0 + 4 * wave + r;
      if (y < g.h) {
#pragma unroll
        for (int c = 0; c < 3; c++)
          atomicAdd(reinterpret_cast<unsigned long long*>(&rows[(int64_t)c * g.h + y]), (unsigned long long)tot[c][r]);
      }
    }
  }
}

}  // namespace

hipError_t launch_cps420s(const CpsArgs& a, hipStream_t stream) {
  if (a.g.cs != 1 || a.g.w % kTW != 0 || (a.g.h & 1) || a.g.w < kTW || a.g.h < 2 || a.g.bd > 10)
    return hipErrorInvalidValue;
  const bool masked = a.mask != nullptr;
  // The unmasked search adds into rows; the masked one writes per-segment partials.
  if (masked ? (a.partials == nullptr || a.msum == nullptr) : a.rows == nullptr) return hipErrorInvalidValue;
  const int bands = (a.g.h + kTH - 1) / kTH;
  if (a.tiles_x != (a.g.w + kTW - 1) / kTW || a.sweep < 1 || a.nseg != (a.tiles_x + a.sweep - 1) / a.sweep)
    return hipErrorInvalidValue;
  dim3 grid(2 * bands * a.nseg), block(kThreads);
  if ((a.scp0 == nullptr) != (a.scp1 == nullptr) || (a.best0 == nullptr) != (a.best1 == nullptr))
    return hipErrorInvalidValue;
  const int out = (a.scp0 ? 1 : 0) | (a.best0 ? 2 : 0);
  if (masked) {  // no optional outputs
    if (out != 0 || (a.mask_bytes_per_sample != 1 && a.mask_bytes_per_sample != 2)) return hipErrorInvalidValue;
    if (a.g.bytes_per_sample == 1) {
      if (a.mask_bytes_per_sample == 1) k_cps420s<uint8_t, uint8_t, true, 0><<<grid, block, 0, stream>>>(a);
      else k_cps420s<uint8_t, uint16_t, true, 0><<<grid, block, 0, stream>>>(a);
    } else {
      if (a.mask_bytes_per_sample == 1) k_cps420s<uint16_t, uint8_t, true, 0><<<grid, block, 0, stream>>>(a);
      else k_cps420s<uint16_t, uint16_t, true, 0><<<grid, block, 0, stream>>>(a);
    }
    return hipGetLastError();
  }
  if (a.g.bytes_per_sample == 1) {
    switch (out) {
      case 0: k_cps420s<uint8_t, uint8_t, false, 0><<<grid, block, 0, stream>>>(a); break;
      case 1: k_cps420s<uint8_t, uint8_t, false, 1><<<grid, block, 0, stream>>>(a); break;
      case 2: k_cps420s<uint8_t, uint8_t, false, 2><<<grid, block, 0, stream>>>(a); break;
      default: k_cps420s<uint8_t, uint8_t, false, 3><<<grid, block, 0, stream>>>(a); break;
    }
  } else {
    switch (out) {
      case 0: k_cps420s<uint16_t, uint8_t, false, 0><<<grid, block, 0, stream>>>(a); break;
      case 1: k_cps420s<uint16_t, uint8_t, false, 1><<<grid, block, 0, stream>>>(a); break;
      case 2: k_cps420s<uint16_t, uint8_t, false, 2><<<grid, block, 0, stream>>>(a); break;
      default: k_cps420s<uint16_t, uint8_t, false, 3><<<grid, block, 0, stream>>>(a); break;
    }
  }
  return hipGetLastError();
}

void cps420s_tile(int32_t* tile_w, int32_t* tile_h) {
  *tile_w = kTW;
  *tile_h = kTH;
}

int cps420s_blocks_per_cu(int bytes_per_sample) {
  int n = 0;
  hipError_t e = bytes_per_sample == 1
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps420s<uint8_t, uint8_t, false, 0>, kThreads, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps420s<uint16_t, uint8_t, false, 0>, kThreads, 0);
  return e == hipSuccess && n > 0 ? n : 1;
}

}  // namespace qmiv
