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
// Workgroup: 4 waves, tile 128 x 16 pels; each lane owns two vertically stacked 2 x 2 blocks
// (one tested chroma cell each). LDS holds the searched picture's luma records {Y0, C} for
// rows [y0 - 2, y0 + 18) x columns [x0 - 4, x0 + 132) and its chroma cells {U | V << 16}.
#include <climits>

#include "device_common.h"
#include "qmiv_internal.h"

namespace qmiv {
namespace {

constexpr int kTW = 128;          // tile columns: 64 lanes x 2
#ifndef QMIV_CPS420_KB
#define QMIV_CPS420_KB 2
#endif
constexpr int kKB = QMIV_CPS420_KB;  // 2 x 2 blocks per lane, stacked vertically
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

template <typename T>
__device__ __forceinline__ void fill420(const Buf<T>& Yp, const Geom& g, int x0, int y0, int32_t w0, int32_t w1,
                                        int32_t w2, uint2* s_luma, uint32_t* s_cell) {
  const uint32_t cb_off = (uint32_t)g.cb_off, cr_off = (uint32_t)g.cr_off;
  const bool interior = x0 >= 4 && x0 + kTW + 4 <= g.w && y0 >= 2 && y0 + kTH + 2 <= g.h && (g.w & 3) == 0;
  uint32_t yv[kFillSteps][4], cu[kFillSteps][2], cv[kFillSteps][2];
#pragma unroll
  for (int k = 0; k < kFillSteps; k++) {
    const int i = threadIdx.x + 256 * k;
    if (kFillItems % 256 != 0 && k == kFillSteps - 1 && i >= kFillItems) break;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    if (interior) {
      const int gy = y0 - 2 + ly;
      const int gx = x0 - 4 + lx;
      const uint32_t lo = (uint32_t)(gy * g.w + gx);
      const uint32_t co = (uint32_t)((gy >> 1) * g.cw + (gx >> 1));
      if constexpr (sizeof(T) == 2) {
        const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(Yp.r, lo * 2u, 0, 0));
        yv[k][0] = w.x & 0xffffu; yv[k][1] = w.x >> 16; yv[k][2] = w.y & 0xffffu; yv[k][3] = w.y >> 16;
        const uint32_t wu = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, (cb_off + co) * 2u, 0, 0);
        const uint32_t wv = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, (cr_off + co) * 2u, 0, 0);
        cu[k][0] = wu & 0xffffu; cu[k][1] = wu >> 16; cv[k][0] = wv & 0xffffu; cv[k][1] = wv >> 16;
      } else {
        const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, lo, 0, 0);
        yv[k][0] = w & 0xffu; yv[k][1] = (w >> 8) & 0xffu; yv[k][2] = (w >> 16) & 0xffu; yv[k][3] = w >> 24;
        const uint32_t wu = __builtin_amdgcn_raw_buffer_load_b16(Yp.r, cb_off + co, 0, 0);
        const uint32_t wv = __builtin_amdgcn_raw_buffer_load_b16(Yp.r, cr_off + co, 0, 0);
        cu[k][0] = wu & 0xffu; cu[k][1] = wu >> 8; cv[k][0] = wv & 0xffu; cv[k][1] = wv >> 8;
      }
    } else {
      const int gy = clampi(y0 - 2 + ly, 0, g.h - 1);
      const uint32_t rowo = (uint32_t)(gy * g.w);
      const uint32_t crow = (uint32_t)((gy >> 1) * g.cw);
#pragma unroll
      for (int m = 0; m < 4; m++) yv[k][m] = Yp.ld(rowo + (uint32_t)clampi(x0 - 4 + lx + m, 0, g.w - 1));
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const uint32_t co = crow + (uint32_t)(clampi(x0 - 4 + lx + 2 * c, 0, g.w - 1) >> 1);
        cu[k][c] = Yp.ld(cb_off + co);
        cv[k][c] = Yp.ld(cr_off + co);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kFillSteps; k++) {
    const int i = threadIdx.x + 256 * k;
    if (kFillItems % 256 != 0 && k == kFillSteps - 1 && i >= kFillItems) break;
    const int ly = i / kQW;
    const int lx = (i - ly * kQW) * 4;
    const int32_t sc0 = w1 * (int32_t)(cu[k][0] * cu[k][0]) + w2 * (int32_t)(cv[k][0] * cv[k][0]);
    const int32_t sc1 = w1 * (int32_t)(cu[k][1] * cu[k][1]) + w2 * (int32_t)(cv[k][1] * cv[k][1]);
    const int32_t h = 5 * ly + lx;
    uint4 r01, r23;
    r01.x = yv[k][0]; r01.y = (uint32_t)(32 * (w0 * (int32_t)(yv[k][0] * yv[k][0]) + sc0) + h);
    r01.z = yv[k][1]; r01.w = (uint32_t)(32 * (w0 * (int32_t)(yv[k][1] * yv[k][1]) + sc0) + h + 1);
    r23.x = yv[k][2]; r23.y = (uint32_t)(32 * (w0 * (int32_t)(yv[k][2] * yv[k][2]) + sc1) + h + 2);
    r23.z = yv[k][3]; r23.w = (uint32_t)(32 * (w0 * (int32_t)(yv[k][3] * yv[k][3]) + sc1) + h + 3);
    uint4* dst = reinterpret_cast<uint4*>(s_luma + ly * kLW + lx);
    dst[0] = r01;
    dst[1] = r23;
    if ((ly & 1) == 0)
      *reinterpret_cast<uint2*>(s_cell + (ly >> 1) * kCW + (lx >> 1)) =
          make_uint2(cu[k][0] | (cv[k][0] << 16), cu[k][1] | (cv[k][1] << 16));
  }
}

// One direction over one tile. acc[r][c]: the lane's per-row, per-component sums (both of its
// columns) kept across the tile sweep.
template <typename T, int DIR>
__device__ __forceinline__ void tile_dir420(const CpsArgs& a, const Geom& g, int x0, int y0, int lane, int wave,
                                            const int32_t (&gcd)[3], uint2* s_luma, uint32_t* s_cell,
                                            uint32_t (&acc)[2 * kKB][3]) {
  // Opaque copies: keeps per-row address arithmetic from being hoisted out of the tile sweep and
  // held in registers across it.
  asm volatile("" : "+v"(lane));
  asm volatile("" : "+v"(wave));
  // R2T (DIR 0): P0 + gcd searched in P1; T2R (DIR 1): P1 - gcd searched in P0.
  const Buf<T> Xp(DIR == 0 ? a.p0 : a.p1, g.frame_bytes);
  const Buf<T> Yp(DIR == 0 ? a.p1 : a.p0, g.frame_bytes);
  uint16_t* scp_ptr = DIR == 0 ? a.scp1 : a.scp0;
  int16_t* best_out = DIR == 0 ? a.best0 : a.best1;
  constexpr int32_t sgn = DIR == 0 ? 1 : -1;
  const int32_t w0 = a.ws[0], w1 = a.ws[1], w2 = a.ws[2];
  const uint32_t cb_off = (uint32_t)g.cb_off, cr_off = (uint32_t)g.cr_off;

  __syncthreads();
  fill420<T>(Yp, g, x0, y0, w0, w1, w2, s_luma, s_cell);
  __syncthreads();

  // Tested values X = pel + sgn * gcd (not clipped: xc_CLIP_CURR_TST_RANGE = false,
  // libs/pmbb-ivqm/src/xCommonDefIVQM.h:36) and the key coefficients a_c = -64 w_c X_c.
  const int32_t m0 = -64 * w0, m1 = -64 * w1, m2 = -64 * w2;
  const int bx = min(x0 + 2 * lane, g.w - 2);  // block origin, clamped for the tile's tail
  int32_t x0v[kKB][2][2], a0[kKB][2][2], x1v[kKB], x2v[kKB], a1[kKB], a2[kKB];
#pragma unroll
  for (int b = 0; b < kKB; b++) {
    const int by = min(y0 + 2 * (kKB * wave + b), g.h - 2);
    const uint32_t co = (uint32_t)((by >> 1) * g.cw + (bx >> 1));
    x1v[b] = (int32_t)Xp.ld(cb_off + co) + sgn * gcd[1];
    x2v[b] = (int32_t)Xp.ld(cr_off + co) + sgn * gcd[2];
    a1[b] = __mul24(m1, x1v[b]);
    a2[b] = __mul24(m2, x2v[b]);
#pragma unroll
    for (int py = 0; py < 2; py++)
#pragma unroll
      for (int px = 0; px < 2; px++) {
        x0v[b][py][px] = (int32_t)Xp.ld((uint32_t)((by + py) * g.w + bx + px)) + sgn * gcd[0];
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

  // Winning candidates: per-component squared errors (and the shift-compensated pels).
#pragma unroll
  for (int b = 0; b < kKB; b++)
#pragma unroll
    for (int py = 0; py < 2; py++)
#pragma unroll
      for (int px = 0; px < 2; px++) {
        const int ly = 2 + 2 * (kKB * wave + b) + py;
        const int lx = 4 + 2 * lane + px;
        const int idx = (best[b][py][px] - (5 * ly + lx) + 12) & 31;
        const int dy = idx / 5;
        const int dx = idx - dy * 5;
        const int qy = ly - 2 + dy, qx = lx - 2 + dx;
        const uint32_t yv = s_luma[qy * kLW + qx].x;
        const uint32_t uv = s_cell[(qy >> 1) * kCW + (qx >> 1)];
        const int32_t d0 = x0v[b][py][px] - (int32_t)yv;
        const int32_t d1 = x1v[b] - (int32_t)(uv & 0xffffu);
        const int32_t d2 = x2v[b] - (int32_t)(uv >> 16);
        const int y = y0 + 2 * (kKB * wave + b) + py;
        const int x = x0 + 2 * lane + px;
        if (x < g.w && y < g.h) {
          acc[2 * b + py][0] += (uint32_t)(d0 * d0);
          acc[2 * b + py][1] += (uint32_t)(d1 * d1);
          acc[2 * b + py][2] += (uint32_t)(d2 * d2);
          if (best_out) best_out[y * g.w + x] = (int16_t)idx;
          if (scp_ptr) {
            const Buf<uint16_t> scp(scp_ptr, 6u * (uint32_t)g.w * (uint32_t)g.h);
            const uint32_t o = (uint32_t)(y * g.w + x);
            const uint32_t plane = (uint32_t)(g.w * g.h);
            scp.st(o, (uint32_t)clampi((int32_t)yv - sgn * gcd[0], 0, a.maxv));
            scp.st(plane + o, (uint32_t)clampi((int32_t)(uv & 0xffffu) - sgn * gcd[1], 0, a.maxv));
            scp.st(2 * plane + o, (uint32_t)clampi((int32_t)(uv >> 16) - sgn * gcd[2], 0, a.maxv));
          }
        }
      }
}

template <typename T>
__global__ __launch_bounds__(256) void k_cps_fast420(CpsArgs a) {
  __shared__ __attribute__((aligned(16))) uint2 s_luma[kLH * kLW];
  __shared__ __attribute__((aligned(16))) uint32_t s_cell[kCH * kCW];

  const FrameSummary* __restrict__ sum = a.sum;
  if (!sum->in_range) return;
  const Geom g = a.g;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int band = blockIdx.x / a.nseg;
  const int seg = blockIdx.x - band * a.nseg;
  const int y0 = band * kTH;
  const int tx_begin = seg * a.sweep;
  const int tx_end = min(a.tiles_x, tx_begin + a.sweep);
  const int32_t gcd[3] = {sum->gcd[0], sum->gcd[1], sum->gcd[2]};

  // Per-lane row sums over the sweep (the host bounds sweep * 2 * max squared error below 2^32).
  uint32_t acc0[2 * kKB][3], acc1[2 * kKB][3];
#pragma unroll
  for (int r = 0; r < 2 * kKB; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc0[r][c] = acc1[r][c] = 0;

  for (int tx = tx_begin; tx < tx_end; tx++) {
    const int x0 = tx * kTW;
    tile_dir420<T, 0>(a, g, x0, y0, lane, wave, gcd, s_luma, s_cell, acc0);
    tile_dir420<T, 1>(a, g, x0, y0, lane, wave, gcd, s_luma, s_cell, acc1);
  }

  // One partial per (segment, direction, component, row): reduce over the wave's 64 lanes.
#pragma unroll
  for (int r = 0; r < 2 * kKB; r++) {
    const int y = y0 + 2 * kKB * wave + r;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint64_t v0 = wave_sum((uint64_t)acc0[r][c]);
      const uint64_t v1 = wave_sum((uint64_t)acc1[r][c]);
      if (lane == 0 && y < g.h) {
        a.partials[(((int64_t)seg * 2 + 0) * 3 + c) * g.h + y] = v0;
        a.partials[(((int64_t)seg * 2 + 1) * 3 + c) * g.h + y] = v1;
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
  if (a.g.bytes_per_sample == 1)
    hipLaunchKernelGGL(k_cps_fast420<uint8_t>, grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL(k_cps_fast420<uint16_t>, grid, block, 0, stream, a);
  return hipGetLastError();
}

void cps_fast420_tile(int32_t* tile_w, int32_t* tile_h) {
  *tile_w = kTW;
  *tile_h = kTH;
}

int cps_fast420_blocks_per_cu(int bytes_per_sample) {
  int n = 0;
  hipError_t e = bytes_per_sample == 1
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps_fast420<uint8_t>, 256, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps_fast420<uint16_t>, 256, 0);
  return e == hipSuccess && n > 0 ? n : 1;
}

}  // namespace qmiv
