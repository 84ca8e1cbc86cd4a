// gfx950 kernels of the PSNR family: per-row statistics, the on-device GCD summary and the
// corresponding-pixel-shift (CPS) search behind IV-PSNR and the shift-compensated pictures.
//
// Frames are read in their packed planar file layout; 4:2:0 chroma is addressed at
// (y >> 1, x >> 1), which is exactly the nearest-neighbour 4:2:0 -> 4:4:4 upsampling of
// libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49, and picture edges are addressed by clamping,
// which reads the same values as the reference's margin replication
// (libs/pmbb-core/src/xPixelOpsSTD.cpp:210-236).
#include "qmiv_internal.h"
#include "device_common.h"

#include <algorithm>
#include <type_traits>
#include <climits>

namespace qmiv {
namespace {

template <typename T>
__device__ __forceinline__ uint32_t ldpel(const T* __restrict__ p, int64_t i) {
  return (uint32_t)p[i];
}

// ---------------------------------------------------------------------------------------------
// Per-row statistics: sum(P1 - P0), sum((P1 - P0)^2) and the largest pel of each picture.
// One wave per row of any plane; 16-byte loads when the row is aligned.
// Restates xDistortion::CalcSD / CalcSSD (libs/pmbb-core/src/xDistortionSTD.cpp:18-62) on the
// 4:2:0 planes: the 4:4:4 sums are 4x the chroma-plane sums (each chroma sample appears 4 times).
// ---------------------------------------------------------------------------------------------
// gate (nullable): the summary of a first pass; when it found both pictures in range this pass has
// nothing to redo (the concealment flow re-runs the statistics only after clipping).
// zero / nzero: an array the pass clears in any case (the CPS kernel's per-row sums, which it adds
// into after this pass).
//
// 16-bit samples of at most 12-bit depth take packed 16-bit arithmetic: per two samples one
// v_pk_sub_i16 difference, one v_dot2 for the two squares, one v_dot2 against (1, 1) for the two
// differences and one v_pk_max_u16 per picture. Exact while every sample is below 2^12 (differences
// below 2^12 in magnitude, squares below 2^24, one group's eight dot products below 2^28); a row
// holding a larger sample (a pel above the bit depth) is summed again by the scalar loop.
typedef short qmiv_s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short qmiv_u16x2 __attribute__((ext_vector_type(2)));

template <typename T>
__global__ __launch_bounds__(256) void k_stats(const T* __restrict__ p0, const T* __restrict__ p1, Geom g,
                                               int64_t* __restrict__ row_sd, uint64_t* __restrict__ row_ssd,
                                               uint32_t* __restrict__ row_max, const FrameSummary* __restrict__ gate,
                                               uint64_t* __restrict__ zero, int32_t nzero) {
  {
    const int32_t zi = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (zi < nzero) zero[zi] = 0;
  }
  if (gate && gate->in_range) return;
  const int lane = threadIdx.x & 63;
  const int nrows = g.h + 2 * g.ch;
  // The row is the wave's: scalar registers hold its address.
  const int r = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (r >= nrows) return;
  int64_t off;
  int32_t width;
  if (r < g.h) {
    off = (int64_t)r * g.w;
    width = g.w;
  } else if (r < g.h + g.ch) {
    off = g.cb_off + (int64_t)(r - g.h) * g.cw;
    width = g.cw;
  } else {
    off = g.cr_off + (int64_t)(r - g.h - g.ch) * g.cw;
    width = g.cw;
  }
  const T* a = p0 + off;
  const T* b = p1 + off;
  int64_t sd = 0;
  uint64_t ssd = 0;
  uint32_t m0 = 0, m1 = 0;
  constexpr int kVec = 16 / sizeof(T);  // samples per 16-byte load
  const bool aligned = (((uintptr_t)a | (uintptr_t)b) & 15) == 0;
  bool scalar = true;
  if constexpr (sizeof(T) == 2) {
    if (g.bd <= 12 && aligned) {
      const int32_t vec_end = (width / kVec) * kVec;
      qmiv_u16x2 mx0 = {0, 0}, mx1 = {0, 0};
      // Two groups per picture in flight per lane.
      int32_t x = lane * kVec;
      auto group = [&](const uint4& va, const uint4& vb, int32_t& sq, int32_t& df) {
        const uint32_t wa[4] = {va.x, va.y, va.z, va.w};
        const uint32_t wb[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const qmiv_u16x2 ua = __builtin_bit_cast(qmiv_u16x2, wa[k]);
          const qmiv_u16x2 ub = __builtin_bit_cast(qmiv_u16x2, wb[k]);
          const qmiv_s16x2 d = __builtin_bit_cast(qmiv_s16x2, wb[k]) - __builtin_bit_cast(qmiv_s16x2, wa[k]);
          sq = __builtin_amdgcn_sdot2(d, d, sq, false);
          df = __builtin_amdgcn_sdot2(d, (qmiv_s16x2){1, 1}, df, false);
          mx0 = __builtin_elementwise_max(mx0, ua);
          mx1 = __builtin_elementwise_max(mx1, ub);
        }
      };
      for (; x + 64 * kVec < vec_end; x += 2 * 64 * kVec) {
        const uint4 va0 = *reinterpret_cast<const uint4*>(a + x), vb0 = *reinterpret_cast<const uint4*>(b + x);
        const uint4 va1 = *reinterpret_cast<const uint4*>(a + x + 64 * kVec);
        const uint4 vb1 = *reinterpret_cast<const uint4*>(b + x + 64 * kVec);
        int32_t sq = 0, df = 0;
        group(va0, vb0, sq, df);
        group(va1, vb1, sq, df);
        ssd += (uint32_t)sq;
        sd += df;
      }
      if (x < vec_end) {
        int32_t sq = 0, df = 0;
        group(*reinterpret_cast<const uint4*>(a + x), *reinterpret_cast<const uint4*>(b + x), sq, df);
        ssd += (uint32_t)sq;
        sd += df;
      }
      for (int32_t t = vec_end + lane; t < width; t += 64) {
        const uint32_t ea = ldpel(a, t), eb = ldpel(b, t);
        const int32_t d = (int32_t)eb - (int32_t)ea;
        sd += d;
        ssd += (uint32_t)(d * d);
        m0 = max(m0, ea);
        m1 = max(m1, eb);
      }
      m0 = max(m0, (uint32_t)max(mx0.x, mx0.y));
      m1 = max(m1, (uint32_t)max(mx1.x, mx1.y));
      m0 = wave_max(m0);
      m1 = wave_max(m1);
      scalar = max(m0, m1) > 4095u;  // a sample outside the packed path's range: sum the row again
      if (scalar) {
        sd = 0;
        ssd = 0;
      }
    }
  }
  if (scalar) {
    int32_t vec_end = 0;
    // Adds one 16-byte group of both pictures to the lane's sums.
    auto group = [&](const uint4& va, const uint4& vb) {
      const uint32_t wa[4] = {va.x, va.y, va.z, va.w};
      const uint32_t wb[4] = {vb.x, vb.y, vb.z, vb.w};
      int32_t sd_l = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int s = 0; s < (int)(4 / sizeof(T)); s++) {
          const uint32_t sh = s * 8 * sizeof(T);
          const uint32_t mask = sizeof(T) == 1 ? 0xffu : 0xffffu;
          const uint32_t ea = (wa[k] >> sh) & mask;
          const uint32_t eb = (wb[k] >> sh) & mask;
          const int32_t d = (int32_t)eb - (int32_t)ea;
          sd_l += d;
          const uint32_t sq = (uint32_t)d * (uint32_t)d;
          ssd += sq;
          m0 = max(m0, ea);
          m1 = max(m1, eb);
        }
      }
      sd += sd_l;
    };
    if (aligned) {
      vec_end = (width / kVec) * kVec;
      // 8-bit samples: four groups per picture in flight per lane before any is used (HBM needs the
      // bytes in flight); 16-bit samples reach here only for rare rows and keep registers low.
      constexpr int U = sizeof(T) == 1 ? 4 : 1;
      int32_t x = lane * kVec;
      for (; x + (U - 1) * 64 * kVec < vec_end; x += U * 64 * kVec) {
        uint4 va[U], vb[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          va[u] = *reinterpret_cast<const uint4*>(a + x + u * 64 * kVec);
          vb[u] = *reinterpret_cast<const uint4*>(b + x + u * 64 * kVec);
        }
#pragma unroll
        for (int u = 0; u < U; u++) group(va[u], vb[u]);
      }
      for (; x < vec_end; x += 64 * kVec)
        group(*reinterpret_cast<const uint4*>(a + x), *reinterpret_cast<const uint4*>(b + x));
    }
    for (int32_t x = vec_end + lane; x < width; x += 64) {
      const uint32_t ea = ldpel(a, x), eb = ldpel(b, x);
      const int32_t d = (int32_t)eb - (int32_t)ea;
      sd += d;
      ssd += (uint32_t)d * (uint32_t)d;
      m0 = max(m0, ea);
      m1 = max(m1, eb);
    }
    m0 = wave_max(m0);
    m1 = wave_max(m1);
  }
  sd = wave_sum(sd);
  ssd = wave_sum(ssd);
  if (lane == 0) {
    row_sd[r] = sd;
    row_ssd[r] = ssd;
    row_max[r] = m0;
    row_max[nrows + r] = m1;
  }
}

// ---------------------------------------------------------------------------------------------
// Frame summary: plane totals, range flags and the global colour difference.
// GCD restates xGlobClrDiff::CalcGlobalColorDiff (libs/pmbb-ivqm/src/xGlobClrDiff.cpp:15-48) as
// the processor calls it (xGlobClrDiff.h:41): the 4:4:4 sum(P1 - P0) wraps in int32
// (xDistortionSTD.cpp:18-28), is divided by the area in double precision and rounded half-even
// (the SSE conversion of xCommonDefCORE.h:129-131), then clipped to +-maxdiff.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_summary(Geom g, const int64_t* __restrict__ row_sd,
                                                 const uint64_t* __restrict__ row_ssd,
                                                 const uint32_t* __restrict__ row_max, int32_t md0, int32_t md1,
                                                 int32_t md2, int32_t md3, int32_t compute_gcd,
                                                 FrameSummary* __restrict__ sum,
                                                 const FrameSummary* __restrict__ gate) {
  // Concealment flow: a first pass that found both pictures in range is this pass's result.
  if (gate && gate->in_range) {
    if (threadIdx.x == 0) *sum = *gate;
    return;
  }
  __shared__ unsigned long long s_sd[3], s_ssd[3];
  __shared__ unsigned int s_max[2];
  if (threadIdx.x < 3) {
    s_sd[threadIdx.x] = 0;
    s_ssd[threadIdx.x] = 0;
  }
  if (threadIdx.x < 2) s_max[threadIdx.x] = 0;
  __syncthreads();
  const int nrows = g.h + 2 * g.ch;
  int64_t sd[3] = {0, 0, 0};
  uint64_t ssd[3] = {0, 0, 0};
  uint32_t m0 = 0, m1 = 0;
  // Raw buffer loads, the offset of a row past nrows replaced by one past every array (frames are
  // below 2 GiB, so 2^31 is): such loads return 0 and need no branch, so four rows' loads per lane
  // issue together.
  const uint32_t nr = (uint32_t)nrows;
  const __amdgpu_buffer_rsrc_t rsd = __builtin_amdgcn_make_buffer_rsrc(const_cast<int64_t*>(row_sd), 0, (int)(8 * nr), kBufDword3);
  const __amdgpu_buffer_rsrc_t rss = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(row_ssd), 0, (int)(8 * nr), kBufDword3);
  const __amdgpu_buffer_rsrc_t rmx = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(row_max), 0, (int)(8 * nr), kBufDword3);
  for (uint32_t r0 = 0; r0 < nr; r0 += 4 * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t r = r0 + (uint32_t)k * blockDim.x + threadIdx.x;
      const bool ok = r < nr;
      const uint32_t o8 = ok ? 8u * r : 0x80000000u, o4 = ok ? 4u * r : 0x80000000u;
      const int c = (int)r < g.h ? 0 : ((int)r < g.h + g.ch ? 1 : 2);
      const int64_t vsd = __builtin_bit_cast(int64_t, __builtin_amdgcn_raw_buffer_load_b64(rsd, o8, 0, 0));
      const uint64_t vss = __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rss, o8, 0, 0));
      const uint32_t v0 = __builtin_amdgcn_raw_buffer_load_b32(rmx, o4, 0, 0);
      const uint32_t v1 = __builtin_amdgcn_raw_buffer_load_b32(rmx, ok ? 4u * (nr + r) : 0x80000000u, 0, 0);
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        sd[cc] += cc == c ? vsd : 0;
        ssd[cc] += cc == c ? vss : 0;
      }
      m0 = max(m0, v0);
      m1 = max(m1, v1);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    sd[c] = wave_sum(sd[c]);
    ssd[c] = wave_sum(ssd[c]);
  }
  m0 = wave_max(m0);
  m1 = wave_max(m1);
  if ((threadIdx.x & 63) == 0) {
    for (int c = 0; c < 3; c++) {
      atomicAdd(&s_sd[c], (unsigned long long)sd[c]);
      atomicAdd(&s_ssd[c], (unsigned long long)ssd[c]);
    }
    atomicMax(&s_max[0], m0);
    atomicMax(&s_max[1], m1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t mult = g.cs ? 4 : 1;
    const int32_t maxv = (1 << g.bd) - 1;
    const int32_t md[4] = {md0, md1, md2, md3};
    int64_t sd444[3];
    for (int c = 0; c < 3; c++) {
      sd444[c] = (int64_t)s_sd[c] * (c ? mult : 1);
      sum->sd[c] = sd444[c];
      sum->ssd[c] = (uint64_t)s_ssd[c] * (uint64_t)(c ? mult : 1);
    }
    sum->max_pel[0] = (int32_t)s_max[0];
    sum->max_pel[1] = (int32_t)s_max[1];
    sum->in_range = g.bd >= 16 ? 1 : ((int32_t)s_max[0] <= maxv && (int32_t)s_max[1] <= maxv);
    if (compute_gcd) {
      const double area = (double)(g.w * g.h);
      for (int c = 0; c < 4; c++) {
        const int32_t sd32 = c < 3 ? (int32_t)(uint32_t)(uint64_t)sd444[c] : 0;
        const double avg = (double)sd32 / area;
        const int32_t rnd = (int32_t)__builtin_rint(avg);
        sum->gcd[c] = max(-md[c], min(rnd, md[c]));
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Fast CPS search, B = 2 (the reference default), for parameter sets whose terms fit the 24-bit
// multiplier operands (checked on the host).
//
// For a pel p with value X (= tested pel + GCD) and candidate q of the searched picture Y:
//   E(p, q) = sum_c w_c (X_c - Y_c(q))^2 = S(q) - 2 <w X, Y(q)> + sum_c w_c X_c^2.
// The last term does not depend on q, so ordering the window by E equals ordering it by
//   t(p, q) = S(q) - 2 <w X, Y(q)>,   S(q) = sum_c w_c Y_c(q)^2.
// Each LDS record holds C(q) = 32 S(q) + h(q), with h(q) = 5 ly + lx its local raster position,
// and each pel holds a_c = -64 w_c X_c, so three v_mad_i32_i24 give
//   key(p, q) = 32 t(p, q) + h(q).
// Within one 5x5 window h(q) - h(p) = 5 dy + dx spans [-12, 12], so the signed minimum of the
// keys is the candidate with the smallest error and, among equal errors, the first in raster
// order: the strict '<' scan of xCorrespPixelShiftSSE::xCalcDistWithinBlock
// (libs/pmbb-ivqm/src/xCorrespPixelShiftSSE.cpp:40-71). The winner's offset is recovered as
// (key - h(p) + 12) mod 32.
//
// Workgroup: 4 waves, tile 64 x kCpsTileH pels (16 by default), each lane owns one column x
// kCpsTileH / 4 rows. The searched picture's haloed records sit in LDS (72 x (kCpsTileH + 4)
// records of 16 bytes), filled one tile row ahead of the search. A workgroup sweeps
// `sweep` tiles along x; per-row per-component sums are reduced over the wave after every tile
// and accumulated in LDS, then written as one partial per (segment, direction, component, row).
// ---------------------------------------------------------------------------------------------

// LDS tile of the fast kernel: rows [y0 - 2, y0 + kCpsTileH + 2), columns [x0 - 4, x0 + 68). The two extra
// columns on each side keep every row start 8-byte aligned for the vectorised fill; the search
// reads columns [x0 - 2, x0 + 66).
constexpr int kCpsLW = kCpsTileW + 8;
constexpr int kCpsLH = kCpsTileH + 4;

// Masked search: a searched pel whose mask is 0 gets the record {0, 0, 0, kCpsMaskedC + (h mod 32)}.
// Its key is then above 2^30 while every real key lies in (-2^29, 2^29) (the host checks the
// bound), so it never beats an unmasked candidate, and a winner still decodes to a window
// position. The tested pel is always its own unmasked candidate, as in
// xCorrespPixelShiftSTD::FindBestPixelWithinBlockM (libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp).
constexpr uint32_t kCpsMaskedC = 1u << 30;
__device__ __forceinline__ uint4 masked_rec(uint4 r, uint32_t m) {
  return m != 0u ? r : make_uint4(0u, 0u, 0u, kCpsMaskedC | (r.w & 31u));
}

// Fills the LDS records of picture Yp for the tile at (x0, y0):
//   rec = {Y, U, V, C},  C = 32 * (w0 Y^2 + w1 U^2 + w2 V^2) + 5 ly + lx;
// with MASKED, the masked record wherever the mask Mp (its luma plane) is 0.
template <typename T, typename TM, bool MASKED>
__device__ __forceinline__ void cps_fill(const Buf<T>& Yp, const Buf<TM>& Mp, const Geom& g, int x0, int y0, int32_t w0,
                                         int32_t w1, int32_t w2, uint4* s_rec) {
  constexpr int LW = kCpsLW, LH = kCpsLH;
  const uint32_t cb_off = (uint32_t)g.cb_off, cr_off = (uint32_t)g.cr_off;
  const bool interior = x0 >= 4 && x0 + kCpsTileW + 4 <= g.w && y0 >= 2 && y0 + kCpsTileH + 2 <= g.h && (g.w & 3) == 0;
  if (interior) {
    // Four records per thread and step: one aligned 8-byte (u16) or 4-byte (u8) luma load; chroma
    // either two samples (4:2:0, each shared by two records) or four (4:4:4).
    constexpr int QW = LW / 4;
    for (int i = threadIdx.x; i < LH * QW; i += 256) {
      const int ly = i / QW;
      const int lx = (i - ly * QW) * 4;
      const int gy = y0 - 2 + ly;
      const int gx = x0 - 4 + lx;
      uint32_t yv[4], uv[4], vv[4];
      const uint32_t lo = (uint32_t)(gy * g.w + gx);
      if constexpr (sizeof(T) == 2) {
        const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(Yp.r, lo * 2u, 0, 0));
        yv[0] = w.x & 0xffffu; yv[1] = w.x >> 16; yv[2] = w.y & 0xffffu; yv[3] = w.y >> 16;
      } else {
        const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, lo, 0, 0);
        yv[0] = w & 0xffu; yv[1] = (w >> 8) & 0xffu; yv[2] = (w >> 16) & 0xffu; yv[3] = w >> 24;
      }
      if (g.cs) {
        const uint32_t co = (uint32_t)((gy >> 1) * g.cw + (gx >> 1));
        uint32_t u0, u1, v0, v1;
        if constexpr (sizeof(T) == 2) {
          const uint32_t wu = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, (cb_off + co) * 2u, 0, 0);
          const uint32_t wv = __builtin_amdgcn_raw_buffer_load_b32(Yp.r, (cr_off + co) * 2u, 0, 0);
          u0 = wu & 0xffffu; u1 = wu >> 16; v0 = wv & 0xffffu; v1 = wv >> 16;
        } else {
          const uint32_t wu = __builtin_amdgcn_raw_buffer_load_b16(Yp.r, cb_off + co, 0, 0);
          const uint32_t wv = __builtin_amdgcn_raw_buffer_load_b16(Yp.r, cr_off + co, 0, 0);
          u0 = wu & 0xffu; u1 = wu >> 8; v0 = wv & 0xffu; v1 = wv >> 8;
        }
        uv[0] = uv[1] = u0; uv[2] = uv[3] = u1;
        vv[0] = vv[1] = v0; vv[2] = vv[3] = v1;
      } else {
#pragma unroll
        for (int m = 0; m < 4; m++) {
          uv[m] = Yp.ld(cb_off + lo + m);
          vv[m] = Yp.ld(cr_off + lo + m);
        }
      }
      uint32_t mk[4] = {1u, 1u, 1u, 1u};
      if constexpr (MASKED) {
        if constexpr (sizeof(TM) == 2) {
          const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(Mp.r, lo * 2u, 0, 0));
          mk[0] = w.x & 0xffffu; mk[1] = w.x >> 16; mk[2] = w.y & 0xffffu; mk[3] = w.y >> 16;
        } else {
          const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(Mp.r, lo, 0, 0);
          mk[0] = w & 0xffu; mk[1] = (w >> 8) & 0xffu; mk[2] = (w >> 16) & 0xffu; mk[3] = w >> 24;
        }
      }
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const int32_t S = w0 * (int32_t)(yv[m] * yv[m]) + w1 * (int32_t)(uv[m] * uv[m]) + w2 * (int32_t)(vv[m] * vv[m]);
        const uint4 r = make_uint4(yv[m], uv[m], vv[m], (uint32_t)(S * 32 + 5 * ly + lx + m));
        s_rec[ly * LW + lx + m] = MASKED ? masked_rec(r, mk[m]) : r;
      }
    }
  } else {
    for (int i = threadIdx.x; i < LH * LW; i += 256) {
      const int ly = i / LW;
      const int lx = i - ly * LW;
      const int gy = clampi(y0 - 2 + ly, 0, g.h - 1);
      const int gx = clampi(x0 - 4 + lx, 0, g.w - 1);
      const uint32_t yv = Yp.ld((uint32_t)(gy * g.w + gx));
      const uint32_t co = (uint32_t)((gy >> g.cs) * g.cw + (gx >> g.cs));
      const uint32_t uv = Yp.ld(cb_off + co);
      const uint32_t vv = Yp.ld(cr_off + co);
      const int32_t S = w0 * (int32_t)(yv * yv) + w1 * (int32_t)(uv * uv) + w2 * (int32_t)(vv * vv);
      const uint4 r = make_uint4(yv, uv, vv, (uint32_t)(S * 32 + 5 * ly + lx));
      s_rec[i] = MASKED ? masked_rec(r, Mp.ld((uint32_t)(gy * g.w + gx))) : r;
    }
  }
}

// One direction over one tile: each of the lane's kCpsTileH / 4 pels finds its best candidate; the winner's
// per-component squared error is added to acc (per-lane row sums kept across the tile sweep).
template <bool MASKED>
using CpsAcc = typename std::conditional<MASKED, uint64_t, uint32_t>::type;

template <typename T, typename TM, bool MASKED, int DIR>
__device__ __forceinline__ void cps_tile_dir(const CpsArgs& a, const Geom& g, int x0, int y0, int lane, int wave,
                                             const int32_t (&gcd)[3], uint4* s_rec,
                                             CpsAcc<MASKED> (&acc)[kCpsTileH / 4][3]) {
  // Opaque copies: keeps per-row address arithmetic from being hoisted out of the tile sweep and
  // held in registers across it.
  asm volatile("" : "+v"(lane));
  asm volatile("" : "+v"(wave));
  constexpr int LW = kCpsLW;
  // R2T (DIR 0): P0 + gcd searched in P1; T2R (DIR 1): P1 - gcd searched in P0.
  const Buf<T> Xp(DIR == 0 ? a.p0 : a.p1, g.frame_bytes);
  const Buf<T> Yp(DIR == 0 ? a.p1 : a.p0, g.frame_bytes);
  uint16_t* scp_ptr = DIR == 0 ? a.scp1 : a.scp0;
  int16_t* best_out = DIR == 0 ? a.best0 : a.best1;
  constexpr int32_t sgn = DIR == 0 ? 1 : -1;
  const int32_t w0 = a.ws[0], w1 = a.ws[1], w2 = a.ws[2];
  const uint32_t cb_off = (uint32_t)g.cb_off, cr_off = (uint32_t)g.cr_off;

  const Buf<TM> Mp(a.mask, a.mask_bytes);
  __syncthreads();
  cps_fill<T, TM, MASKED>(Yp, Mp, g, x0, y0, w0, w1, w2, s_rec);
  __syncthreads();

  // The lane's pels: X = pel + sgn * gcd (not clipped: xc_CLIP_CURR_TST_RANGE = false,
  // libs/pmbb-ivqm/src/xCommonDefIVQM.h:36), kept packed for the winner's error, and the scaled
  // coefficients a_c = -64 w_c X_c of the key.
  const int x = x0 + lane;
  const int xc = min(x, g.w - 1);
  const int32_t m0 = -64 * w0, m1 = -64 * w1, m2 = -64 * w2;
  constexpr int K = kCpsTileH / 4;  // rows per lane
  int32_t a0[K], a1[K], a2[K];
  uint32_t x01[K];
  int32_t x2[K];
  uint32_t mx[K];  // masked search: the tested pels' mask values
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int y = min(y0 + K * wave + k, g.h - 1);
    if constexpr (MASKED) mx[k] = Mp.ld((uint32_t)(y * g.w + xc));
    const uint32_t co = (uint32_t)((y >> g.cs) * g.cw + (xc >> g.cs));
    const int32_t v0 = (int32_t)Xp.ld((uint32_t)(y * g.w + xc)) + sgn * gcd[0];
    const int32_t v1 = (int32_t)Xp.ld(cb_off + co) + sgn * gcd[1];
    const int32_t v2 = (int32_t)Xp.ld(cr_off + co) + sgn * gcd[2];
    a0[k] = __mul24(m0, v0);
    a1[k] = __mul24(m1, v1);
    a2[k] = __mul24(m2, v2);
    x01[k] = ((uint32_t)v0 & 0xffffu) | ((uint32_t)v1 << 16);
    x2[k] = v2;
  }

  int32_t best[K];
#pragma unroll
  for (int k = 0; k < K; k++) best[k] = INT_MAX;
  // Window rows are read one ahead of use: the next row's records load while this row's keys are
  // computed.
  uint4 qn[5];
#pragma unroll
  for (int m = 0; m < 5; m++) qn[m] = s_rec[(K * wave) * LW + lane + 2 + m];
#pragma unroll
  for (int rr = 0; rr < K + 4; rr++) {
    uint4 q[5];
#pragma unroll
    for (int m = 0; m < 5; m++) q[m] = qn[m];
    if (rr + 1 < K + 4) {
#pragma unroll
      for (int m = 0; m < 5; m++) qn[m] = s_rec[(K * wave + rr + 1) * LW + lane + 2 + m];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int dy = rr - 2 - k;
      if (dy < -2 || dy > 2) continue;
      int32_t key[5];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        int32_t t = mad24(a0[k], (int32_t)q[m].x, (int32_t)q[m].w);
        t = mad24(a1[k], (int32_t)q[m].y, t);
        key[m] = mad24(a2[k], (int32_t)q[m].z, t);
      }
      best[k] = min(min(best[k], key[0]), key[1]);
      best[k] = min(min(best[k], key[2]), key[3]);
      best[k] = min(best[k], key[4]);
      // Materialise the running minimum now: otherwise the min chain is re-associated across
      // window rows and the keys wait in scratch.
      asm volatile("" : "+v"(best[k]));
    }
    // One window row at a time: bounds the live keys and records.
    __builtin_amdgcn_sched_barrier(0);
  }

  // Winning candidate: per-component squared error of X against it (and the SCP pel).
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int ly = K * wave + k + 2;
    const int hp = 5 * ly + lane + 4;
    const int idx = (best[k] - hp + 12) & 31;
    const int dy = idx / 5;
    const int dx = idx - dy * 5;
    const uint4 r = s_rec[(ly - 2 + dy) * LW + lane + 2 + dx];
    const int32_t xv0 = (int32_t)(int16_t)(x01[k] & 0xffffu);
    const int32_t xv1 = (int32_t)(int16_t)(x01[k] >> 16);
    const int32_t d0 = xv0 - (int32_t)r.x, d1 = xv1 - (int32_t)r.y, d2 = x2[k] - (int32_t)r.z;
    const int y = y0 + K * wave + k;
    if (x < g.w && y < g.h) {
      if constexpr (MASKED) {
        // xCorrespPixelShiftSTD::CalcDistAsymmetricRowM: the winner's int32 squared errors times the
        // tested pel's mask value, in 64 bits (the host keeps the squares below 2^31).
        acc[k][0] += (uint64_t)(uint32_t)(d0 * d0) * mx[k];
        acc[k][1] += (uint64_t)(uint32_t)(d1 * d1) * mx[k];
        acc[k][2] += (uint64_t)(uint32_t)(d2 * d2) * mx[k];
      } else {
        acc[k][0] += (uint32_t)(d0 * d0);
        acc[k][1] += (uint32_t)(d1 * d1);
        acc[k][2] += (uint32_t)(d2 * d2);
      }
      if (best_out) best_out[y * g.w + x] = (int16_t)idx;
      if (scp_ptr) {
        const Buf<uint16_t> scp(scp_ptr, 6u * (uint32_t)g.w * (uint32_t)g.h);
        const uint32_t o = (uint32_t)(y * g.w + x);
        const uint32_t plane = (uint32_t)(g.w * g.h);
        scp.st(o, (uint32_t)clampi((int32_t)r.x - sgn * gcd[0], 0, a.maxv));
        scp.st(plane + o, (uint32_t)clampi((int32_t)r.y - sgn * gcd[1], 0, a.maxv));
        scp.st(2 * plane + o, (uint32_t)clampi((int32_t)r.z - sgn * gcd[2], 0, a.maxv));
      }
    }
  }
}

template <typename T, typename TM, bool MASKED>
__global__ __launch_bounds__(256) void k_cps_fast(CpsArgs a) {
  __shared__ uint4 s_rec[kCpsLH * kCpsLW];

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
  const int band = blockIdx.x / a.nseg;
  const int seg = blockIdx.x - band * a.nseg;
  const int y0 = band * kCpsTileH;
  const int tx_begin = seg * a.sweep;
  const int tx_end = min(a.tiles_x, tx_begin + a.sweep);
  const int32_t gcd[3] = {gsrc[0], gsrc[1], gsrc[2]};

  // Per-lane row sums over the sweep (the host bounds sweep * max squared error below 2^32;
  // masked sums are kept in 64 bits).
  constexpr int K = kCpsTileH / 4;
  CpsAcc<MASKED> acc0[K][3], acc1[K][3];
#pragma unroll
  for (int k = 0; k < K; k++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc0[k][c] = acc1[k][c] = 0;

  for (int tx = tx_begin; tx < tx_end; tx++) {
    const int x0 = tx * kCpsTileW;
    cps_tile_dir<T, TM, MASKED, 0>(a, g, x0, y0, lane, wave, gcd, s_rec, acc0);
    cps_tile_dir<T, TM, MASKED, 1>(a, g, x0, y0, lane, wave, gcd, s_rec, acc1);
  }

  // One partial per (segment, direction, component, row): reduce over the wave's 64 lanes.
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int y = y0 + K * wave + k;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint64_t v0 = wave_sum((uint64_t)acc0[k][c]);
      const uint64_t v1 = wave_sum((uint64_t)acc1[k][c]);
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

// ---------------------------------------------------------------------------------------------
// Generic CPS search: any search range, weights and bit depth. One thread per pel, the window
// read straight from memory, errors in wrapping 32-bit arithmetic over the four interleaved
// lanes exactly as xCorrespPixelShiftSSE::xCalcDistWithinBlock computes them (the fourth lane
// holds 0 in the picture and the GCD's fourth entry in the tested value).
// grid = (ceil(w / 256), h, 2 directions).
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_cps_generic(CpsArgs a) {
  __shared__ uint64_t red[3][256];
  const FrameSummary* __restrict__ sum = a.sum;
  if (!cps_gate_open(a.gate, sum->in_range)) return;
  const Geom g = a.g;
  const int y = blockIdx.y;
  const int dir = blockIdx.z;
  const int x = blockIdx.x * kGenericCols + threadIdx.x;
  const int32_t sgn = dir == 0 ? 1 : -1;
  const T* __restrict__ Xp = (const T*)(dir == 0 ? a.p0 : a.p1);
  const T* __restrict__ Yp = (const T*)(dir == 0 ? a.p1 : a.p0);
  uint16_t* __restrict__ scp = dir == 0 ? a.scp1 : a.scp0;
  int16_t* __restrict__ best_out = dir == 0 ? a.best0 : a.best1;
  uint32_t dist_best[3] = {0, 0, 0};
  if (x < g.w) {
    int32_t t[4];
    const int64_t co = (int64_t)(y >> g.cs) * g.cw + (x >> g.cs);
    t[0] = (int32_t)ldpel(Xp, (int64_t)y * g.w + x) + sgn * sum->gcd[0];
    t[1] = (int32_t)ldpel(Xp, g.cb_off + co) + sgn * sum->gcd[1];
    t[2] = (int32_t)ldpel(Xp, g.cr_off + co) + sgn * sum->gcd[2];
    t[3] = sgn * sum->gcd[3];
    int32_t best_err = INT_MAX;
    int64_t best_q = -1, best_co = -1;
    int32_t best_idx = 0, idx = 0;
    for (int dy = -a.B; dy <= a.B; dy++) {
      const int qy = clampi(y + dy, 0, g.h - 1);
      for (int dx = -a.B; dx <= a.B; dx++) {
        const int qx = clampi(x + dx, 0, g.w - 1);
        const int64_t q = (int64_t)qy * g.w + qx;
        const int64_t qc = (int64_t)(qy >> g.cs) * g.cw + (qx >> g.cs);
        const int32_t r[4] = {(int32_t)ldpel(Yp, q), (int32_t)ldpel(Yp, g.cb_off + qc),
                              (int32_t)ldpel(Yp, g.cr_off + qc), 0};
        uint32_t dist[4], err = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const uint32_t d = (uint32_t)(t[c] - r[c]);
          dist[c] = d * d;
          err += dist[c] * (uint32_t)a.ws[c];
        }
        if ((int32_t)err < best_err) {
          best_err = (int32_t)err;
          dist_best[0] = dist[0];
          dist_best[1] = dist[1];
          dist_best[2] = dist[2];
          best_q = q;
          best_co = qc;
          best_idx = idx;
        }
        idx++;
      }
    }
    if (best_out) best_out[(int64_t)y * g.w + x] = (int16_t)best_idx;
    if (scp && best_q >= 0) {
      const int64_t o = (int64_t)y * g.w + x;
      const int64_t plane = (int64_t)g.w * g.h;
      const int32_t sel[3] = {(int32_t)ldpel(Yp, best_q), (int32_t)ldpel(Yp, g.cb_off + best_co),
                              (int32_t)ldpel(Yp, g.cr_off + best_co)};
      for (int c = 0; c < 3; c++)
        scp[c * plane + o] = (uint16_t)clampi(sel[c] - sgn * sum->gcd[c], 0, a.maxv);
    }
  }
  // The winner's int32 squares, sign-extended into the 64-bit row sums (xCorrespPixelShiftSTD.cpp:35-36:
  // a square wraps only for differences of 46341 or more, i.e. samples beyond the bit depth).
  for (int c = 0; c < 3; c++) red[c][threadIdx.x] = (uint64_t)(int64_t)(int32_t)dist_best[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    uint64_t* part = a.partials;
    part[(((int64_t)blockIdx.x * 2 + dir) * 3 + threadIdx.x) * g.h + y] = red[threadIdx.x][0];
  }
}

// rows[d][c][y] = sum over segments of the partials.
// flag / mode: 0 always; 1 only when *flag != 0; 2 only when *flag == 0 (the range-gated search paths).
__global__ __launch_bounds__(256) void k_cps_rows(const uint64_t* __restrict__ part, int32_t nseg, int32_t h,
                                                  uint64_t* __restrict__ rows, const int32_t* __restrict__ flag,
                                                  int32_t mode) {
  if (mode == 1 && *flag == 0) return;
  if (mode == 2 && *flag != 0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = 6 * (int64_t)h;
  if (i >= n) return;
  uint64_t s = 0;
  for (int k = 0; k < nseg; k++) s += part[(int64_t)k * n + i];
  rows[i] = s;
}

}  // namespace

hipError_t launch_stats(const void* p0, const void* p1, const Geom& g, int64_t* row_sd, uint64_t* row_ssd,
                        uint32_t* row_max, hipStream_t stream, const FrameSummary* gate, uint64_t* zero,
                        int32_t nzero) {
  const int nrows = g.h + 2 * g.ch;
  if (nzero < 0 || (nzero > 0 && !zero)) return hipErrorInvalidValue;
  const int blocks = std::max((nrows + 3) / 4, (nzero + 255) / 256);
  dim3 grid(blocks), block(256);
  if (g.bytes_per_sample == 1)
    hipLaunchKernelGGL(k_stats<uint8_t>, grid, block, 0, stream, (const uint8_t*)p0, (const uint8_t*)p1, g, row_sd,
                       row_ssd, row_max, gate, zero, nzero);
  else
    hipLaunchKernelGGL(k_stats<uint16_t>, grid, block, 0, stream, (const uint16_t*)p0, (const uint16_t*)p1, g,
                       row_sd, row_ssd, row_max, gate, zero, nzero);
  return hipGetLastError();
}

hipError_t launch_summary(const Geom& g, const int64_t* row_sd, const uint64_t* row_ssd, const uint32_t* row_max,
                          const int32_t maxdiff[4], int32_t compute_gcd, FrameSummary* sum, hipStream_t stream,
                          const FrameSummary* gate) {
  if (gate == sum) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_summary, dim3(1), dim3(1024), 0, stream, g, row_sd, row_ssd, row_max, maxdiff[0], maxdiff[1],
                     maxdiff[2], maxdiff[3], compute_gcd, sum, gate);
  return hipGetLastError();
}

hipError_t launch_cps_fast(const CpsArgs& a, hipStream_t stream) {
  const int bands = (a.g.h + kCpsTileH - 1) / kCpsTileH;
  if (a.tiles_x != (a.g.w + kCpsTileW - 1) / kCpsTileW) return hipErrorInvalidValue;
  dim3 grid(bands * a.nseg), block(256);
  if (a.mask) {  // masked search: no optional outputs
    if (a.scp0 || a.scp1 || a.best0 || a.best1 || !a.msum || (a.mask_bytes_per_sample != 1 && a.mask_bytes_per_sample != 2))
      return hipErrorInvalidValue;
    if (a.g.bytes_per_sample == 1) {
      if (a.mask_bytes_per_sample == 1) hipLaunchKernelGGL((k_cps_fast<uint8_t, uint8_t, true>), grid, block, 0, stream, a);
      else hipLaunchKernelGGL((k_cps_fast<uint8_t, uint16_t, true>), grid, block, 0, stream, a);
    } else {
      if (a.mask_bytes_per_sample == 1) hipLaunchKernelGGL((k_cps_fast<uint16_t, uint8_t, true>), grid, block, 0, stream, a);
      else hipLaunchKernelGGL((k_cps_fast<uint16_t, uint16_t, true>), grid, block, 0, stream, a);
    }
    return hipGetLastError();
  }
  if (a.g.bytes_per_sample == 1)
    hipLaunchKernelGGL((k_cps_fast<uint8_t, uint8_t, false>), grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL((k_cps_fast<uint16_t, uint8_t, false>), grid, block, 0, stream, a);
  return hipGetLastError();
}

void cps_fast_tile(int32_t* tile_w, int32_t* tile_h) {
  *tile_w = kCpsTileW;
  *tile_h = kCpsTileH;
}

int cps_fast_blocks_per_cu(int bytes_per_sample) {
  int n = 0;
  hipError_t e = bytes_per_sample == 1
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps_fast<uint8_t, uint8_t, false>, 256, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_cps_fast<uint16_t, uint8_t, false>, 256, 0);
  return e == hipSuccess && n > 0 ? n : 1;
}

hipError_t launch_cps_generic(const CpsArgs& a, hipStream_t stream) {
  if ((int64_t)a.nseg * kGenericCols < a.g.w) return hipErrorInvalidValue;  // every column needs a block
  dim3 grid(a.nseg, a.g.h, 2), block(kGenericCols);
  if (a.g.bytes_per_sample == 1)
    hipLaunchKernelGGL(k_cps_generic<uint8_t>, grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL(k_cps_generic<uint16_t>, grid, block, 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_cps_rows(const uint64_t* partials, int32_t nseg, int32_t h, uint64_t* rows, hipStream_t stream,
                           const int32_t* flag, int32_t mode) {
  if (mode != 0 && !flag) return hipErrorInvalidValue;
  const int64_t n = 6 * (int64_t)h;
  hipLaunchKernelGGL(k_cps_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, partials, nseg, h, rows,
                     flag, mode);
  return hipGetLastError();
}

}  // namespace qmiv
