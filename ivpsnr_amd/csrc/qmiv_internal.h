// Internal declarations shared by the gfx950 kernels (*.hip) and the host runtime (*.cpp).
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace qmiv {

// Geometry of one packed planar frame as it sits in device memory.
struct Geom {
  int32_t w, h;      // luma size
  int32_t cw, ch;    // chroma plane size
  int32_t cs;        // chroma subsampling shift: 1 for 4:2:0, 0 for 4:4:4
  int32_t bd;        // bit depth
  int32_t bytes_per_sample;
  uint32_t frame_bytes;  // bytes of one packed frame (< 4 GiB)
  int64_t cb_off;    // sample offset of the Cb plane
  int64_t cr_off;    // sample offset of the Cr plane
  // 420, 444, or 400: a luma plane only (a mask picture), whose chroma views alias the luma plane
  // (cs 0, cw x ch = w x h, both chroma offsets 0).
  int32_t cf;
};

// Per-frame summary written by the statistics pass (device memory, read back by the host).
struct FrameSummary {
  int32_t gcd[4];          // GCD R2T (also the GCD consumed by the CPS kernels)
  int32_t max_pel[2];      // largest pel of P0 / P1 over all planes
  int32_t in_range;        // both pictures within [0, 2^bd - 1]
  int32_t pad_;
  int64_t sd[3];           // 4:4:4-domain sum(P1 - P0), exact
  uint64_t ssd[3];         // 4:4:4-domain sum((P1 - P0)^2), exact
};

// Row layout of the per-row statistics: luma rows [0, h), Cb rows [h, h + ch), Cr rows [h + ch, h + 2ch).
inline int32_t stat_rows(const Geom& g) { return g.h + 2 * g.ch; }

struct MaskSummary;

// Arguments of the corresponding-pixel-shift kernels.
struct CpsArgs {
  const void* p0;
  const void* p1;
  Geom g;
  const FrameSummary* sum;   // gcd and in_range
  int32_t ws[4];             // search weights
  int32_t B;                 // search range (generic kernel)
  int32_t tiles_x;           // tiles per band
  int32_t sweep;             // tiles per workgroup
  int32_t nseg;              // segments per band (= ceil(tiles_x / sweep)); generic kernel: column blocks
  uint64_t* partials;        // [nseg][2][3][h] per-segment row sums
  // 4:2:0 fast kernel: when set, the row sums are added atomically into rows[2][3][h] (cleared
  // beforehand) instead of being written to partials, so no reduction pass follows.
  uint64_t* rows = nullptr;
  uint16_t* scp0;            // optional [3][h][w]: clipU(P0(best of P1 - g) + g)   (SCP of InputFile0)
  uint16_t* scp1;            // optional [3][h][w]: clipU(P1(best of P0 + g) - g)   (SCP of InputFile1)
  int16_t* best0;            // optional [h][w]: R2T winner, window raster index (dy + B)(2B + 1) + dx + B
  int16_t* best1;            // optional [h][w]: T2R winner
  int32_t maxv;
  // Masked search (4:2:0 fast kernel only): the mask picture (its luma plane at offset 0, one sample
  // per pel), the masked summary (GCD, range flags) in place of sum, no optional outputs.
  const void* mask = nullptr;
  uint32_t mask_bytes = 0;
  int32_t mask_bytes_per_sample = 0;
  const MaskSummary* msum = nullptr;
  // When the search runs, by the summary's range flag (all inputs within the bit depth):
  // kGateInRange (the fast kernels' only mode), kGateOutOfRange (the general kernels as the fallback
  // of a frame that holds pels above the bit depth) or kGateAlways.
  int32_t gate = 0;
};
constexpr int32_t kGateInRange = 0, kGateOutOfRange = 1, kGateAlways = 2;
__host__ __device__ inline bool cps_gate_open(int32_t gate, int32_t in_range) {
  return gate == kGateAlways || (gate == kGateInRange ? in_range != 0 : in_range == 0);
}

// One plane read through its 4:4:4 view: value(x, y) = base[(y >> shift) * stride + (x >> shift)].
struct PlaneView {
  const void* base;
  int32_t stride;   // samples per row
  int32_t shift;    // 1 for 4:2:0 chroma of a packed frame, else 0
  int32_t bytes;    // bytes per sample (1 or 2)
  uint32_t nbytes;  // readable bytes from base
};

// Arguments of the SSIM kernel (the 10 x 10 window travels with the launch).
struct SsimArgs {
  PlaneView tst, ref;
  int32_t w, h;          // 4:4:4 size of this (pyramid level's) picture
  int32_t calc_l;        // 1: luminance x contrast-structure; 0: contrast-structure only
  int32_t tiles_x;       // ssim_tiles_x(w)
  double c1, c2;
  double* partials;      // [tiles_x][h]
  double f[100];         // window weights F[dy + 5][dx + 5], dy, dx in [-5, 4]
};

// Arguments of the fused full-resolution SSIM kernel: one component, pictures P0, P1 and the
// shift-compensated SCP1 / SCP0 (IV-SSIM), and one tile-partials array per output (null: not asked).
struct SsimMultiArgs {
  PlaneView p0, p1, scp1, scp0;
  int32_t w, h;
  int32_t tiles_x;       // ssim_tiles_x(w)
  double c1, c2;
  double* part_lcs;      // SSIM(P0, P1), L·CS
  double* part_cs;       // CS(P0, P1) (MS-SSIM level 0, or SSIM without L)
  double* part_c;        // SSIM(P0, SCP1), L·CS
  double* part_d;        // SSIM(P1, SCP0), L·CS
  double f[100];
};

// Arguments of the batched SSIM kernel: several evaluations (pyramid levels, components) in one
// launch. tile0 is filled in by the launcher.
constexpr int kMaxSsimBatch = 16;
struct SsimBatchJob {
  PlaneView tst, ref;
  int32_t w, h;
  int32_t tiles_x;       // ssim_tiles_x(w)
  int32_t calc_l;        // 1: L x CS; 0: CS only
  int32_t tile0;         // first workgroup of this job
  double* partials;      // [tiles_x][h]
};
struct SsimBatchArgs {
  int32_t n;
  double c1, c2;
  double f[100];
  SsimBatchJob job[kMaxSsimBatch];
};

// Kernel launchers (kernels.hip). All launch on `stream` and return hipSuccess or the launch error.
// gate (nullable): a first pass's summary; a pass whose gate found both pictures in range does
// nothing (statistics) or copies the gate (summary): the statistics re-run after concealment.
// zero / nzero (launch_stats): an array that the statistics pass also clears (the per-row CPS sums
// that the 4:2:0 fast kernel adds into), whatever the gate says.
hipError_t launch_stats(const void* p0, const void* p1, const Geom& g, int64_t* row_sd, uint64_t* row_ssd,
                        uint32_t* row_max, hipStream_t stream, const FrameSummary* gate = nullptr,
                        uint64_t* zero = nullptr, int32_t nzero = 0);
hipError_t launch_summary(const Geom& g, const int64_t* row_sd, const uint64_t* row_ssd, const uint32_t* row_max,
                          const int32_t maxdiff[4], int32_t compute_gcd, FrameSummary* sum, hipStream_t stream,
                          const FrameSummary* gate = nullptr);
hipError_t launch_cps_fast(const CpsArgs& a, hipStream_t stream);
hipError_t launch_cps_generic(const CpsArgs& a, hipStream_t stream);
// flag / mode: 0 always; 1 only when *flag != 0; 2 only when *flag == 0.
hipError_t launch_cps_rows(const uint64_t* partials, int32_t nseg, int32_t h, uint64_t* rows, hipStream_t stream,
                           const int32_t* flag = nullptr, int32_t mode = 0);
// ---- Differentiable SSIM family on float planes (kernels_grad.hip) ----
// Forward moments of one plane pair; samples are src * scale (src fp32, or fp64 when *_f64).
struct SsimFArgs {
  const void* tst;
  const void* ref;
  int32_t tst_f64, ref_f64;
  double scale_t, scale_r;
  int32_t w, h;
  int32_t tiles_x;       // ssimf_tiles_x(w)
  double c1, c2;
  double* partials;      // [tiles_x][h] row sums of the SSIM map
  double* maps;          // [4][h][w] A_T, A_R, B, Cm over W H (interior pels; others untouched), or null
  double f[100];         // window weights F[dy + 5][dx + 5]
};
// Gradient from the maps: g = coef * scale * (...), stored or added (accumulate) where not null.
struct SsimGArgs {
  const void* tst;
  const void* ref;
  int32_t tst_f64, ref_f64;
  double scale_t, scale_r;
  int32_t w, h;
  const double* maps;
  float* grad_t;
  float* grad_r;
  double coef_t, coef_r;
  int32_t accumulate;
  double f[100];
};
// Shift-compensated float picture (3 planes) from the winners of one search direction.
struct ScpFArgs {
  const float* src;      // searched picture (3 planes)
  const int16_t* best;   // [h][w] winner indices
  int32_t w, h, B;
  double scale;
  int32_t sgn;           // +1: Y - gcd (SCP of InputFile1); -1: Y + gcd (SCP of InputFile0)
  int32_t gcd[3];
  const int32_t* gcd_dev = nullptr;  // when set, the GCD is read from device memory instead (3 values)
  int32_t maxv;
  float* out;            // [3][h][w] in sample units
};
int32_t ssimf_tiles_x(int32_t w);
hipError_t launch_ssimf(const SsimFArgs& a, int32_t calc_l, int32_t grad, hipStream_t stream);
hipError_t launch_ssimf_grad(const SsimGArgs& a, hipStream_t stream);
// dst (fp64, w/2 x h/2) = exact mean of the 2 x 2 blocks of src (fp32, or fp64 when src_f64).
hipError_t launch_down2(const void* src, int32_t src_f64, int32_t w, int32_t h, double* dst, hipStream_t stream);
// fine = coef * fine + coarse(parent) / 4 (coarse may be null: scaling only).
hipError_t launch_up2_addf(const float* coarse, double coef, int32_t w, int32_t h, float* fine, hipStream_t stream);
hipError_t launch_quantf(const float* src, int64_t n, double scale, int32_t maxv, void* dst, int32_t bytes,
                         hipStream_t stream);
hipError_t launch_scp_gatherf(const ScpFArgs& a, hipStream_t stream);
// The adjoint of launch_scp_gatherf: grad_src[q] = coef * scale * (sum of g over the pels whose winner
// is q, for the components whose shifted value at q is not clamped), written for every pel q (0 where
// nothing applies).
hipError_t launch_scp_scatterf(const ScpFArgs& a, const float* g, double coef, float* grad_src, hipStream_t stream);

// Tile shape of the fast CPS kernel (pels per workgroup tile).
void cps_fast_tile(int32_t* tile_w, int32_t* tile_h);
// Resident workgroups of the fast CPS kernel per CU (from the HIP occupancy calculator).
int cps_fast_blocks_per_cu(int bytes_per_sample);
// The 4:2:0 fast CPS kernel (kernels_cps420.hip): chroma cells factored out of the key.
hipError_t launch_cps_fast420(const CpsArgs& a, hipStream_t stream);
void cps_fast420_tile(int32_t* tile_w, int32_t* tile_h);
int cps_fast420_blocks_per_cu(int bytes_per_sample);
// The 4:2:0 CPS kernel with 16-bit key operands (kernels_cps420s.hip): samples of at most 10 bits,
// row sums added into CpsArgs::rows, no mask.
hipError_t launch_cps420s(const CpsArgs& a, hipStream_t stream);
void cps420s_tile(int32_t* tile_w, int32_t* tile_h);
int cps420s_blocks_per_cu(int bytes_per_sample);
// Fast CPS kernels, as the host plans them.
constexpr int kCps444 = 0, kCps420 = 1, kCps420s = 2;

// Nearest-neighbour 4:2:0 -> 4:4:4 upsampling (a copy for 4:4:4 sources) of a packed frame into a
// planar 4:4:4 frame of the same sample type.
// Masked mode: per-row statistics of the 4:4:4 domain (device memory).
struct MaskRows {
  int32_t* nnz;    // [h] non-zero mask pels
  int64_t* sd;     // [3][h] sum((P1 - P0) * m), int32 products
  uint64_t* ssd;   // [3][h] sum(sq(P1 - P0) * m)
  uint32_t* max;   // [3][h] largest pel of P0, P1, mask (all planes)
  uint32_t* diff;  // [h] bit c: component c differs somewhere in the row
};

// Masked mode: plane totals, flags and the GCD used by the masked search (device memory, read
// back by the host).
struct MaskSummary {
  int32_t gcd[4];
  int32_t nnm;
  int32_t in_range[3];     // P0, P1, mask
  int32_t all_in_range;
  int32_t exact[3];
  int64_t sd[3];
  uint64_t ssd[3];
  int32_t max_pel[3];      // largest pel of P0, P1 (all planes) and the mask
};

// gate (nullable): as launch_stats / launch_summary, by all_in_range.
hipError_t launch_mstats(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm,
                         const MaskRows& r, hipStream_t st, const MaskSummary* gate = nullptr);
// gcd_given: null to compute the masked GCD, else the four values to store; nnm_given >= 0
// replaces the counted pels in the GCD's denominator.
hipError_t launch_msummary(const Geom& g, int32_t maxm, const MaskRows& r, const int32_t maxdiff[4], int32_t nnm_given,
                           const int32_t* gcd_given, MaskSummary* out, hipStream_t st,
                           const MaskSummary* gate = nullptr);
constexpr int32_t kMaxMaskedSearchRange = 12;
// partials: [masked_cps_segments(g)][2][3][h], reduced by launch_cps_rows.
hipError_t launch_cps_masked(const void* p0, const void* p1, const Geom& g, const void* mask, const Geom& gm,
                             const MaskSummary* sum, int32_t B, const int32_t w[4], uint64_t* partials,
                             hipStream_t st, int32_t gate = kGateInRange);
size_t masked_cps_lds_bytes(int32_t B);
int32_t masked_cps_segments(const Geom& g);

// Q16 matrix (row-major) and range of an RGB <-> YCbCr conversion.
struct ColourArgs {
  int32_t m[9];
  int32_t mid;   // 2^(bd - 1)
  int32_t maxv;  // 2^bd - 1
};

hipError_t launch_colour(const void* src, void* dst, const Geom& g, const ColourArgs& k, bool to_rgb,
                         hipStream_t stream);
hipError_t launch_swap_planes(void* pic, const Geom& g, int a, int b, hipStream_t stream);
// xPixelOps::ClipToRange over every sample of a 16-bit frame (libs/pmbb-core/src/xPixelOpsSTD.cpp:198-209),
// when *max_pel exceeds maxv (null: always).
hipError_t launch_clip(void* pic, const Geom& g, int32_t maxv, const int32_t* max_pel, hipStream_t stream);
hipError_t launch_upsample(const void* src, const Geom& gs, void* dst, hipStream_t stream);
// One packed raw-file frame (src: src_bytes in device memory, file chroma format fcf 400 / 420 / 422 /
// 444, any w and h) unpacked into a 4:4:4 frame of the same sample size (xSeqBase::xUnpackFrame).
hipError_t launch_unpack(const void* src, uint64_t src_bytes, int32_t w, int32_t h, int32_t bd, int32_t fcf,
                         void* dst, hipStream_t stream);
// dst[i] = (uint8_t)src[i] for i < n.
hipError_t launch_narrow(const uint16_t* src, int64_t n, uint8_t* dst, hipStream_t stream);
// dst[i] = src[i] for i < n.
hipError_t launch_widen(const uint8_t* src, int64_t n, uint16_t* dst, hipStream_t stream);
// A planar 4:4:4 frame of g to (to_interleaved) or from the reference's interleaved layout (4 x uint16
// per pel, rows packed, A = 0).
hipError_t launch_interleave(const void* planar, const Geom& g, uint16_t* inter, bool to_interleaved,
                             hipStream_t stream);

// Windowed moments of one evaluation into a.partials ([tiles_x][h] row sums per tile column).
hipError_t launch_ssim(const SsimArgs& a, hipStream_t stream);
hipError_t launch_ssim_multi(const SsimMultiArgs& a, hipStream_t stream);
hipError_t launch_ssim_batch(SsimBatchArgs& a, hipStream_t stream);
int32_t ssim_tiles_x(int32_t w);

// Per-row reduction of every SSIM evaluation of one frame in one launch:
// rows_j[y] = sum over tile columns t of part_j[t * h_j + y], in tile order (zero outside the
// interior rows [5, h_j - 5)).
constexpr int kMaxSsimEvals = 32;
struct SsimRowsJobs {
  int32_t n;
  int32_t max_h;
  int32_t tiles_x[kMaxSsimEvals];
  int32_t h[kMaxSsimEvals];
  const double* part[kMaxSsimEvals];
  double* rows[kMaxSsimEvals];
};
hipError_t launch_ssim_rows(const SsimRowsJobs& j, hipStream_t stream);

// One MS-SSIM pyramid level for up to 6 planes in one launch:
// dst_i[y][x] = (s(2x, 2y) + s(2x + 1, 2y) + s(2x, 2y + 1) + s(2x + 1, 2y + 1) + 2) >> 2.
// MS-SSIM levels 1-4 of up to six planes in one launch: dst[p][l - 1] is level l, (w >> l) x (h >> l)
// uint16 (stride w >> l), of the 4:4:4 view src[p] of size w x h.
struct PyramidJobs {
  int32_t n;
  int32_t w, h;
  PlaneView src[6];
  uint16_t* dst[6][4];
};
hipError_t launch_pyramid(const PyramidJobs& j, hipStream_t stream);

// Tile shape of the fast CPS kernel (device code only; the host asks cps_fast_tile()).
constexpr int kCpsTileW = 64;   // 64 lanes x 1 column
#ifndef QMIV_CPS_TILE_H
#define QMIV_CPS_TILE_H 16
#endif
constexpr int kCpsTileH = QMIV_CPS_TILE_H;  // 4 waves x (kCpsTileH / 4) rows
constexpr int kGenericCols = 256;

// Fixed device scratch of the allocation-free frame entry point (fixed_scratch.hip): its size, and
// its device address on `device`, which must be the current device. The address is resolved once
// per device and kept; later calls return it.
size_t fixed_scratch_bytes();
hipError_t fixed_scratch_setup(int device, void** ptr);

}  // namespace qmiv
