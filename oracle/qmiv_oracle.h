/*
  Plain-C restatement of the reference's per-frame metric engine (the oracle).

  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU-baseline leg
  of the benchmark load this library as the checker. The product library never
  links, loads or calls it.

  Pictures are 4:4:4 uint16 planes without margin (stride = width); the
  reference's margin replication (libs/pmbb-core/src/xPixelOpsSTD.cpp:210-236)
  is restated as clamp-to-edge addressing, which reads identical values while
  the search range does not exceed the margin.
*/
#ifndef QMIV_ORACLE_H
#define QMIV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Metric selection bits (same order as the reference application's metric list). */
enum {
  ORC_PSNR = 1, ORC_WSPSNR = 2, ORC_IVPSNR = 4, ORC_SSIM = 8, ORC_MSSSIM = 16, ORC_IVSSIM = 32
};

typedef struct {
  int32_t width, height, bit_depth, chroma_format; /* chroma_format: 400, 420, 422 or 444 (file layout) */
  int32_t search_range;                           /* B */
  int32_t cmp_weights_search[4];
  int32_t cmp_weights_average[4];
  float unnoticeable_coef[4];
  int32_t erp;                                    /* equirectangular: WS weights on */
  int32_t lon_range_deg, lat_range_deg;
  int32_t cps_int32_rows;                         /* 1: SSE-path int32 per-row sums (app default), 0: uint64 */
  int32_t metrics;                                /* ORC_* bits */
  int32_t conceal;                                /* 1: InvalidPelActn CNCL, else the samples as read */
} orc_params;

typedef struct {
  int32_t in_range[2];
  int32_t exact[3];
  int32_t gcd[4];
  double psnr[3], psnr_pic;
  double wspsnr[3], wspsnr_pic;
  double ivpsnr, r2t, t2r;
  uint64_t cps_sum[2][3];     /* direction (R2T, T2R) x component: sum of the per-row values */
  double ssim[3], ssim_pic;
  double msssim[3], msssim_pic;
  double msssim_scales[5][3];
  double ivssim, ivssim_r2t, ivssim_t2r;
} orc_frame_result;

void orc_default_params(orc_params* p, int32_t width, int32_t height, int32_t bit_depth, int32_t chroma_format);

/* xSeqBase::xUnpackFrame (libs/pmbb-core/src/xSeq.cpp:179-249): packed planar file frame (400, 420,
   422 or 444, any size) -> 4:4:4 planes. */
int orc_unpack(const void* packed, int32_t w, int32_t h, int32_t bd, int32_t cf, uint16_t* y, uint16_t* cb, uint16_t* cr);

/* xPixelOpsSTD::CheckIfInRange (xPixelOpsSTD.cpp:163-174). */
int orc_in_range(const uint16_t* p, int32_t n, int32_t bd);

/* xGlobClrDiff::CalcGlobalColorDiff (libs/pmbb-ivqm/src/xGlobClrDiff.cpp:15-48) called as the
   processor does it (xGlobClrDiff.h:41): g = clip(rint(SD(p1 - p0) / area)). */
void orc_gcd(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd,
             const float ucoef[4], int32_t g[4]);

/* CPS search + row distortion (xCorrespPixelShiftSSE.cpp:16-71, xCorrespPixelShiftSTD.cpp:18-75):
   each pel of X (+g) is matched inside the +-B window of Y. rows[c*h + y]. best (optional, w*h)
   receives the raster index of the winning offset. */
void orc_cps_rows(const uint16_t* const x[3], const uint16_t* const y[3], int32_t w, int32_t h, const int32_t g[4],
                  int32_t B, const int32_t wts[4], int32_t int32_rows, uint64_t* rows, int32_t* best);

/* xShftCompPic::xGenShftCompRow (libs/pmbb-ivqm/src/xShftCompPic.cpp:69-97):
   dst(p) = clipU(Y(p + best(X(p) + g)) - g, max). */
void orc_scp(const uint16_t* const y[3], const uint16_t* const x[3], int32_t w, int32_t h, int32_t bd,
             const int32_t g[4], int32_t B, const int32_t wts[4], uint16_t* const dst[3]);

/* Per-row SSD (xDistortionSTD.cpp:46-62 per row, as xPSNR.cpp:93-112 / xWSPSNR.cpp:101-126). */
void orc_ssd_rows(const uint16_t* a, const uint16_t* b, int32_t w, int32_t h, uint64_t* rows);

/* xPSNR::CalcPSNRfromSSD (libs/pmbb-ivqm/src/xPSNR.cpp:139-146). */
double orc_psnr_from_ssd(double ssd, int32_t area, int32_t bd);

/* xWeightedSpherically::initWS (libs/pmbb-ivqm/src/xWeightedSpherically.cpp:13-29). */
void orc_ws_init(int32_t h, int32_t lat_range_deg, double* weights, double* correction);

/* xKBNS::Accumulate (libs/pmbb-core/src/xMathUtils.cpp:12-26) and xKBNS::acc (xMathUtils.h:23-29). */
double orc_kbns(const double* x, size_t n);

/* xIVPSNR::xCalcQualAsymmetricPic tail (libs/pmbb-ivqm/src/xIVPSNR.cpp:115-132) from per-row values. */
double orc_qual_from_rows(const uint64_t* rows, int32_t h, const double* ws_weights, int32_t area, int32_t bd,
                          const int32_t wavg[4]);

/* The 11x11 Gaussian window (xStructSimConsts.h:30-43), derived from its closed form. */
void orc_ssim_filter(float f[121]);

/* xSSIM::xCalcCmpSSIM (libs/pmbb-ivqm/src/xSSIM.cpp:91-141) with xStructSim::CalcPel (xStructSim.cpp:15-60).
   row_sums (optional, h entries) receives the per-row sums. */
double orc_ssim_plane(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t bd, int32_t calc_l,
                      double* row_sums);

/* xPixelOpsSTD::DownsampleHV (xPixelOpsSTD.cpp:50-67). */
void orc_downsample(uint16_t* dst, const uint16_t* src, int32_t dw, int32_t dh, int32_t src_stride);

/* xSSIM::calcPicMSSSIM for one component (libs/pmbb-ivqm/src/xSSIM.cpp:67-87). */
double orc_msssim_plane(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t bd, double scales[5]);

/* One frame through the reference application's frame flow (apps/QMIV/src/xAppQMIV.cpp:661-935).
   p0 / p1 are packed frames (InputFile0 / InputFile1). scp0 / scp1 (optional, 3*w*h each) receive
   the shift-compensated pictures when IV-SSIM is selected. */
int orc_frame(const void* p0, const void* p1, const orc_params* prm, orc_frame_result* out, uint16_t* scp0,
              uint16_t* scp1);

/* Masked mode (apps/QMIV/src/xAppQMIV.cpp:319-337,695,825-889): PSNR, WS-PSNR and IV-PSNR over
   the pels whose mask (luma plane of the mask picture) is non-zero, each weighted by the mask
   value. r2t / t2r carry the masked processor's own labels (libs/pmbb-ivqm/src/xIVPSNR.cpp:152-165):
   R2T searches P1 - g inside P0, T2R searches P0 + g inside P1. cps_sum[0] is R2T. */
typedef struct {
  int32_t in_range[3];        /* P0, P1, mask */
  int32_t exact[3];
  int32_t num_non_masked;
  int32_t gcd[4];
  double psnr[3], psnr_pic;
  double wspsnr[3], wspsnr_pic;
  double ivpsnr, r2t, t2r;
  uint64_t cps_sum[2][3];
} orc_masked_result;

/* xGlobClrDiff::CalcGlobalColorDiffM (libs/pmbb-ivqm/src/xGlobClrDiff.cpp:52-83) as the processor
   calls it (xGlobClrDiff.h:42): g = clip(cvt(sum((p1 - p0) * m) / (nnm * maxm))). */
void orc_gcd_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* mask, int32_t w,
                    int32_t h, int32_t bd, int32_t bdm, int32_t nnm, const float ucoef[4], int32_t g[4]);

/* xCorrespPixelShiftSTD::CalcDistAsymmetricRowM + FindBestPixelWithinBlockM
   (libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp:80-144): rows[c*h + y] (uint64). */
void orc_cps_rows_masked(const uint16_t* const x[3], const uint16_t* const y[3], const uint16_t* mask, int32_t w,
                         int32_t h, const int32_t g[4], int32_t B, const int32_t wts[4], uint64_t* rows);

/* xPSNR::CalcPSNRfromMaskedSSD (libs/pmbb-ivqm/src/xPSNR.cpp:148-155). */
double orc_psnr_from_masked_ssd(double ssd, int32_t nnm, int32_t bd, int32_t bdm);

/* One frame of the masked flow. mask is a packed frame of bit depth bdm and chroma format cfm. */
int orc_frame_masked(const void* p0, const void* p1, const void* mask, int32_t bdm, int32_t cfm,
                     const orc_params* prm, orc_masked_result* out);

/* KBNS average over frames (apps/QMIV/src/xAppQMIV.h:256-260). */
double orc_frame_average(const double* v, size_t n);

/* Colour space: Q16 matrix of space 0..3 (BT.601, BT.709, SMPTE 240M, BT.2020) for RGB -> YCbCr
   (to_rgb = 0) or YCbCr -> RGB (to_rgb = 1), row-major; and the conversion of three planar planes
   of n samples (xColorSpaceSTD.cpp:195-272). */
int orc_colorspace_coeffs(int32_t space, int32_t to_rgb, int32_t m[9]);
int orc_colorspace(const uint16_t* src, uint16_t* dst, size_t n, int32_t bd, int32_t space, int32_t to_rgb);

#ifdef __cplusplus
}
#endif
#endif
