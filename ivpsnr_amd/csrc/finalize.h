// Host-side double-precision finalisation of the metric engine (finalize.cpp is compiled with
// -ffp-contract=off so every expression rounds like the reference's ISO C++ build).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace qmiv {
namespace fin {

// xPSNR::CalcPSNRfromSSD (libs/pmbb-ivqm/src/xPSNR.cpp:139-146): DBL_MAX for SSD <= 0.
double psnr_from_ssd(double ssd, int32_t area, int32_t bit_depth);

// Neumaier compensated summation as xKBNS::Accumulate (libs/pmbb-core/src/xMathUtils.cpp:12-26).
double kbns(const double* x, size_t n);

// xWeightedSpherically::initWS (libs/pmbb-ivqm/src/xWeightedSpherically.cpp:13-29).
void ws_weights(int32_t height, int32_t lat_range_deg, std::vector<double>& weights, double& correction);

// xIVPSNR::xCalcQualAsymmetricPic tail (libs/pmbb-ivqm/src/xIVPSNR.cpp:115-132): rows is
// [3][height] per-row values; ws (nullable) the WS weights, applied without correction.
double qual_from_rows(const uint64_t* rows, int32_t height, const double* ws, int32_t area, int32_t bit_depth,
                      const int32_t wavg[4]);

// WS-PSNR of one component from its 4:4:4 per-row SSD (libs/pmbb-ivqm/src/xWSPSNR.cpp:101-126).
double wspsnr_from_rows(const uint64_t* rows444, int32_t height, const double* ws, double correction,
                        int32_t area, int32_t bit_depth);

// xPSNR::CalcPSNRfromMaskedSSD (libs/pmbb-ivqm/src/xPSNR.cpp:148-155): peak = nnm * max^2 * maxm in
// int64 arithmetic; DBL_MAX for SSD <= 0.
double psnr_from_masked_ssd(double ssd, int32_t nnm, int32_t bit_depth, int32_t mask_bit_depth);

// xIVPSNRM::xCalcQualAsymmetricPicM tail (libs/pmbb-ivqm/src/xIVPSNR.cpp:170-206): as qual_from_rows
// with the masked peak.
double qual_from_rows_masked(const uint64_t* rows, int32_t height, const double* ws, int32_t nnm, int32_t bit_depth,
                             int32_t mask_bit_depth, const int32_t wavg[4]);

// Masked WS-PSNR of one component from its per-row masked SSD (libs/pmbb-ivqm/src/xWSPSNR.cpp:127-155).
double wspsnr_from_rows_masked(const uint64_t* rows444, int32_t height, const double* ws, double correction,
                               int32_t nnm, int32_t bit_depth, int32_t mask_bit_depth);

// WS-PSNR of the three components at once (their weighted KBNS sums interleaved, each one's
// arithmetic unchanged): row y of component c is rows[c][y >> shift[c]] << shift[c] (a 4:2:0
// chroma plane's row SSD doubled horizontally and repeated vertically onto the 4:4:4 rows).
void wspsnr3_from_rows(const uint64_t* const rows[3], const int shift[3], int32_t height, const double* ws,
                       double correction, int32_t area, int32_t bit_depth, double out[3]);

// Component-weighted picture value as xMetricStat::setPerCmpMeric (apps/QMIV/src/xAppQMIV.h:245-251).
double pic_value(const double v[3], const int32_t wavg[4]);

// (v * (flt64V4)w).getSum() * (1 / sum(w)) over four lanes with v[3] = 0 (xVec.h:308).
double weighted_average4(const double v[3], const int32_t wavg[4]);

// GCD clip bound: xRoundFltToInt32(coef * (flt32)max) in single precision, half-even
// (libs/pmbb-ivqm/src/xGlobClrDiff.cpp:21).
int32_t gcd_max_diff(float coef, int32_t bit_depth);

// The SSIM window: the normalised 11 x 11 Gaussian (sigma 1.5) with each tap rounded to ten
// decimals and stored in single precision, as the reference's table holds it
// (libs/pmbb-ivqm/src/xStructSimConsts.h:30-43); out[dy * 10 + dx] = F[dy][dx] for the
// 10 x 10 part the window uses (libs/pmbb-ivqm/src/xStructSim.cpp:19-21).
void ssim_window(double out[100]);

// SSIM of one picture from its per-row sums: KBNS over the rows / (w * h)
// (libs/pmbb-ivqm/src/xSSIM.cpp:104-111).
double ssim_from_rows(const double* rows, int32_t h, int32_t w);

// MS-SSIM: product of score_i ^ weight_i in scale order with std::pow (xSSIM.cpp:83-84).
double msssim_combine(const double scores[5]);

// std::min argument order (xCommonDefCORE.h:93).
double min2(double a, double b);

// Q16 RGB -> YCbCr (to_rgb false) or YCbCr -> RGB matrix, row-major, of colour space 0..3 (BT.601,
// BT.709, SMPTE 240M, BT.2020): the textbook matrices from the luma weights (Kr, Kg, Kb) in
// double, each entry scaled by 65536 and rounded half away from zero
// (libs/pmbb-core/src/xColorSpaceCoeff.h:36-100,171-213). Returns false for an unknown space.
bool colour_matrix(int32_t space, bool to_rgb, int32_t m[9]);

}  // namespace fin
}  // namespace qmiv
