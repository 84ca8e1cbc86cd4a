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

// Component-weighted picture value as xMetricStat::setPerCmpMeric (apps/QMIV/src/xAppQMIV.h:245-251).
double pic_value(const double v[3], const int32_t wavg[4]);

// (v * (flt64V4)w).getSum() * (1 / sum(w)) over four lanes with v[3] = 0 (xVec.h:308).
double weighted_average4(const double v[3], const int32_t wavg[4]);

// GCD clip bound: xRoundFltToInt32(coef * (flt32)max) in single precision, half-even
// (libs/pmbb-ivqm/src/xGlobClrDiff.cpp:21).
int32_t gcd_max_diff(float coef, int32_t bit_depth);

// std::min argument order (xCommonDefCORE.h:93).
double min2(double a, double b);

}  // namespace fin
}  // namespace qmiv
