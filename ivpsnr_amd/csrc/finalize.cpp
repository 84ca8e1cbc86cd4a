// Host-side double-precision finalisation. Compiled with -ffp-contract=off (no fused
// multiply-add), so each expression below rounds exactly like the reference's ISO C++ build;
// the integer inputs come bit-exact from the device, hence the doubles are bit-exact too.
#include "finalize.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace qmiv {
namespace fin {

double psnr_from_ssd(double ssd, int32_t area, int32_t bit_depth) {
  const uint64_t points = (uint64_t)(int64_t)area;
  const uint64_t maxv = (uint64_t)((1 << bit_depth) - 1);
  const uint64_t peak = points * (maxv * maxv);
  return ssd > 0 ? 10.0 * std::log10((double)peak / ssd) : DBL_MAX;
}

namespace {
inline void neumaier(double& s, double& c, double v) {
  const double t = s + v;
  if (std::fabs(s) >= std::fabs(v)) c += ((s - t) + v);
  else c += ((v - t) + s);
  s = t;
}
}  // namespace

double kbns(const double* x, size_t n) {
  if (n == 0) return 0.0;
  double s = x[0], c = 0.0;
  for (size_t i = 1; i < n; i++) neumaier(s, c, x[i]);
  return s + c;
}

void ws_weights(int32_t height, int32_t lat_range_deg, std::vector<double>& weights, double& correction) {
  const double pi = 3.14159265358979323846;
  const double eh = 180.0 * (double)height / (double)lat_range_deg;
  const double off = (eh - height) / 2.0;
  weights.resize((size_t)height);
  for (int32_t i = 0; i < height; i++) weights[(size_t)i] = std::cos((i + off - (eh / 2 - 0.5)) * pi / eh);
  correction = height / kbns(weights.data(), weights.size());
}

namespace {
// One component's error of the asymmetric quality: lane-wise xKBNS4 of row * w[y] from zero
// with WS, else the uint64 sum of the rows (std::accumulate).
// Three Neumaier sums from zero of r_c[y >> sh_c] << sh_c times ws[y], interleaved so that their
// dependency chains overlap; each sum performs the operations of its own loop in the same order.
void kbns3_rows(const uint64_t* const r[3], const int sh[3], int32_t height, const double* ws, double out[3]) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, k0 = 0.0, k1 = 0.0, k2 = 0.0;
  for (int32_t y = 0; y < height; y++) {
    const double w = ws[y];
    const double v0 = (double)(r[0][y >> sh[0]] << sh[0]) * w;
    const double v1 = (double)(r[1][y >> sh[1]] << sh[1]) * w;
    const double v2 = (double)(r[2][y >> sh[2]] << sh[2]) * w;
    const double t0 = s0 + v0, t1 = s1 + v1, t2 = s2 + v2;
    k0 += std::fabs(s0) >= std::fabs(v0) ? ((s0 - t0) + v0) : ((v0 - t0) + s0);
    k1 += std::fabs(s1) >= std::fabs(v1) ? ((s1 - t1) + v1) : ((v1 - t1) + s1);
    k2 += std::fabs(s2) >= std::fabs(v2) ? ((s2 - t2) + v2) : ((v2 - t2) + s2);
    s0 = t0;
    s1 = t1;
    s2 = t2;
  }
  out[0] = s0 + k0;
  out[1] = s1 + k1;
  out[2] = s2 + k2;
}

// Errors of the three components of an asymmetric quality from rows [3][height].
void rows_errors3(const uint64_t* rows, int32_t height, const double* ws, double err[3]) {
  const uint64_t* r[3] = {rows, rows + (size_t)height, rows + 2 * (size_t)height};
  if (ws) {
    const int sh[3] = {0, 0, 0};
    kbns3_rows(r, sh, height, ws, err);
    return;
  }
  for (int c = 0; c < 3; c++) {
    uint64_t s = 0;
    for (int32_t y = 0; y < height; y++) s += r[c][y];
    err[c] = (double)s;
  }
}

}  // namespace

double qual_from_rows(const uint64_t* rows, int32_t height, const double* ws, int32_t area, int32_t bit_depth,
                      const int32_t wavg[4]) {
  double err[3], q[3];
  rows_errors3(rows, height, ws, err);
  for (int c = 0; c < 3; c++) q[c] = psnr_from_ssd(err[c] > 0 ? err[c] : 1.0, area, bit_depth);
  return weighted_average4(q, wavg);
}

void wspsnr3_from_rows(const uint64_t* const rows[3], const int shift[3], int32_t height, const double* ws,
                       double correction, int32_t area, int32_t bit_depth, double out[3]) {
  double err[3];
  kbns3_rows(rows, shift, height, ws, err);
  for (int c = 0; c < 3; c++) out[c] = psnr_from_ssd(err[c] * correction, area, bit_depth);
}

double psnr_from_masked_ssd(double ssd, int32_t nnm, int32_t bit_depth, int32_t mask_bit_depth) {
  const uint64_t maxv = (uint64_t)((1 << bit_depth) - 1);
  const uint64_t maxm = (uint64_t)((1 << mask_bit_depth) - 1);
  const uint64_t peak = (uint64_t)(int64_t)nnm * (maxv * maxv) * maxm;  // int64 products, wrapping
  return ssd > 0 ? 10.0 * std::log10((double)(int64_t)peak / ssd) : DBL_MAX;
}

double qual_from_rows_masked(const uint64_t* rows, int32_t height, const double* ws, int32_t nnm, int32_t bit_depth,
                             int32_t mask_bit_depth, const int32_t wavg[4]) {
  double err[3], q[3];
  rows_errors3(rows, height, ws, err);
  for (int c = 0; c < 3; c++) q[c] = psnr_from_masked_ssd(err[c] > 0 ? err[c] : 1.0, nnm, bit_depth, mask_bit_depth);
  return weighted_average4(q, wavg);
}

double wspsnr_from_rows_masked(const uint64_t* rows444, int32_t height, const double* ws, double correction,
                               int32_t nnm, int32_t bit_depth, int32_t mask_bit_depth) {
  double s = 0.0, k = 0.0;
  for (int32_t y = 0; y < height; y++) neumaier(s, k, (double)rows444[y] * ws[y]);
  return psnr_from_masked_ssd((s + k) * correction, nnm, bit_depth, mask_bit_depth);
}

double wspsnr_from_rows(const uint64_t* rows444, int32_t height, const double* ws, double correction,
                        int32_t area, int32_t bit_depth) {
  double s = 0.0, k = 0.0;  // xKBNS::acc from zero
  for (int32_t y = 0; y < height; y++) neumaier(s, k, (double)rows444[y] * ws[y]);
  return psnr_from_ssd((s + k) * correction, area, bit_depth);
}

double pic_value(const double v[3], const int32_t wavg[4]) {
  const double inv = 1.0 / (double)(wavg[0] + wavg[1] + wavg[2] + wavg[3]);
  return (v[0] * wavg[0] + v[1] * wavg[1] + v[2] * wavg[2]) * inv;
}

double weighted_average4(const double v[3], const int32_t wavg[4]) {
  const double inv = 1.0 / (double)(wavg[0] + wavg[1] + wavg[2] + wavg[3]);
  const double sum = ((v[0] * (double)wavg[0] + v[1] * (double)wavg[1]) + v[2] * (double)wavg[2]) + 0.0 * (double)wavg[3];
  return sum * inv;
}

int32_t gcd_max_diff(float coef, int32_t bit_depth) {
  const float prod = coef * (float)((1 << bit_depth) - 1);
  return (int32_t)std::nearbyint(prod);
}

double min2(double a, double b) { return (b < a) ? b : a; }

bool colour_matrix(int32_t space, bool to_rgb, int32_t m[9]) {
  static const double kLuma[4][3] = {
      {0.299, 0.587, 0.114}, {0.2126, 0.7152, 0.0722}, {0.212, 0.701, 0.087}, {0.2627, 0.678, 0.0593}};
  if (space < 0 || space > 3) return false;
  const double kr = kLuma[space][0], kg = kLuma[space][1], kb = kLuma[space][2];
  double f[9];
  if (!to_rgb) {
    const double cb = 2.0 * (1.0 - kb), cr = 2.0 * (1.0 - kr);
    const double fwd[9] = {kr, kg, kb, -kr / cb, -kg / cb, 0.5, 0.5, -kg / cr, -kb / cr};
    for (int i = 0; i < 9; i++) f[i] = fwd[i];
  } else {
    const double inv[9] = {1.0, 0.0, 2.0 * (1.0 - kr),
                           1.0, -2.0 * kb * (1.0 - kb) / kg, -2.0 * kr * (1.0 - kr) / kg,
                           1.0, 2.0 * (1.0 - kb), 0.0};
    for (int i = 0; i < 9; i++) f[i] = inv[i];
  }
  for (int i = 0; i < 9; i++) {
    const double v = 65536.0 * f[i];
    m[i] = v >= 0.0 ? (int32_t)(v + 0.5) : -(int32_t)(-v + 0.5);
  }
  return true;
}

void ssim_window(double out[100]) {
  double g[11];
  for (int i = 0; i < 11; i++) g[i] = std::exp(-((double)((i - 5) * (i - 5))) / (2.0 * 1.5 * 1.5));
  double m[121], tot = 0.0;
  for (int i = 0; i < 11; i++)
    for (int j = 0; j < 11; j++) m[i * 11 + j] = g[i] * g[j];
  for (int i = 0; i < 121; i++) tot += m[i];
  for (int dy = 0; dy < 10; dy++)
    for (int dx = 0; dx < 10; dx++) {
      char buf[64];
      std::snprintf(buf, sizeof buf, "%.10f", m[dy * 11 + dx] / tot);
      out[dy * 10 + dx] = (double)std::strtof(buf, nullptr);
    }
}

double ssim_from_rows(const double* rows, int32_t h, int32_t w) {
  return kbns(rows, (size_t)(h > 0 ? h : 0)) / (double)((int64_t)w * (int64_t)h);
}

double msssim_combine(const double scores[5]) {
  static const double wts[5] = {0.0448, 0.2856, 0.3001, 0.2363, 0.1333};  // xStructSimConsts.h:27
  double prod = 1.0;
  for (int i = 0; i < 5; i++) prod = prod * std::pow(scores[i], wts[i]);
  return prod;
}

}  // namespace fin
}  // namespace qmiv
