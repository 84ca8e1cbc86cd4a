/*
  Plain-C restatement of the reference's per-frame metric engine (the oracle).

  TEST INFRASTRUCTURE ONLY. Loaded by the tests, the smoke check and the
  CPU-baseline leg of the benchmark as the checker; never part of the product.
  Every function cites the reference file:line whose semantics it restates.

  Pinning (each function is checked by at least one of these, see tests/):
    - golden vectors produced by the reference itself (oracle/_ref, fixtures in tests/golden):
      unpack, gcd, cps_rows (both row semantics), ssd_rows, ws_init, psnr, qual_from_rows,
      ssim_plane, msssim_plane, scp, frame;
    - the published example values (example/README.md:46-47);
    - closed forms: identical pictures, constant offsets, SSIM interior fraction,
      the Gaussian window's regeneration, KBNS exactness (Neumaier 1974);
    - brute-force comparisons on tiny inputs (argmin tie order).
  Built with -ffp-contract=off so double expressions round like the reference's
  ISO C++ build.
*/
#include "qmiv_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXV(bd) ((1 << (bd)) - 1)

static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : v > hi ? hi : v; }
/* std::min / std::max argument order (xCommonDefCORE.h:93-94). */
static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline int32_t imin(int32_t a, int32_t b) { return (b < a) ? b : a; }
static inline int32_t imax(int32_t a, int32_t b) { return (a < b) ? b : a; }

void orc_default_params(orc_params* p, int32_t width, int32_t height, int32_t bit_depth, int32_t chroma_format) {
  /* Defaults: xCorrespPixelShift.h:28-45, xGlobClrDiff.h:17-24, xAppQMIV.cpp (ilp on -> SSE rows). */
  memset(p, 0, sizeof(*p));
  p->width = width;
  p->height = height;
  p->bit_depth = bit_depth;
  p->chroma_format = chroma_format;
  p->search_range = 2;
  p->cmp_weights_search[0] = 4; p->cmp_weights_search[1] = 1; p->cmp_weights_search[2] = 1;
  p->cmp_weights_average[0] = 4; p->cmp_weights_average[1] = 1; p->cmp_weights_average[2] = 1;
  p->unnoticeable_coef[0] = p->unnoticeable_coef[1] = p->unnoticeable_coef[2] = 0.01f;
  p->lon_range_deg = 360;
  p->lat_range_deg = 180;
  p->cps_int32_rows = 1;
  p->metrics = ORC_PSNR | ORC_WSPSNR | ORC_IVPSNR;
}

/* ---------------------------------------------------------------------------------------------- */

int orc_unpack(const void* packed, int32_t w, int32_t h, int32_t bd, int32_t cf, uint16_t* y, uint16_t* cb,
               uint16_t* cr) {
  /* File layout of xSeq::create (xSeq.cpp:408-428): luma, then the chroma planes, each
     (luma bytes >> 2) for 4:2:0 and (luma bytes >> 1) for 4:2:2; chroma rows of w >> 1 samples
     (xSeq.cpp:200,216), so for an odd width the last column reads the next row's first sample. A
     byte past the frame reads as 0. 4:0:0 has no chroma planes; they are set to the mid value. */
  if (w <= 0 || h <= 0 || bd < 8 || bd > 16) return -1;
  const size_t n = (size_t)w * (size_t)h;
  const size_t bps = bd <= 8 ? 1 : 2;
  const size_t cmp = n * bps;
  size_t cr_byte, frame;
  if (cf == 420) {
    cr_byte = cmp + (cmp >> 2);
    frame = cmp + (cmp >> 1);
  } else if (cf == 422) {
    cr_byte = cmp + (cmp >> 1);
    frame = cmp << 1;
  } else if (cf == 444) {
    cr_byte = 2 * cmp;
    frame = 3 * cmp;
  } else if (cf == 400) {
    cr_byte = cmp;
    frame = cmp;
  } else {
    return -1;
  }
  const uint8_t* b = (const uint8_t*)packed;
  uint16_t* dst[3] = {y, cb, cr};
  for (int c = 0; c < 3; c++) {
    const size_t base = c == 0 ? 0 : (c == 1 ? cmp : cr_byte);
    for (int32_t yy = 0; yy < h; yy++) {
      for (int32_t xx = 0; xx < w; xx++) {
        uint16_t v;
        if (c > 0 && cf == 400) {
          v = (uint16_t)(1u << (bd - 1));
        } else {
          size_t idx;
          if (c == 0 || cf == 444) idx = (size_t)yy * w + xx;
          else if (cf == 422) idx = (size_t)yy * (size_t)(w >> 1) + (size_t)(xx >> 1); /* UpsampleH */
          else idx = (size_t)(yy >> 1) * (size_t)(w >> 1) + (size_t)(xx >> 1);          /* UpsampleHV, xPixelOpsSTD.cpp:30-49 */
          const size_t o = base + idx * bps;
          v = o < frame ? b[o] : 0;
          if (bps == 2) v = (uint16_t)(v | (o + 1 < frame ? (uint16_t)b[o + 1] << 8 : 0));
        }
        dst[c][(size_t)yy * w + xx] = v;
      }
    }
  }
  return 0;
}

/* xPixelOpsSTD::ClipToRange (xPixelOpsSTD.cpp:198-209), the concealment of xPicP::conceal. */
static void clip_to_range(uint16_t* p, size_t n, int32_t bd) {
  const uint16_t mx = (uint16_t)ORC_MAXV(bd);
  for (size_t i = 0; i < n; i++)
    if (p[i] > mx) p[i] = mx;
}

int orc_in_range(const uint16_t* p, int32_t n, int32_t bd) {
  if (bd == 16) return 1;
  const int32_t mx = ORC_MAXV(bd);
  for (int32_t i = 0; i < n; i++)
    if (p[i] > mx) return 0;
  return 1;
}

/* ---------------------------------------------------------------------------------------------- */

void orc_gcd(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd,
             const float ucoef[4], int32_t g[4]) {
  const int32_t mx = ORC_MAXV(bd);
  const double area = (double)w * (double)h; /* Ref->getArea() as flt64 */
  for (int c = 0; c < 4; c++) {
    /* MaxDiff = xRoundFltToInt32(CmpUnntcbCoef * (flt32)MaxValue): fp32 product, half-even. */
    const float prod = ucoef[c] * (float)mx;
    const int32_t maxdiff = (int32_t)rintf(prod);
    int32_t sd = 0;
    if (c < 3) {
      /* xDistortionSTD::CalcSD (xDistortionSTD.cpp:18-28): int32 accumulation that wraps. */
      uint32_t acc = 0;
      const size_t n = (size_t)w * (size_t)h;
      for (size_t i = 0; i < n; i++) acc += (uint32_t)((int32_t)p1[c][i] - (int32_t)p0[c][i]);
      sd = (int32_t)acc;
    }
    const double avg = (double)(int64_t)sd / area;
    int32_t r = (int32_t)rint(avg); /* _mm_cvtsd_si32 under round-to-nearest-even */
    g[c] = imax(-maxdiff, imin(r, maxdiff)); /* xClip = xMax(min, xMin(x, max)) */
  }
}

/* ---------------------------------------------------------------------------------------------- */

void orc_cps_rows(const uint16_t* const x[3], const uint16_t* const y[3], int32_t w, int32_t h, const int32_t g[4],
                  int32_t B, const int32_t wts[4], int32_t int32_rows, uint64_t* rows, int32_t* best) {
  for (int32_t py = 0; py < h; py++) {
    uint32_t row32[4] = {0, 0, 0, 0};
    uint64_t row64[4] = {0, 0, 0, 0};
    for (int32_t px = 0; px < w; px++) {
      const size_t o = (size_t)py * w + px;
      /* CurrTstValue = (int32V4)TstPel + GlobalColorShift; the interleaved picture's 4th lane is 0. */
      int32_t t[4];
      for (int c = 0; c < 3; c++) t[c] = (int32_t)x[c][o] + g[c];
      t[3] = g[3];
      int32_t best_err = 0x7fffffff; /* std::numeric_limits<int32>::max() */
      uint32_t best_dist[4] = {0, 0, 0, 0};
      int32_t best_idx = -1, idx = 0;
      for (int32_t dy = -B; dy <= B; dy++) {
        const int32_t qy = clampi(py + dy, 0, h - 1);
        for (int32_t dx = -B; dx <= B; dx++, idx++) {
          const int32_t qx = clampi(px + dx, 0, w - 1);
          const size_t q = (size_t)qy * w + qx;
          uint32_t dist[4];
          uint32_t err = 0;
          for (int c = 0; c < 4; c++) {
            const int32_t r = c < 3 ? (int32_t)y[c][q] : 0;
            const uint32_t d = (uint32_t)(t[c] - r);
            dist[c] = d * d;                     /* _mm_mullo_epi32(Diff, Diff) */
            err += dist[c] * (uint32_t)wts[c];   /* _mm_mullo_epi32 + two _mm_hadd_epi32 */
          }
          if ((int32_t)err < best_err) { /* strict <: first candidate in raster order wins ties */
            best_err = (int32_t)err;
            for (int c = 0; c < 4; c++) best_dist[c] = dist[c];
            best_idx = idx;
          }
        }
      }
      if (best) best[o] = best_idx;
      for (int c = 0; c < 4; c++) {
        row32[c] += best_dist[c];                       /* _mm_add_epi32(RowDistV, BestDist) */
        row64[c] += (uint64_t)(int64_t)(int32_t)best_dist[c]; /* (uint64V4)Dist, STD path */
      }
    }
    for (int c = 0; c < 3; c++)
      rows[(size_t)c * h + py] = int32_rows ? (uint64_t)(int64_t)(int32_t)row32[c] /* (uint64V4)int32V4 */
                                            : row64[c];
  }
}

void orc_scp(const uint16_t* const y[3], const uint16_t* const x[3], int32_t w, int32_t h, int32_t bd,
             const int32_t g[4], int32_t B, const int32_t wts[4], uint16_t* const dst[3]) {
  const int32_t mx = ORC_MAXV(bd);
  for (int32_t py = 0; py < h; py++) {
    for (int32_t px = 0; px < w; px++) {
      const size_t o = (size_t)py * w + px;
      int32_t t[3];
      for (int c = 0; c < 3; c++) t[c] = (int32_t)x[c][o] + g[c];
      /* xCorrespPixelShift::xFindBestPixelWithinBlock (xCorrespPixelShift.cpp:48-85), planar, 3 components */
      int32_t best_err = 0x7fffffff;
      size_t best_q = o;
      for (int32_t dy = -B; dy <= B; dy++) {
        const int32_t qy = clampi(py + dy, 0, h - 1);
        for (int32_t dx = -B; dx <= B; dx++) {
          const int32_t qx = clampi(px + dx, 0, w - 1);
          const size_t q = (size_t)qy * w + qx;
          uint32_t err = 0;
          for (int c = 0; c < 3; c++) {
            const uint32_t d = (uint32_t)(t[c] - (int32_t)y[c][q]);
            err += (d * d) * (uint32_t)wts[c];
          }
          if ((int32_t)err < best_err) { best_err = (int32_t)err; best_q = q; }
        }
      }
      for (int c = 0; c < 3; c++) {
        const int32_t v = (int32_t)y[c][best_q] - g[c];
        dst[c][o] = (uint16_t)imax(0, imin(v, mx)); /* xClipU */
      }
    }
  }
}

/* ---------------------------------------------------------------------------------------------- */

void orc_ssd_rows(const uint16_t* a, const uint16_t* b, int32_t w, int32_t h, uint64_t* rows) {
  for (int32_t yy = 0; yy < h; yy++) {
    uint64_t s = 0;
    for (int32_t xx = 0; xx < w; xx++) {
      const int32_t d = (int32_t)a[(size_t)yy * w + xx] - (int32_t)b[(size_t)yy * w + xx];
      s += (uint64_t)(d * d);
    }
    rows[yy] = s;
  }
}

double orc_psnr_from_ssd(double ssd, int32_t area, int32_t bd) {
  const uint64_t num_points = (uint64_t)(int64_t)area;
  const uint64_t max_value = (uint64_t)ORC_MAXV(bd);
  const uint64_t mx = num_points * (max_value * max_value);
  return ssd > 0 ? 10.0 * log10((double)mx / ssd) : 1.7976931348623157e308; /* flt64_max */
}

/* Neumaier's compensated summation step (xMathUtils.h:23-29). */
static inline void kbns_acc(double* s, double* c, double v) {
  const double t = *s + v;
  if (fabs(*s) >= fabs(v)) *c += ((*s - t) + v);
  else *c += ((v - t) + *s);
  *s = t;
}

double orc_kbns(const double* x, size_t n) {
  if (n == 0) return 0.0;
  double s = x[0], c = 0.0;
  for (size_t i = 1; i < n; i++) kbns_acc(&s, &c, x[i]);
  return s + c;
}

void orc_ws_init(int32_t h, int32_t lat_range_deg, double* weights, double* correction) {
  const double pi = 3.14159265358979323846;
  const double eh = 180.0 * (double)h / (double)lat_range_deg;
  const double off = (eh - h) / 2.0;
  for (int32_t i = 0; i < h; i++) weights[i] = cos((i + off - (eh / 2 - 0.5)) * pi / eh);
  *correction = h / orc_kbns(weights, (size_t)h);
}

double orc_qual_from_rows(const uint64_t* rows, int32_t h, const double* ws_weights, int32_t area, int32_t bd,
                          const int32_t wavg[4]) {
  double err[3];
  for (int c = 0; c < 3; c++) {
    if (ws_weights) {
      /* xKBNS4 over rows of (flt64V4)RowDists * weight: lane-wise acc starting at s = c = 0 */
      double s = 0.0, k = 0.0;
      for (int32_t yy = 0; yy < h; yy++) kbns_acc(&s, &k, (double)rows[(size_t)c * h + yy] * ws_weights[yy]);
      err[c] = s + k;
    } else {
      uint64_t s = 0; /* std::accumulate in uint64 */
      for (int32_t yy = 0; yy < h; yy++) s += rows[(size_t)c * h + yy];
      err[c] = (double)s;
    }
  }
  double q[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++) q[c] = orc_psnr_from_ssd(err[c] > 0 ? err[c] : 1.0, area, bd);
  const int32_t sumw = wavg[0] + wavg[1] + wavg[2] + wavg[3];
  const double inv = 1.0 / (double)sumw;
  /* (CmpQuality * (flt64V4)Weights).getSum(): ((a + b) + c) + d (xVec.h:308) */
  const double sum = ((q[0] * (double)wavg[0] + q[1] * (double)wavg[1]) + q[2] * (double)wavg[2]) + q[3] * (double)wavg[3];
  return sum * inv;
}

/* ---------------------------------------------------------------------------------------------- */

void orc_ssim_filter(float f[121]) {
  /* Normalised 11x11 Gaussian, sigma 1.5, each tap rounded to 10 decimal places and stored as fp32. */
  double g[11], tot = 0.0;
  for (int i = 0; i < 11; i++) g[i] = exp(-((double)((i - 5) * (i - 5))) / (2.0 * 1.5 * 1.5));
  double m[121];
  for (int i = 0; i < 11; i++)
    for (int j = 0; j < 11; j++) { m[i * 11 + j] = g[i] * g[j]; }
  for (int i = 0; i < 121; i++) tot += m[i];
  for (int i = 0; i < 121; i++) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.10f", m[i] / tot);
    f[i] = strtof(buf, NULL);
  }
}

double orc_ssim_plane(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t bd, int32_t calc_l,
                      double* row_sums) {
  float filt[121];
  orc_ssim_filter(filt);
  const double maxv = (double)ORC_MAXV(bd);
  const double k1 = 0.01 * maxv, k2 = 0.03 * maxv;
  const double C1 = k1 * k1, C2 = k2 * k2;
  const int R = 5; /* c_FilterRange; taps dy, dx in [-5, 4] */
  double* rs = (double*)calloc((size_t)(h > 0 ? h : 1), sizeof(double));
  for (int32_t yy = R; yy < h - R; yy++) {
    double row = 0.0;
    for (int32_t xx = R; xx < w - R; xx++) {
      double sR = 0, sT = 0, sR2 = 0, sT2 = 0, sRT = 0;
      for (int dy = -R; dy < R; dy++) {
        for (int dx = -R; dx < R; dx++) {
          const double Rv = ref[(size_t)(yy + dy) * w + (xx + dx)];
          const double Tv = tst[(size_t)(yy + dy) * w + (xx + dx)];
          const double C = (double)filt[(dy + R) * 11 + (dx + R)];
          sR += Rv * C;
          sT += Tv * C;
          sR2 += (Rv * Rv) * C;
          sT2 += (Tv * Tv) * C;
          sRT += (Rv * Tv) * C;
        }
      }
      const double aR = sR, aT = sT;
      const double vR = sR2 - aR * aR;
      const double vT = sT2 - aT * aT;
      const double cov = sRT - aR * aT;
      double pel;
      if (calc_l) {
        const double L = (2 * aR * aT + C1) / (aR * aR + aT * aT + C1);
        const double CS = (2 * cov + C2) / (vR + vT + C2);
        pel = L * CS;
      } else {
        pel = (2 * cov + C2) / (vR + vT + C2);
      }
      row += pel;
    }
    rs[yy] = row;
  }
  if (row_sums) memcpy(row_sums, rs, sizeof(double) * (size_t)h);
  const double sum = orc_kbns(rs, (size_t)h);
  free(rs);
  return sum / (double)((int64_t)w * (int64_t)h);
}

void orc_downsample(uint16_t* dst, const uint16_t* src, int32_t dw, int32_t dh, int32_t src_stride) {
  for (int32_t yy = 0; yy < dh; yy++) {
    const uint16_t* l0 = src + (size_t)(2 * yy) * src_stride;
    const uint16_t* l1 = l0 + src_stride;
    for (int32_t xx = 0; xx < dw; xx++) {
      const int32_t sx = xx << 1;
      dst[(size_t)yy * dw + xx] = (uint16_t)(((int32_t)l0[sx] + l0[sx + 1] + l1[sx] + l1[sx + 1] + 2) >> 2);
    }
  }
}

double orc_msssim_plane(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t bd, double scales[5]) {
  static const double wts[5] = {0.0448, 0.2856, 0.3001, 0.2363, 0.1333}; /* c_MultiScaleWghts */
  double s[5];
  s[0] = orc_ssim_plane(tst, ref, w, h, bd, 0, NULL);
  const uint16_t* pt = tst;
  const uint16_t* pr = ref;
  int32_t cw = w, ch = h;
  uint16_t* bufT = NULL;
  uint16_t* bufR = NULL;
  for (int i = 1; i < 5; i++) {
    const int32_t nw = cw >> 1, nh = ch >> 1;
    uint16_t* nt = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(nw > 0 ? nw : 1) * (size_t)(nh > 0 ? nh : 1));
    uint16_t* nr = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)(nw > 0 ? nw : 1) * (size_t)(nh > 0 ? nh : 1));
    orc_downsample(nt, pt, nw, nh, cw);
    orc_downsample(nr, pr, nw, nh, cw);
    free(bufT);
    free(bufR);
    bufT = nt;
    bufR = nr;
    pt = nt;
    pr = nr;
    cw = nw;
    ch = nh;
    s[i] = orc_ssim_plane(pt, pr, cw, ch, bd, i != 4, NULL);
  }
  free(bufT);
  free(bufR);
  double prod = 1.0;
  for (int i = 0; i < 5; i++) prod = prod * pow(s[i], wts[i]);
  if (scales) memcpy(scales, s, sizeof s);
  return prod;
}

/* ---------------------------------------------------------------------------------------------- */

static double pic_value(const double v[3], const int32_t wavg[4]) {
  /* xMetricStat::setPerCmpMeric (apps/QMIV/src/xAppQMIV.h:245-251) */
  const double inv = 1.0 / (double)(wavg[0] + wavg[1] + wavg[2] + wavg[3]);
  return (v[0] * wavg[0] + v[1] * wavg[1] + v[2] * wavg[2]) * inv;
}

int orc_frame(const void* p0, const void* p1, const orc_params* prm, orc_frame_result* out, uint16_t* scp0,
              uint16_t* scp1) {
  const int32_t w = prm->width, h = prm->height, bd = prm->bit_depth;
  const size_t n = (size_t)w * (size_t)h;
  const int32_t area = w * h;
  const int m = prm->metrics;
  memset(out, 0, sizeof(*out));
  uint16_t* buf = (uint16_t*)malloc(sizeof(uint16_t) * n * 6);
  if (!buf) return -2;
  const uint16_t* P0[3] = {buf, buf + n, buf + 2 * n};
  const uint16_t* P1[3] = {buf + 3 * n, buf + 4 * n, buf + 5 * n};
  if (orc_unpack(p0, w, h, bd, prm->chroma_format, buf, buf + n, buf + 2 * n) ||
      orc_unpack(p1, w, h, bd, prm->chroma_format, buf + 3 * n, buf + 4 * n, buf + 5 * n)) {
    free(buf);
    return -1;
  }
  out->in_range[0] = orc_in_range(buf, (int32_t)(3 * n), bd);
  out->in_range[1] = orc_in_range(buf + 3 * n, (int32_t)(3 * n), bd);
  /* InvalidPelActn CNCL (xAppQMIV.cpp:753-762): the pictures that fail the check are clipped and the
     frame goes on with them; otherwise (WARN, SKIP) with the samples as read. */
  if (prm->conceal) {
    if (!out->in_range[0]) clip_to_range(buf, 3 * n, bd);
    if (!out->in_range[1]) clip_to_range(buf + 3 * n, 3 * n, bd);
  }
  for (int c = 0; c < 3; c++) out->exact[c] = memcmp(P0[c], P1[c], n * sizeof(uint16_t)) == 0;

  const int calc_ivs = (m & (ORC_IVPSNR | ORC_IVSSIM)) != 0;
  if (calc_ivs) orc_gcd(P0, P1, w, h, bd, prm->unnoticeable_coef, out->gcd);
  const int32_t* g = out->gcd;
  const int32_t ng[4] = {-g[0], -g[1], -g[2], -g[3]};

  uint16_t* scp = NULL;
  if (m & ORC_IVSSIM) {
    scp = (uint16_t*)malloc(sizeof(uint16_t) * n * 6);
    uint16_t* d1[3] = {scp + 3 * n, scp + 4 * n, scp + 5 * n};
    uint16_t* d0[3] = {scp, scp + n, scp + 2 * n};
    /* GenShftCompPics(&SCP1, &SCP0, &P1, &P0, GCD) (xAppQMIV.cpp:703, xShftCompPic.cpp:12-21) */
    orc_scp(P1, P0, w, h, bd, g, prm->search_range, prm->cmp_weights_search, d1);
    orc_scp(P0, P1, w, h, bd, ng, prm->search_range, prm->cmp_weights_search, d0);
    if (scp0) memcpy(scp0, scp, sizeof(uint16_t) * 3 * n);
    if (scp1) memcpy(scp1, scp + 3 * n, sizeof(uint16_t) * 3 * n);
  }

  const double fake = orc_psnr_from_ssd(1.0, area, bd);
  double* ws = NULL;
  double corr = 1.0;
  if (prm->erp) {
    ws = (double*)malloc(sizeof(double) * (size_t)h);
    orc_ws_init(h, prm->lat_range_deg, ws, &corr);
  }
  uint64_t* rows = (uint64_t*)malloc(sizeof(uint64_t) * 3 * (size_t)h);

  if (m & ORC_PSNR) {
    for (int c = 0; c < 3; c++) {
      orc_ssd_rows(P1[c], P0[c], w, h, rows); /* CalcSSD(RefPtr, TstPtr) per row */
      uint64_t s = 0;
      for (int32_t yy = 0; yy < h; yy++) s += rows[yy];
      out->psnr[c] = out->exact[c] ? fake : orc_psnr_from_ssd((double)s, area, bd);
    }
    out->psnr_pic = pic_value(out->psnr, prm->cmp_weights_average);
  }
  if (m & ORC_WSPSNR) {
    for (int c = 0; c < 3; c++) {
      double v;
      if (!ws) {
        uint64_t s = 0;
        orc_ssd_rows(P1[c], P0[c], w, h, rows);
        for (int32_t yy = 0; yy < h; yy++) s += rows[yy];
        v = orc_psnr_from_ssd((double)s, area, bd);
      } else {
        orc_ssd_rows(P1[c], P0[c], w, h, rows);
        double s = 0.0, k = 0.0;
        for (int32_t yy = 0; yy < h; yy++) kbns_acc(&s, &k, (double)rows[yy] * ws[yy]);
        v = orc_psnr_from_ssd((s + k) * corr, area, bd);
      }
      out->wspsnr[c] = out->exact[c] ? fake : v;
    }
    out->wspsnr_pic = pic_value(out->wspsnr, prm->cmp_weights_average);
  }
  if (m & ORC_IVPSNR) {
    /* calcPicIVPSNR(I0, I1, GCD): R2T = Q(P0 + g in P1), T2R = Q(P1 - g in P0) (xIVPSNR.cpp:45-57) */
    for (int d = 0; d < 2; d++) {
      orc_cps_rows(d == 0 ? P0 : P1, d == 0 ? P1 : P0, w, h, d == 0 ? g : ng, prm->search_range,
                   prm->cmp_weights_search, prm->cps_int32_rows, rows, NULL);
      for (int c = 0; c < 3; c++) {
        uint64_t s = 0;
        for (int32_t yy = 0; yy < h; yy++) s += rows[(size_t)c * h + yy];
        out->cps_sum[d][c] = s;
      }
      const double q = orc_qual_from_rows(rows, h, ws, area, bd, prm->cmp_weights_average);
      if (d == 0) out->r2t = q;
      else out->t2r = q;
    }
    out->ivpsnr = dmin(out->r2t, out->t2r);
  }
  if (m & ORC_SSIM) {
    for (int c = 0; c < 3; c++) out->ssim[c] = orc_ssim_plane(P0[c], P1[c], w, h, bd, 1, NULL);
    out->ssim_pic = pic_value(out->ssim, prm->cmp_weights_average);
  }
  if (m & ORC_MSSSIM) {
    for (int c = 0; c < 3; c++) {
      double sc[5];
      out->msssim[c] = orc_msssim_plane(P0[c], P1[c], w, h, bd, sc);
      for (int i = 0; i < 5; i++) out->msssim_scales[i][c] = sc[i];
    }
    out->msssim_pic = pic_value(out->msssim, prm->cmp_weights_average);
  }
  if (m & ORC_IVSSIM) {
    /* calcPicIVSSIM(P0, P1, SCP0, SCP1): T2R = SSIM(P0, SCP1), R2T = SSIM(P1, SCP0) (xSSIM.cpp:192-212) */
    const int32_t* wa = prm->cmp_weights_average;
    const double inv = 1.0 / (double)(wa[0] + wa[1] + wa[2] + wa[3]);
    double st[4] = {0, 0, 0, 0}, sr[4] = {0, 0, 0, 0};
    for (int c = 0; c < 3; c++) {
      st[c] = orc_ssim_plane(P0[c], scp + 3 * n + (size_t)c * n, w, h, bd, 1, NULL);
      sr[c] = orc_ssim_plane(P1[c], scp + (size_t)c * n, w, h, bd, 1, NULL);
    }
    const double t2r = (((st[0] * wa[0] + st[1] * wa[1]) + st[2] * wa[2]) + st[3] * wa[3]) * inv;
    const double r2t = (((sr[0] * wa[0] + sr[1] * wa[1]) + sr[2] * wa[2]) + sr[3] * wa[3]) * inv;
    out->ivssim_t2r = t2r;
    out->ivssim_r2t = r2t;
    out->ivssim = dmin(t2r, r2t);
  }
  free(rows);
  free(ws);
  free(scp);
  free(buf);
  return 0;
}

double orc_frame_average(const double* v, size_t n) { return n ? orc_kbns(v, n) / (double)n : 0.0; }

/* ---------------------------------------------------------------------------------------------
   Masked mode
   --------------------------------------------------------------------------------------------- */

/* _mm_cvtsd_si32 under round-to-nearest-even (xCommonDefCORE.h:130): NaN and results outside
   int32 give the "integer indefinite" value INT32_MIN. */
static int32_t cvtsd_si32(double v) {
  const double r = rint(v);
  if (!(r >= -2147483648.0 && r <= 2147483647.0)) return (int32_t)0x80000000u;
  return (int32_t)r;
}

void orc_gcd_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* mask, int32_t w,
                    int32_t h, int32_t bd, int32_t bdm, int32_t nnm, const float ucoef[4], int32_t g[4]) {
  const int32_t mx = ORC_MAXV(bd);
  const size_t n = (size_t)w * (size_t)h;
  /* (flt64)((int64)NumNonMasked * (int64)MaxMsk) */
  const double denom = (double)((int64_t)nnm * (int64_t)ORC_MAXV(bdm));
  for (int c = 0; c < 4; c++) {
    const int32_t maxdiff = (int32_t)rintf(ucoef[c] * (float)mx);
    int64_t sd = 0;
    if (c < 3) {
      /* xDistortionSTD::CalcWeightedSD (xDistortionSTD.cpp:69-80): each (Tst - Ref) * Msk is an
         int32 product (wrapping), accumulated in int64. Tst = P1, Ref = P0 here. */
      for (size_t i = 0; i < n; i++) {
        const int32_t d = (int32_t)p1[c][i] - (int32_t)p0[c][i];
        sd += (int64_t)(int32_t)((uint32_t)d * (uint32_t)mask[i]);
      }
    }
    const int32_t r = cvtsd_si32((double)sd / denom);
    g[c] = imax(-maxdiff, imin(r, maxdiff));
  }
}

void orc_cps_rows_masked(const uint16_t* const x[3], const uint16_t* const y[3], const uint16_t* mask, int32_t w,
                         int32_t h, const int32_t g[4], int32_t B, const int32_t wts[4], uint64_t* rows) {
  for (int32_t py = 0; py < h; py++) {
    uint64_t row[3] = {0, 0, 0};
    for (int32_t px = 0; px < w; px++) {
      const size_t o = (size_t)py * w + px;
      const uint32_t m = mask[o];
      if (m == 0) continue; /* masked pel: skipped */
      int32_t t[4];
      for (int c = 0; c < 3; c++) t[c] = (int32_t)x[c][o] + g[c];
      t[3] = g[3];
      int32_t best_err = 0x7fffffff;
      uint32_t best_dist[3] = {0, 0, 0};
      for (int32_t dy = -B; dy <= B; dy++) {
        const int32_t qy = clampi(py + dy, 0, h - 1);
        for (int32_t dx = -B; dx <= B; dx++) {
          const int32_t qx = clampi(px + dx, 0, w - 1);
          const size_t q = (size_t)qy * w + qx;
          if (mask[q] == 0) continue; /* masked candidate (the margin replicates the mask's edge) */
          uint32_t dist[4], err = 0;
          for (int c = 0; c < 4; c++) {
            const int32_t r = c < 3 ? (int32_t)y[c][q] : 0;
            const uint32_t d = (uint32_t)(t[c] - r);
            dist[c] = d * d;                    /* (Diff).getVecPow2() in int32 */
            err += dist[c] * (uint32_t)wts[c];  /* (Dist * CmpWeights).getSum() */
          }
          if ((int32_t)err < best_err) {
            best_err = (int32_t)err;
            for (int c = 0; c < 3; c++) best_dist[c] = dist[c];
          }
        }
      }
      /* RowDist += ((uint64V4)Dist) * CurrMskValue: sign-extending conversion, uint64 product */
      for (int c = 0; c < 3; c++) row[c] += (uint64_t)(int64_t)(int32_t)best_dist[c] * (uint64_t)m;
    }
    for (int c = 0; c < 3; c++) rows[(size_t)c * h + py] = row[c];
  }
}

double orc_psnr_from_masked_ssd(double ssd, int32_t nnm, int32_t bd, int32_t bdm) {
  /* (int64)NumNonMasked * xPow2(MaxValuePic) * MaxValueMsk, int64 arithmetic (wrapping) */
  const uint64_t mx = (uint64_t)(int64_t)nnm * ((uint64_t)ORC_MAXV(bd) * (uint64_t)ORC_MAXV(bd)) * (uint64_t)ORC_MAXV(bdm);
  return ssd > 0 ? 10.0 * log10((double)(int64_t)mx / ssd) : 1.7976931348623157e308;
}

/* xDistortionSTD::CalcWeightedSSD per row (xDistortionSTD.cpp:81-86): (uint64)xPow2(int32 d) * m. */
static void ssd_rows_masked(const uint16_t* a, const uint16_t* b, const uint16_t* mask, int32_t w, int32_t h,
                            uint64_t* rows) {
  for (int32_t yy = 0; yy < h; yy++) {
    uint64_t s = 0;
    for (int32_t xx = 0; xx < w; xx++) {
      const size_t o = (size_t)yy * w + xx;
      const uint32_t d = (uint32_t)((int32_t)a[o] - (int32_t)b[o]);
      s += (uint64_t)(int64_t)(int32_t)(d * d) * (uint64_t)mask[o];
    }
    rows[yy] = s;
  }
}

int orc_frame_masked(const void* p0, const void* p1, const void* mask, int32_t bdm, int32_t cfm,
                     const orc_params* prm, orc_masked_result* out) {
  const int32_t w = prm->width, h = prm->height, bd = prm->bit_depth;
  const size_t n = (size_t)w * (size_t)h;
  const int32_t area = w * h;
  const int m = prm->metrics;
  memset(out, 0, sizeof(*out));
  if (m & (ORC_SSIM | ORC_MSSSIM | ORC_IVSSIM)) return -3; /* xAppQMIV.cpp:337 */
  uint16_t* buf = (uint16_t*)malloc(sizeof(uint16_t) * n * 9);
  if (!buf) return -2;
  const uint16_t* P0[3] = {buf, buf + n, buf + 2 * n};
  const uint16_t* P1[3] = {buf + 3 * n, buf + 4 * n, buf + 5 * n};
  const uint16_t* M = buf + 6 * n; /* luma plane of the mask (eCmp::LM) */
  if (orc_unpack(p0, w, h, bd, prm->chroma_format, buf, buf + n, buf + 2 * n) ||
      orc_unpack(p1, w, h, bd, prm->chroma_format, buf + 3 * n, buf + 4 * n, buf + 5 * n) ||
      orc_unpack(mask, w, h, bdm, cfm, buf + 6 * n, buf + 7 * n, buf + 8 * n)) {
    free(buf);
    return -1;
  }
  out->in_range[0] = orc_in_range(buf, (int32_t)(3 * n), bd);
  out->in_range[1] = orc_in_range(buf + 3 * n, (int32_t)(3 * n), bd);
  out->in_range[2] = orc_in_range(buf + 6 * n, (int32_t)(3 * n), bdm);
  if (prm->conceal) {
    if (!out->in_range[0]) clip_to_range(buf, 3 * n, bd);
    if (!out->in_range[1]) clip_to_range(buf + 3 * n, 3 * n, bd);
    if (!out->in_range[2]) clip_to_range(buf + 6 * n, 3 * n, bdm);
  }
  for (int c = 0; c < 3; c++) out->exact[c] = memcmp(P0[c], P1[c], n * sizeof(uint16_t)) == 0;
  int32_t nnm = 0; /* xPixelOpsSTD::CountNonZero (xPixelOpsSTD.cpp:273-284) */
  for (size_t i = 0; i < n; i++) nnm += M[i] != 0;
  out->num_non_masked = nnm;

  if (m & ORC_IVPSNR) orc_gcd_masked(P0, P1, M, w, h, bd, bdm, nnm, prm->unnoticeable_coef, out->gcd);
  const int32_t* g = out->gcd;
  const int32_t ng[4] = {-g[0], -g[1], -g[2], -g[3]};

  const double fake = orc_psnr_from_ssd(1.0, area, bd);
  double* ws = NULL;
  double corr = 1.0;
  if (prm->erp) {
    ws = (double*)malloc(sizeof(double) * (size_t)h);
    orc_ws_init(h, prm->lat_range_deg, ws, &corr);
  }
  uint64_t* rows = (uint64_t*)malloc(sizeof(uint64_t) * 3 * (size_t)h);
  double plain[3] = {0, 0, 0};
  if (m & (ORC_PSNR | ORC_WSPSNR)) {
    /* xPSNR::calcPicPSNRM (libs/pmbb-ivqm/src/xPSNR.cpp:42-71,84-136) */
    for (int c = 0; c < 3; c++) {
      ssd_rows_masked(P1[c], P0[c], M, w, h, rows);
      uint64_t s = 0;
      for (int32_t yy = 0; yy < h; yy++) s += rows[yy];
      plain[c] = orc_psnr_from_masked_ssd((double)s, nnm, bd, bdm);
    }
  }
  if (m & ORC_PSNR) {
    for (int c = 0; c < 3; c++) out->psnr[c] = out->exact[c] ? fake : plain[c];
    out->psnr_pic = pic_value(out->psnr, prm->cmp_weights_average);
  }
  if (m & ORC_WSPSNR) {
    /* xWSPSNR::calcPicWSPSNRM (libs/pmbb-ivqm/src/xWSPSNR.cpp:61-97,127-155): plain masked PSNR
       without WS; with WS, KBNS over rows of rowSSD * w[y], times the correction. */
    for (int c = 0; c < 3; c++) {
      double v = plain[c];
      if (ws) {
        ssd_rows_masked(P1[c], P0[c], M, w, h, rows);
        double s = 0.0, k = 0.0;
        for (int32_t yy = 0; yy < h; yy++) kbns_acc(&s, &k, (double)rows[yy] * ws[yy]);
        v = orc_psnr_from_masked_ssd((s + k) * corr, nnm, bd, bdm);
      }
      out->wspsnr[c] = out->exact[c] ? fake : v;
    }
    out->wspsnr_pic = pic_value(out->wspsnr, prm->cmp_weights_average);
  }
  if (m & ORC_IVPSNR) {
    /* xIVPSNRM::calcPicIVPSNRM(I0, I1, Msk, NNM, GCD) (libs/pmbb-ivqm/src/xIVPSNR.cpp:152-209):
       R2T = Q(P1 - g in P0), T2R = Q(P0 + g in P1). */
    const int32_t* wa = prm->cmp_weights_average;
    const double inv = 1.0 / (double)(wa[0] + wa[1] + wa[2] + wa[3]);
    for (int d = 0; d < 2; d++) {
      orc_cps_rows_masked(d == 0 ? P1 : P0, d == 0 ? P0 : P1, M, w, h, d == 0 ? ng : g, prm->search_range,
                          prm->cmp_weights_search, rows);
      double q[4] = {0, 0, 0, 0};
      for (int c = 0; c < 3; c++) {
        uint64_t s = 0;
        for (int32_t yy = 0; yy < h; yy++) s += rows[(size_t)c * h + yy];
        out->cps_sum[d][c] = s;
        double err;
        if (ws) {
          double a = 0.0, k = 0.0;
          for (int32_t yy = 0; yy < h; yy++) kbns_acc(&a, &k, (double)rows[(size_t)c * h + yy] * ws[yy]);
          err = a + k;
        } else {
          err = (double)s;
        }
        q[c] = orc_psnr_from_masked_ssd(err > 0 ? err : 1.0, nnm, bd, bdm);
      }
      const double v = (((q[0] * (double)wa[0] + q[1] * (double)wa[1]) + q[2] * (double)wa[2]) + q[3] * (double)wa[3]) * inv;
      if (d == 0) out->r2t = v;
      else out->t2r = v;
    }
    out->ivpsnr = dmin(out->r2t, out->t2r);
  }
  free(rows);
  free(ws);
  free(buf);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
   Colour space (libs/pmbb-core/src/xColorSpaceSTD.cpp:195-272). Q16 matrices of the four
   standard luma/chroma spaces (index = eClrSpcLC, xCommonDefCORE.h:356-370: BT.601, BT.709,
   SMPTE 240M, BT.2020), derived from the standards' luma weights (Kr, Kg, Kb):
     RGB -> YCbCr: Y = Kr R + Kg G + Kb B, Cb = (B - Y) / (2 (1 - Kb)), Cr = (R - Y) / (2 (1 - Kr))
     YCbCr -> RGB: R = Y + 2 (1 - Kr) Cr, B = Y + 2 (1 - Kb) Cb,
                   G = Y - 2 Kb (1 - Kb) / Kg Cb - 2 Kr (1 - Kr) / Kg Cr
   each coefficient scaled by 65536 and rounded half away from zero.
   --------------------------------------------------------------------------------------------- */
static const double k_luma_weights[4][3] = {
    {0.299, 0.587, 0.114},    /* ITU-R BT.601 / SMPTE 170M */
    {0.2126, 0.7152, 0.0722}, /* ITU-R BT.709 */
    {0.212, 0.701, 0.087},    /* SMPTE 240M */
    {0.2627, 0.678, 0.0593},  /* ITU-R BT.2020 */
};

static int32_t q16(double v) { return v >= 0.0 ? (int32_t)(v * 65536.0 + 0.5) : -(int32_t)(-v * 65536.0 + 0.5); }

int orc_colorspace_coeffs(int32_t space, int32_t to_rgb, int32_t m[9]) {
  if (space < 0 || space > 3) return -1;
  const double kr = k_luma_weights[space][0], kg = k_luma_weights[space][1], kb = k_luma_weights[space][2];
  double f[9];
  if (!to_rgb) {
    f[0] = kr;                        f[1] = kg;                        f[2] = kb;
    f[3] = -kr / (2.0 * (1.0 - kb));  f[4] = -kg / (2.0 * (1.0 - kb));  f[5] = 0.5;
    f[6] = 0.5;                       f[7] = -kg / (2.0 * (1.0 - kr));  f[8] = -kb / (2.0 * (1.0 - kr));
  } else {
    f[0] = 1.0; f[1] = 0.0;                                f[2] = 2.0 * (1.0 - kr);
    f[3] = 1.0; f[4] = -2.0 * kb * (1.0 - kb) / kg;        f[5] = -2.0 * kr * (1.0 - kr) / kg;
    f[6] = 1.0; f[7] = 2.0 * (1.0 - kb);                   f[8] = 0.0;
  }
  for (int i = 0; i < 9; i++) m[i] = q16(f[i]);
  return 0;
}

static int32_t clipu(int32_t v, int32_t maxv) { return v < 0 ? 0 : (v > maxv ? maxv : v); }
/* int32 arithmetic with two's-complement wrap and arithmetic right shift, as the reference's
   int32 expressions compile. */
static int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static int32_t wrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
static int32_t wrap_shl(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }

int orc_colorspace(const uint16_t* src, uint16_t* dst, size_t n, int32_t bd, int32_t space, int32_t to_rgb) {
  int32_t m[9];
  if (orc_colorspace_coeffs(space, to_rgb, m)) return -1;
  const int32_t mid = 1 << (bd - 1), maxv = (1 << bd) - 1, add = 1 << 15;
  const uint16_t *s0 = src, *s1 = src + n, *s2 = src + 2 * n;
  uint16_t *d0 = dst, *d1 = dst + n, *d2 = dst + 2 * n;
  for (size_t i = 0; i < n; i++) {
    if (!to_rgb) {  /* ConvertRGB2YCbCr_I32 (xColorSpaceSTD.cpp:195-235): the 0.5 taps as << 15 */
      const int32_t r = s0[i], g = s1[i], b = s2[i];
      const int32_t ty = wrap_add(wrap_add(wrap_add(wrap_mul(m[0], r), wrap_mul(m[1], g)), wrap_mul(m[2], b)), add) >> 16;
      const int32_t tu = wrap_add(wrap_add(wrap_add(wrap_mul(m[3], r), wrap_mul(m[4], g)), wrap_shl(b, 15)), add) >> 16;
      const int32_t tv = wrap_add(wrap_add(wrap_add(wrap_shl(r, 15), wrap_mul(m[7], g)), wrap_mul(m[8], b)), add) >> 16;
      d0[i] = (uint16_t)clipu(ty, maxv);
      d1[i] = (uint16_t)clipu(tu + mid, maxv);
      d2[i] = (uint16_t)clipu(tv + mid, maxv);
    } else {        /* ConvertYCbCr2RGB_I32 (xColorSpaceSTD.cpp:236-272) */
      const int32_t iy = s0[i], iu = (int32_t)s1[i] - mid, iv = (int32_t)s2[i] - mid;
      const int32_t sy = wrap_add(wrap_shl(iy, 16), add);
      const int32_t r = wrap_add(sy, wrap_mul(m[2], iv)) >> 16;
      const int32_t g = wrap_add(wrap_add(sy, wrap_mul(m[4], iu)), wrap_mul(m[5], iv)) >> 16;
      const int32_t b = wrap_add(sy, wrap_mul(m[7], iu)) >> 16;
      d0[i] = (uint16_t)clipu(r, maxv);
      d1[i] = (uint16_t)clipu(g, maxv);
      d2[i] = (uint16_t)clipu(b, maxv);
    }
  }
  return 0;
}
