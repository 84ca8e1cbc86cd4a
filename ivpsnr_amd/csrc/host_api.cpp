// Host runtime of the metric engine: contexts, device pictures, the per-frame flow and the
// C ABI declared in include/qmiv.h.
#include "../../include/qmiv.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "finalize.h"
#include "host_internal.h"
#include "qmiv_internal.h"

using namespace qmiv;

namespace {

thread_local std::string g_last_error;

qmiv_status fail(qmiv_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

#define QMIV_HIP(expr)                                                                          \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail(QMIV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));              \
  } while (0)

#define QMIV_TRY(expr)                 \
  do {                                 \
    qmiv_status s_ = (expr);           \
    if (s_ != QMIV_OK) return s_;      \
  } while (0)

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Pinned results block of the allocation-free frame entry point, allocated with each context.
constexpr size_t kFixedResultsBytes = (size_t)2 << 20;

}  // namespace

qmiv_status qmiv::host_fail(qmiv_status st, const std::string& msg) { return fail(st, msg); }

// Bytes of one packed frame of a raw file: xSeq::create (libs/pmbb-core/src/xSeq.cpp:408-428), for
// any size (odd 4:2:0 and 4:2:2 widths included); 0 for an invalid format.
size_t qmiv::file_frame_bytes(int32_t w, int32_t h, int32_t bd, int32_t cf) {
  if (w <= 0 || h <= 0 || bd < 8 || bd > 16) return 0;
  const size_t cmp = (size_t)w * h * (bd <= 8 ? 1 : 2);
  switch (cf) {
    case 444: return 3 * cmp;
    case 422: return cmp << 1;
    case 420: return cmp + (cmp >> 1);
    case 400: return cmp;
    default: return 0;
  }
}

qmiv_status qmiv::file_route(const qmiv_pic* pic, int32_t w, int32_t h, int32_t bd, int32_t fcf, bool& unpack) {
  if (!pic) return fail(QMIV_ERR_INVALID_ARG, "null picture");
  if (file_frame_bytes(w, h, bd, fcf) == 0) return fail(QMIV_ERR_INVALID_ARG, "file format must be 400, 420, 422 or 444");
  if (pic->g.w != w || pic->g.h != h || pic->g.bd != bd)
    return fail(QMIV_ERR_INVALID_ARG, "picture size or bit depth differs from the frame's");
  unpack = pic->g.cf != fcf;
  if (unpack && pic->g.cf != 444)
    return fail(QMIV_ERR_INVALID_ARG, "frames of another chroma format unpack into 4:4:4 pictures only");
  return QMIV_OK;
}

namespace {

qmiv_status make_geom(int32_t w, int32_t h, int32_t bd, int32_t cf, Geom& g) {
  if (w <= 0 || h <= 0) return fail(QMIV_ERR_INVALID_ARG, "width and height must be positive");
  if (bd < 8 || bd > 16) return fail(QMIV_ERR_INVALID_ARG, "bit depth must be within 8..16");
  if (cf != 420 && cf != 444 && cf != 400)
    return fail(QMIV_ERR_UNSUPPORTED, "device pictures are 4:2:0, 4:4:4 or 4:0:0 (other file formats unpack into 4:4:4)");
  if (cf == 420 && ((w & 1) || (h & 1)))
    return fail(QMIV_ERR_UNSUPPORTED, "4:2:0 device pictures need even width and height (odd sizes unpack into 4:4:4)");
  if (h > 65535) return fail(QMIV_ERR_UNSUPPORTED, "height above 65535");
  g = Geom{};
  g.w = w;
  g.h = h;
  g.cf = cf;
  g.cs = cf == 420 ? 1 : 0;
  g.cw = w >> g.cs;
  g.ch = h >> g.cs;
  g.bd = bd;
  g.bytes_per_sample = bd <= 8 ? 1 : 2;
  const size_t samples = (size_t)w * h + (cf == 400 ? 0 : 2 * (size_t)g.cw * g.ch);
  const size_t bytes = samples * (size_t)g.bytes_per_sample;
  if (bytes >= (size_t)1 << 31) return fail(QMIV_ERR_UNSUPPORTED, "frame larger than 2 GiB");
  g.frame_bytes = (uint32_t)bytes;
  g.cb_off = cf == 400 ? 0 : (int64_t)w * h;
  g.cr_off = cf == 400 ? 0 : g.cb_off + (int64_t)g.cw * g.ch;
  return QMIV_OK;
}

bool same_geom(const Geom& a, const Geom& b) {
  return a.w == b.w && a.h == b.h && a.bd == b.bd && a.cf == b.cf;
}

bool is444(const Geom& g) { return g.cf == 444; }



// xWeightedSpherically::initWS for this height and latitude range, computed once per context.
const qmiv_ctx::WsTable& ws_table(qmiv_ctx* ctx, int32_t h, int32_t lat) {
  for (const auto& t : ctx->ws_cache)
    if (t.h == h && t.lat == lat) return t;
  qmiv_ctx::WsTable t;
  t.h = h;
  t.lat = lat;
  fin::ws_weights(h, lat, t.w, t.corr);
  ctx->ws_cache.push_back(std::move(t));
  return ctx->ws_cache.back();
}

qmiv_status ensure_scratch(qmiv_ctx* ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return QMIV_OK;
  if (ctx->d_scratch) {
    QMIV_HIP(hipStreamSynchronize(ctx->stream));
    for (hipStream_t a : ctx->aux_stream) QMIV_HIP(hipStreamSynchronize(a));
    QMIV_HIP(hipStreamSynchronize(ctx->copy_stream));
    QMIV_HIP(hipFree(ctx->d_scratch));
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  const size_t sz = align_up(bytes, 1 << 20);
  if (hipMalloc(&ctx->d_scratch, sz) != hipSuccess) return fail(QMIV_ERR_OUT_OF_MEMORY, "device scratch allocation failed");
  ctx->scratch_bytes = sz;
  return QMIV_OK;
}

qmiv_status ensure_results(qmiv_ctx* ctx, size_t bytes) {
  if (ctx->results_bytes >= bytes) return QMIV_OK;
  if (ctx->h_results) {
    QMIV_HIP(hipStreamSynchronize(ctx->stream));
    for (hipStream_t a : ctx->aux_stream) QMIV_HIP(hipStreamSynchronize(a));
    QMIV_HIP(hipStreamSynchronize(ctx->copy_stream));
    QMIV_HIP(hipHostFree(ctx->h_results));
    ctx->h_results = nullptr;
    ctx->results_bytes = 0;
  }
  const size_t sz = align_up(bytes, 1 << 16);
  if (hipHostMalloc(&ctx->h_results, sz, hipHostMallocDefault) != hipSuccess)
    return fail(QMIV_ERR_OUT_OF_MEMORY, "pinned result buffer allocation failed");
  ctx->results_bytes = sz;
  return QMIV_OK;
}

qmiv_status ensure_stage(qmiv_ctx* ctx, size_t bytes) {
  if (ctx->stage_bytes >= bytes) return QMIV_OK;
  if (ctx->d_stage) {
    QMIV_HIP(hipStreamSynchronize(ctx->stream));
    QMIV_HIP(hipFree(ctx->d_stage));
    ctx->d_stage = nullptr;
    ctx->stage_bytes = 0;
  }
  const size_t sz = align_up(bytes, 1 << 20);
  if (hipMalloc(&ctx->d_stage, sz) != hipSuccess) return fail(QMIV_ERR_OUT_OF_MEMORY, "staging buffer allocation failed");
  ctx->stage_bytes = sz;
  return QMIV_OK;
}

qmiv_status ensure_scp(qmiv_ctx* ctx, size_t bytes) {
  if (ctx->scp_bytes >= bytes) return QMIV_OK;
  if (ctx->d_scp) {
    QMIV_HIP(hipStreamSynchronize(ctx->stream));
    for (hipStream_t a : ctx->aux_stream) QMIV_HIP(hipStreamSynchronize(a));
    QMIV_HIP(hipStreamSynchronize(ctx->copy_stream));
    QMIV_HIP(hipFree(ctx->d_scp));
    ctx->d_scp = nullptr;
    ctx->scp_bytes = 0;
  }
  const size_t sz = align_up(bytes, 1 << 20);
  if (hipMalloc(&ctx->d_scp, sz) != hipSuccess) return fail(QMIV_ERR_OUT_OF_MEMORY, "SCP allocation failed");
  ctx->scp_bytes = sz;
  return QMIV_OK;
}

// ---------------------------------------------------------------------------------------------
// Scratch layout of one frame evaluation.
//   device: [results block | row_sd | row_max | partials]
//   results block (copied back in one transfer): FrameSummary | cps rows [2][3][h] | row_ssd [nrows]
// ---------------------------------------------------------------------------------------------
// Internal flag beside the QMIV_* bits: the SSIM evaluation without the luminance term.
constexpr uint32_t kSsimNoLuminance = 1u << 30;

// One SSIM evaluation (one component of one picture pair at one pyramid level).
struct SsimEval {
  int kind;       // 0 SSIM (calc_l as asked), 1 MS-SSIM level, 2 IV-SSIM T2R, 3 IV-SSIM R2T, 4 SSIM without L
  int cmp;        // component
  int level;      // pyramid level (MS-SSIM)
  int32_t w, h;   // size at this level
  size_t off_rows;  // per-row sums in the results block
  size_t off_part;  // device: its [tiles_x][h] tile partials
};

struct Layout {
  size_t off_summary, off_cps_rows, off_row_ssd, results_bytes;
  size_t off_row_sd, off_row_max, off_partials, total;
  int32_t nseg;
  // The summary every later kernel and the host read: off_summary, or under concealment the summary
  // of the statistics over the clipped pictures (off_summary then keeps the range flags as read).
  size_t off_summary2 = 0;
  // The general search's partials when it runs as the fallback of an out-of-range frame.
  size_t off_partials_gen = 0;
  int32_t nseg_gen = 0;
  std::vector<SsimEval> ssim;
  size_t off_pyr = 0;         // device: MS-SSIM pyramid levels 1..4, [picture][component][level]
  size_t pyr_level_off[5] = {0, 0, 0, 0, 0};  // element offsets of the levels within one plane chain
  size_t pyr_chain = 0;       // elements of one (picture, component) chain
};

// Fast CPS path eligibility: every operand of the packed key must stay inside the ranges the
// kernel's 24-bit multiplies and 32-bit keys assume, and the reference's int32 error must not
// wrap (the generic kernel reproduces wrapping).
bool fast_cps_ok(const Geom& g, const qmiv_params& p, const int32_t* gcd_bound) {
  if (p.search_range != 2) return false;
  const int64_t maxv = (1 << g.bd) - 1;
  int64_t wsum = 0;
  for (int c = 0; c < 3; c++) {
    if (p.cmp_weights_search[c] < 0) return false;
    wsum += p.cmp_weights_search[c];
  }
  if (p.cmp_weights_search[3] != 0 && gcd_bound[3] != 0) return false;
  int64_t m = 0;
  for (int c = 0; c < 3; c++) m = std::max<int64_t>(m, maxv + std::abs((int64_t)gcd_bound[c]));
  for (int c = 0; c < 3; c++)
    if (64 * (int64_t)p.cmp_weights_search[c] * m >= ((int64_t)1 << 23)) return false;
  const int64_t key_bound = 32 * (2 * wsum * m * maxv + wsum * maxv * maxv) + 512;
  if (key_bound >= ((int64_t)1 << 31)) return false;
  int64_t ref_err = 0;
  for (int c = 0; c < 3; c++) ref_err += (int64_t)p.cmp_weights_search[c] * (m + maxv) * (m + maxv);
  if (ref_err >= ((int64_t)1 << 31)) return false;
  return true;
}

// 4:2:0 search with 16-bit key operands (kernels_cps420s.hip): samples below 2^10, so 32 Y fits a
// signed 16-bit half, and -2 w_c X_c fits one for every |X_c| up to maxv + |gcd| (w_c U fits an
// unsigned half for the cell constant); widths that are a multiple of its tile width. Holds for the
// reference's default weights at 8 to 10 bits and the common picture widths (1280, 1920, 2048, 3840,
// 4096); other widths use the 4:2:0 kernel with 32-bit operands.
bool cps420s_ok(const Geom& g, const qmiv_params& p, const int32_t* gcd_bound) {
  int32_t tw, th;
  cps420s_tile(&tw, &th);
  if (g.cs != 1 || g.bd > 10 || g.bytes_per_sample > 2 || g.w % tw != 0 || (g.h & 1)) return false;
  const int64_t maxv = (1 << g.bd) - 1;
  if (32 * maxv > 32767) return false;
  for (int c = 0; c < 3; c++) {
    const int64_t m = maxv + std::abs((int64_t)gcd_bound[c]);
    if (2 * (int64_t)p.cmp_weights_search[c] * m > 32767) return false;
    if ((int64_t)p.cmp_weights_search[c] * maxv > 65535) return false;
  }
  return true;
}

// Tile of the fast CPS kernel that serves this geometry: the 4:4:4 kernel (one column per lane),
// the 4:2:0 kernel (two columns per lane) or the 4:2:0 kernel with 16-bit operands (2 x 4 pels per lane).
void cps_tile(int kind, int32_t* tw, int32_t* th) {
  if (kind == kCps420s) cps420s_tile(tw, th);
  else if (kind == kCps420) cps_fast420_tile(tw, th);
  else cps_fast_tile(tw, th);
}

// Tiles per workgroup along x: just enough that every workgroup is resident at once
// (CUs x resident workgroups per CU), which gives one round with no tail, and within the bound
// that keeps the per-lane row sums below 2^32: one pel per tile and lane column for the 4:4:4
// kernel, two for the 4:2:0 kernel, and for the 16-bit 4:2:0 kernel the wave totals too
// (128 pels per tile and row).
int32_t pick_sweep(qmiv_ctx* ctx, const Geom& g, int kind, const int32_t* gcd_bound) {
  int32_t tw, th;
  cps_tile(kind, &tw, &th);
  const int32_t bands = (g.h + th - 1) / th;
  const int32_t tiles_x = (g.w + tw - 1) / tw;
  const int64_t tiles = (int64_t)bands * tiles_x;
  int32_t& bpc = ctx->cps_blocks_per_cu[kind][g.bytes_per_sample == 1 ? 0 : 1];
  if (bpc == 0) {
    bpc = kind == kCps420s ? cps420s_blocks_per_cu(g.bytes_per_sample)
          : kind == kCps420 ? cps_fast420_blocks_per_cu(g.bytes_per_sample)
                            : cps_fast_blocks_per_cu(g.bytes_per_sample);
  }
  // The 16-bit kernel runs each (band, segment) as two workgroups, one per direction. One round
  // suits frame sets, where the next frame's kernels run beside this one: alone, two rounds of
  // shorter sweeps measure faster (4096x2048: 55.0 us at 4 tiles per workgroup, 57.5 us at 8), but
  // the frame-set and batched workloads run slower with them.
  const int64_t resident = std::max<int64_t>(1, (int64_t)ctx->num_cus * bpc / (kind == kCps420s ? 2 : 1));
  int64_t sweep = (tiles + resident - 1) / resident;
  const int64_t maxv = (1 << g.bd) - 1;
  int64_t m = 0;
  for (int c = 0; c < 3; c++) m = std::max<int64_t>(m, maxv + std::abs((int64_t)gcd_bound[c]));
  // Largest squared error of one pel: |X - Y| <= maxv + |gcd| for the 16-bit kernel's in-range
  // samples; the other kernels keep their wider bound.
  const int64_t sq = kind == kCps420s ? m * m * 128 : (m + maxv) * (m + maxv) * (kind == kCps420 ? 2 : 1);
  sweep = std::min<int64_t>(sweep, std::max<int64_t>(1, (int64_t)0xffffffffLL / std::max<int64_t>(sq, 1)));
  return (int32_t)std::max<int64_t>(1, std::min<int64_t>(tiles_x, sweep));
}

// conceal: a second summary (concealment); fallback: partials of the general search beside the
// fast one's (a frame flow that must search frames with out-of-range pels).
Layout make_layout(const Geom& g, bool fast, int kind, int32_t sweep, uint32_t ssim_metrics, bool conceal = false,
                   bool fallback = false) {
  Layout L{};
  const int32_t nrows = g.h + 2 * g.ch;
  int32_t tw, th;
  cps_tile(kind, &tw, &th);
  const int32_t tiles_x = (g.w + tw - 1) / tw;
  L.nseg = fast ? (tiles_x + sweep - 1) / sweep : (g.w + kGenericCols - 1) / kGenericCols;
  size_t o = 0;
  L.off_summary = o;
  o = align_up(o + sizeof(FrameSummary), 256);
  L.off_summary2 = L.off_summary;
  if (conceal) {
    L.off_summary2 = o;
    o = align_up(o + sizeof(FrameSummary), 256);
  }
  L.off_cps_rows = o;
  o = align_up(o + sizeof(uint64_t) * 6 * (size_t)g.h, 256);
  L.off_row_ssd = o;
  o = align_up(o + sizeof(uint64_t) * (size_t)nrows, 256);
  // SSIM family: per-row sums of every evaluation, in the results block.
  auto add_eval = [&](int kind, int cmp, int level) {
    SsimEval e{kind, cmp, level, g.w >> level, g.h >> level, 0, 0};
    e.off_rows = o;
    o = align_up(o + sizeof(double) * (size_t)std::max(e.h, 1), 256);
    L.ssim.push_back(e);
  };
  if (ssim_metrics & QMIV_SSIM)
    for (int c = 0; c < 3; c++) add_eval((ssim_metrics & kSsimNoLuminance) ? 4 : 0, c, 0);
  if (ssim_metrics & QMIV_MSSSIM)
    for (int c = 0; c < 3; c++)
      for (int l = 0; l < 5; l++) add_eval(1, c, l);
  if (ssim_metrics & QMIV_IVSSIM)
    for (int c = 0; c < 3; c++) {
      add_eval(2, c, 0);
      add_eval(3, c, 0);
    }
  L.results_bytes = o;
  L.off_row_sd = o;
  o = align_up(o + sizeof(int64_t) * (size_t)nrows, 256);
  L.off_row_max = o;
  o = align_up(o + sizeof(uint32_t) * 2 * (size_t)nrows, 256);
  L.off_partials = o;
  o = align_up(o + sizeof(uint64_t) * 6 * (size_t)g.h * (size_t)L.nseg, 256);
  if (fallback) {
    L.nseg_gen = (g.w + kGenericCols - 1) / kGenericCols;
    L.off_partials_gen = o;
    o = align_up(o + sizeof(uint64_t) * 6 * (size_t)g.h * (size_t)L.nseg_gen, 256);
  }
  for (SsimEval& e : L.ssim) {
    e.off_part = o;
    o = align_up(o + sizeof(double) * std::max<size_t>((size_t)ssim_tiles_x(e.w) * (size_t)std::max(e.h, 1), 1), 256);
  }
  if (ssim_metrics & QMIV_MSSSIM) {
    size_t chain = 0;
    for (int l = 1; l < 5; l++) {
      L.pyr_level_off[l] = chain;
      chain += (size_t)(g.w >> l) * (size_t)(g.h >> l);
    }
    L.pyr_chain = align_up(chain, 128);
    L.off_pyr = o;
    o = align_up(o + sizeof(uint16_t) * L.pyr_chain * 6, 256);
  }
  L.total = o;
  return L;
}

// Brackets one kernel launch with events when profiling is on.
struct KernelTimer {
  qmiv_ctx* ctx;
  int kind;
  hipStream_t st;
  hipEvent_t end = nullptr;
  KernelTimer(qmiv_ctx* c, int k, hipStream_t s) : ctx(c), kind(k), st(s) {
    if (!ctx->profiling) return;
    if (ctx->ev_used == ctx->ev_pool.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess) return;
      if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
      ctx->ev_pool.emplace_back(a, b);
      ctx->ev_kind.push_back(0);
    }
    auto& pr = ctx->ev_pool[ctx->ev_used];
    ctx->ev_kind[ctx->ev_used] = kind;
    ctx->ev_used++;
    (void)hipEventRecord(pr.first, st);
    end = pr.second;
  }
  ~KernelTimer() {
    if (end) (void)hipEventRecord(end, st);
  }
};

// Adds the elapsed times of the recorded event pairs (after the stream has been synchronised).
void harvest_events(qmiv_ctx* ctx) {
  for (size_t i = 0; i < ctx->ev_used; i++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev_pool[i].first, ctx->ev_pool[i].second) == hipSuccess) {
      ctx->prof.kernel_ms[ctx->ev_kind[i]] += ms;
      ctx->prof.launches[ctx->ev_kind[i]] += 1;
    }
  }
  ctx->ev_used = 0;
}

struct FrameWork {
  int32_t ipa = QMIV_PEL_STOP;   // invalid-pel action of the frame flow
  bool need_gcd = false;         // compute the GCD on the device
  const int32_t* gcd = nullptr;  // or use this one
  bool need_cps = false;
  bool need_scp = false;
  uint32_t ssim_metrics = 0;     // QMIV_SSIM | QMIV_MSSSIM | QMIV_IVSSIM
  int16_t* best0 = nullptr;      // optional per-pel winner indices of the two directions
  int16_t* best1 = nullptr;
};

// The device work a frame's metrics need (the application's per-frame flow, xAppQMIV.cpp:686-704).
FrameWork frame_work(uint32_t metrics, const qmiv_params& prm) {
  FrameWork fw;
  fw.ipa = prm.invalid_pel_action;
  fw.need_gcd = (metrics & (QMIV_IVPSNR | QMIV_IVSSIM)) != 0;
  fw.need_cps = (metrics & QMIV_IVPSNR) != 0;
  fw.need_scp = (metrics & QMIV_IVSSIM) != 0;
  fw.ssim_metrics = metrics & (QMIV_SSIM | QMIV_MSSSIM | QMIV_IVSSIM | kSsimNoLuminance);
  return fw;
}

// Everything decided on the host before a frame's kernels are enqueued.
struct FramePlan {
  Layout L;
  bool fast = false;     // fast CPS kernel (else the generic one)
  bool fast420 = false;  // ... its 4:2:0 form
  int kind = kCps444;     // which fast kernel (cps_tile)
  int32_t sweep = 1;
  int32_t md[4];
  bool conceal = false;   // CNCL: clip out-of-range pictures, then the statistics again
  bool fallback = false;  // WARN / SKIP: the general search, for a frame with out-of-range pels
};

// Samples of this geometry can exceed the bit-depth maximum (16-bit storage below 16 bits).
bool may_exceed(const Geom& g) { return g.bytes_per_sample == 2 && g.bd < 16; }

// sweep_override > 0 caps the CPS tiles per workgroup (the batch autotuner's knob); 0 keeps
// the occupancy rule of pick_sweep.
FramePlan plan_frame(qmiv_ctx* ctx, const Geom& g, const qmiv_params& prm, const FrameWork& fw,
                     int32_t sweep_override = 0) {
  FramePlan P;
  for (int c = 0; c < 4; c++) P.md[c] = fin::gcd_max_diff(prm.unnoticeable_coef[c], g.bd);
  const int32_t* gbound = fw.gcd ? fw.gcd : P.md;
  P.fast = fast_cps_ok(g, prm, gbound);
  P.fast420 = P.fast && g.cs == 1;
  P.kind = !P.fast420 ? kCps444 : (cps420s_ok(g, prm, gbound) ? kCps420s : kCps420);
  P.sweep = pick_sweep(ctx, g, P.kind, gbound);
  if (sweep_override > 0) P.sweep = std::min(P.sweep, sweep_override);
  P.conceal = fw.ipa == QMIV_PEL_CNCL && may_exceed(g);
  P.fallback = (fw.ipa == QMIV_PEL_WARN || fw.ipa == QMIV_PEL_SKIP) && may_exceed(g) && P.fast &&
               (fw.need_cps || fw.need_scp);
  P.L = make_layout(g, P.fast, P.kind, P.sweep, fw.ssim_metrics, P.conceal, P.fallback);
  return P;
}

PlaneView pic_view(const qmiv_pic* p, int c) {
  const Geom& g = p->g;
  PlaneView v;
  const int64_t off = c == 0 ? 0 : (c == 1 ? g.cb_off : g.cr_off);
  v.base = (const char*)p->dev + off * g.bytes_per_sample;
  v.stride = c == 0 ? g.w : g.cw;
  v.shift = c == 0 ? 0 : g.cs;
  v.bytes = g.bytes_per_sample;
  v.nbytes = (uint32_t)((c == 0 ? (size_t)g.w * g.h : (size_t)g.cw * g.ch) * g.bytes_per_sample);
  return v;
}

PlaneView u16_view(const uint16_t* base, int32_t w, int32_t h) {
  PlaneView v;
  v.base = base;
  v.stride = w;
  v.shift = 0;
  v.bytes = 2;
  v.nbytes = (uint32_t)((size_t)w * h * 2);
  return v;
}

// Enqueues the SSIM-family evaluations of one frame (after the CPS kernel has written the SCP
// pictures when IV-SSIM is asked): pyramid levels, then one windowed-moment pass per evaluation.
qmiv_status enqueue_ssim(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const FramePlan& P, char* dbase,
                         const uint16_t* scp, hipStream_t st) {
  const Geom& g = p0->g;
  const Layout& L = P.L;
  uint16_t* pyr = L.pyr_chain ? (uint16_t*)(dbase + L.off_pyr) : nullptr;
  auto level_ptr = [&](int pic, int c, int l) { return pyr + ((size_t)pic * 3 + c) * L.pyr_chain + L.pyr_level_off[l]; };
  bool need_pyr = false;
  for (const SsimEval& e : L.ssim) need_pyr |= e.kind == 1;
  if (need_pyr) {
    KernelTimer t(ctx, QMIV_K_PYRAMID, st);
    PyramidJobs j{};
    j.n = 6;
    j.w = g.w;
    j.h = g.h;
    for (int pic = 0; pic < 2; pic++)
      for (int c = 0; c < 3; c++) {
        j.src[pic * 3 + c] = pic_view(pic ? p1 : p0, c);
        for (int l = 1; l < 5; l++) j.dst[pic * 3 + c][l - 1] = level_ptr(pic, c, l);
      }
    QMIV_HIP(launch_pyramid(j, st));
  }
  SsimArgs a{};
  fin::ssim_window(a.f);
  const double maxv = (double)((1 << g.bd) - 1);
  const double k1 = 0.01 * maxv, k2 = 0.03 * maxv;
  a.c1 = k1 * k1;
  a.c2 = k2 * k2;
  const uint16_t* scp0 = scp;
  const uint16_t* scp1 = scp0 ? scp0 + 3 * (size_t)g.w * g.h : nullptr;
  // Full-resolution evaluations of one component that share pictures run as one fused pass
  // (launch_ssim_multi): SSIM(P0, P1), its CS (MS-SSIM level 0 or SSIM without L), and IV-SSIM's
  // SSIM(P0, SCP1) and SSIM(P1, SCP0). The rest run one by one.
  std::vector<char> fused(L.ssim.size(), 0);
  for (int c = 0; c < 3; c++) {
    int i_lcs = -1, i_cs = -1, i_c = -1, i_d = -1;
    for (size_t i = 0; i < L.ssim.size(); i++) {
      const SsimEval& e = L.ssim[i];
      if (e.cmp != c || e.level != 0) continue;
      if (e.kind == 0) i_lcs = (int)i;
      else if ((e.kind == 4 || e.kind == 1) && i_cs < 0) i_cs = (int)i;
      else if (e.kind == 2 && scp1) i_c = (int)i;
      else if (e.kind == 3 && scp0) i_d = (int)i;
    }
    const int n = (i_lcs >= 0) + (i_cs >= 0) + (i_c >= 0) + (i_d >= 0);
    if ((i_lcs < 0 && i_cs < 0) || n < 2) continue;
    SsimMultiArgs m{};
    std::memcpy(m.f, a.f, sizeof(m.f));
    m.c1 = a.c1;
    m.c2 = a.c2;
    m.w = g.w;
    m.h = g.h;
    m.tiles_x = ssim_tiles_x(g.w);
    const size_t plane = (size_t)g.w * g.h;
    m.p0 = pic_view(p0, c);
    m.p1 = pic_view(p1, c);
    if (i_c >= 0) m.scp1 = u16_view(scp1 + c * plane, g.w, g.h);
    if (i_d >= 0) m.scp0 = u16_view(scp0 + c * plane, g.w, g.h);
    auto part = [&](int i) { return i >= 0 ? (double*)(dbase + L.ssim[i].off_part) : nullptr; };
    m.part_lcs = part(i_lcs);
    m.part_cs = part(i_cs);
    m.part_c = part(i_c);
    m.part_d = part(i_d);
    for (int i : {i_lcs, i_cs, i_c, i_d})
      if (i >= 0) fused[i] = 1;
    KernelTimer t(ctx, QMIV_K_SSIM, st);
    QMIV_HIP(launch_ssim_multi(m, st));
  }
  // The remaining evaluations (mostly the MS-SSIM pyramid levels, each too small to fill the GPU)
  // go into batched launches of up to kMaxSsimBatch evaluations; a lone one runs by itself.
  std::vector<SsimArgs> rest;
  for (size_t ei = 0; ei < L.ssim.size(); ei++) {
    if (fused[ei]) continue;
    const SsimEval& e = L.ssim[ei];
    a.w = e.w;
    a.h = e.h;
    a.tiles_x = ssim_tiles_x(e.w);
    a.calc_l = ((e.kind == 1 && (e.level == 0 || e.level == 4)) || e.kind == 4) ? 0 : 1;  // xSSIM.cpp:74,80
    const size_t plane = (size_t)g.w * g.h;
    switch (e.kind) {
      case 0:  // calcPicSSIM(P0, P1)
      case 4:
        a.tst = pic_view(p0, e.cmp);
        a.ref = pic_view(p1, e.cmp);
        break;
      case 1:
        if (e.level == 0) {
          a.tst = pic_view(p0, e.cmp);
          a.ref = pic_view(p1, e.cmp);
        } else {
          a.tst = u16_view(level_ptr(0, e.cmp, e.level), e.w, e.h);
          a.ref = u16_view(level_ptr(1, e.cmp, e.level), e.w, e.h);
        }
        break;
      case 2:  // T2R = calcPicSSIM(P0, SCP1)
        a.tst = pic_view(p0, e.cmp);
        a.ref = u16_view(scp1 + e.cmp * plane, g.w, g.h);
        break;
      default:  // R2T = calcPicSSIM(P1, SCP0)
        a.tst = pic_view(p1, e.cmp);
        a.ref = u16_view(scp0 + e.cmp * plane, g.w, g.h);
        break;
    }
    a.partials = (double*)(dbase + e.off_part);
    rest.push_back(a);
  }
  for (size_t i0 = 0; i0 < rest.size(); i0 += kMaxSsimBatch) {
    const size_t n = std::min<size_t>(kMaxSsimBatch, rest.size() - i0);
    KernelTimer t(ctx, QMIV_K_SSIM, st);
    if (n == 1) {
      QMIV_HIP(launch_ssim(rest[i0], st));
      continue;
    }
    SsimBatchArgs b{};
    b.n = (int32_t)n;
    b.c1 = a.c1;
    b.c2 = a.c2;
    std::memcpy(b.f, a.f, sizeof(b.f));
    for (size_t k = 0; k < n; k++) {
      const SsimArgs& r = rest[i0 + k];
      SsimBatchJob& J = b.job[k];
      J.tst = r.tst;
      J.ref = r.ref;
      J.w = r.w;
      J.h = r.h;
      J.tiles_x = r.tiles_x;
      J.calc_l = r.calc_l;
      J.partials = r.partials;
    }
    QMIV_HIP(launch_ssim_batch(b, st));
  }
  SsimRowsJobs j{};
  for (const SsimEval& e : L.ssim) {
    if (j.n == kMaxSsimEvals) return fail(QMIV_ERR_UNSUPPORTED, "too many SSIM evaluations in one frame");
    j.tiles_x[j.n] = ssim_tiles_x(e.w);
    j.h[j.n] = e.h;
    j.part[j.n] = (const double*)(dbase + e.off_part);
    j.rows[j.n] = (double*)(dbase + e.off_rows);
    j.max_h = std::max(j.max_h, e.h);
    j.n++;
  }
  KernelTimer t(ctx, QMIV_K_SSIM_ROWS, st);
  QMIV_HIP(launch_ssim_rows(j, st));
  return QMIV_OK;
}

// Enqueues one frame's statistics (+ GCD) and, when asked, its CPS search and SSIM family on
// stream `st`, with device scratch at `dbase`, the shift-compensated pictures (when needed) at
// `scp` (2 x [3][h][w] uint16) and the results block copied into pinned `hbase`.
qmiv_status enqueue_frame(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_params& prm,
                          const FrameWork& fw, const FramePlan& P, char* dbase, char* hbase, uint16_t* scp,
                          hipStream_t st, hipStream_t copy_st = nullptr, hipEvent_t copy_ev = nullptr,
                          hipEvent_t done_ev = nullptr) {
  const Geom& g = p0->g;
  const Layout& L = P.L;
  FrameSummary* sum = (FrameSummary*)(dbase + L.off_summary);
  uint64_t* cps_rows = (uint64_t*)(dbase + L.off_cps_rows);
  uint64_t* row_ssd = (uint64_t*)(dbase + L.off_row_ssd);
  int64_t* row_sd = (int64_t*)(dbase + L.off_row_sd);
  uint32_t* row_max = (uint32_t*)(dbase + L.off_row_max);
  uint64_t* partials = (uint64_t*)(dbase + L.off_partials);

  FrameSummary* sum2 = (FrameSummary*)(dbase + L.off_summary2);
  // The 4:2:0 fast search adds its row sums straight into cps_rows, which the statistics pass
  // clears first; the other searches write per-segment partials that a reduction pass sums.
  const bool fused_rows = (fw.need_cps || fw.need_scp) && P.fast420;
  if (fw.gcd) QMIV_HIP(hipMemcpyAsync(sum->gcd, fw.gcd, sizeof(int32_t) * 4, hipMemcpyHostToDevice, st));
  {
    KernelTimer t(ctx, QMIV_K_STATS, st);
    QMIV_HIP(launch_stats(p0->dev, p1->dev, g, row_sd, row_ssd, row_max, st, nullptr, fused_rows ? cps_rows : nullptr,
                          fused_rows ? 6 * g.h : 0));
  }
  {
    KernelTimer t(ctx, QMIV_K_SUMMARY, st);
    QMIV_HIP(launch_summary(g, row_sd, row_ssd, row_max, P.md, fw.need_gcd ? 1 : 0, sum, st));
  }
  if (P.conceal) {
    // xAppQMIV::validateFrames with CNCL (apps/QMIV/src/xAppQMIV.cpp:753-762): a picture holding a
    // pel above the maximum is clipped in place, then the statistics and the GCD are computed
    // again over the clipped pictures. Every step reads the first summary and does nothing when the
    // frame was in range, so the frame needs no wait on the host.
    const int32_t maxv = (1 << g.bd) - 1;
    QMIV_HIP(launch_clip(p0->dev, g, maxv, &sum->max_pel[0], st));
    QMIV_HIP(launch_clip(p1->dev, g, maxv, &sum->max_pel[1], st));
    if (fw.gcd) QMIV_HIP(hipMemcpyAsync(sum2->gcd, fw.gcd, sizeof(int32_t) * 4, hipMemcpyHostToDevice, st));
    {
      KernelTimer t(ctx, QMIV_K_STATS, st);
      QMIV_HIP(launch_stats(p0->dev, p1->dev, g, row_sd, row_ssd, row_max, st, sum));
    }
    {
      KernelTimer t(ctx, QMIV_K_SUMMARY, st);
      QMIV_HIP(launch_summary(g, row_sd, row_ssd, row_max, P.md, fw.need_gcd ? 1 : 0, sum2, st, sum));
    }
  }
  if (fw.need_cps || fw.need_scp) {
    uint16_t* scp0 = nullptr;
    uint16_t* scp1 = nullptr;
    if (fw.need_scp) {
      if (!scp) return fail(QMIV_ERR_INVALID_ARG, "no buffer for the shift-compensated pictures");
      scp0 = scp;
      scp1 = scp0 + 3 * (size_t)g.w * g.h;
    }
    CpsArgs a{};
    a.p0 = p0->dev;
    a.p1 = p1->dev;
    a.g = g;
    a.sum = sum2;
    for (int c = 0; c < 4; c++) a.ws[c] = prm.cmp_weights_search[c];
    a.B = prm.search_range;
    int32_t tw, th;
    cps_tile(P.kind, &tw, &th);
    a.tiles_x = (g.w + tw - 1) / tw;
    a.sweep = P.sweep;
    a.nseg = L.nseg;
    a.partials = partials;
    if (fused_rows) a.rows = cps_rows;
    a.scp0 = scp0;
    a.scp1 = scp1;
    a.best0 = fw.best0;
    a.best1 = fw.best1;
    a.maxv = (1 << g.bd) - 1;
    // WARN / SKIP search frames with pels above the maximum too: the fast kernels run when the frame
    // is in range, the general kernel (any 16-bit samples, the reference's wrapping int32 errors)
    // when it is not; each gate reads the summary's range flag.
    const bool any_range = fw.ipa == QMIV_PEL_WARN || fw.ipa == QMIV_PEL_SKIP;
    if (!P.fast) a.gate = any_range ? kGateAlways : kGateInRange;
    {
      KernelTimer t(ctx, P.fast ? QMIV_K_CPS : QMIV_K_CPS_GENERAL, st);
      if (P.kind == kCps420s) QMIV_HIP(launch_cps420s(a, st));
      else if (P.fast420) QMIV_HIP(launch_cps_fast420(a, st));
      else if (P.fast) QMIV_HIP(launch_cps_fast(a, st));
      else QMIV_HIP(launch_cps_generic(a, st));
    }
    if (!fused_rows) {
      KernelTimer t(ctx, QMIV_K_CPS_ROWS, st);
      QMIV_HIP(launch_cps_rows(partials, L.nseg, g.h, cps_rows, st, &sum2->in_range, P.fallback ? 1 : 0));
    }
    if (P.fallback) {
      CpsArgs b = a;
      b.rows = nullptr;
      b.gate = kGateOutOfRange;
      b.nseg = L.nseg_gen;
      b.partials = (uint64_t*)(dbase + L.off_partials_gen);
      {
        KernelTimer t(ctx, QMIV_K_CPS_GENERAL, st);
        QMIV_HIP(launch_cps_generic(b, st));
      }
      KernelTimer t(ctx, QMIV_K_CPS_ROWS, st);
      QMIV_HIP(launch_cps_rows(b.partials, L.nseg_gen, g.h, cps_rows, st, &sum2->in_range, 2));
    }
  }
  if (!L.ssim.empty()) QMIV_TRY(enqueue_ssim(ctx, p0, p1, P, dbase, scp, st));
  if (copy_st) {
    QMIV_HIP(hipEventRecord(copy_ev, st));
    QMIV_HIP(hipStreamWaitEvent(copy_st, copy_ev, 0));
    QMIV_HIP(hipMemcpyAsync(hbase, dbase, L.results_bytes, hipMemcpyDeviceToHost, copy_st));
    if (done_ev) QMIV_HIP(hipEventRecord(done_ev, copy_st));
  } else {
    QMIV_HIP(hipMemcpyAsync(hbase, dbase, L.results_bytes, hipMemcpyDeviceToHost, st));
    if (done_ev) QMIV_HIP(hipEventRecord(done_ev, st));
  }
  return QMIV_OK;
}

qmiv_status wait_stream(qmiv_ctx* ctx) {
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  for (hipStream_t a : ctx->aux_stream) QMIV_HIP(hipStreamSynchronize(a));
  QMIV_HIP(hipStreamSynchronize(ctx->copy_stream));
  harvest_events(ctx);
  return QMIV_OK;
}

// Single frame through enqueue + wait, results at the start of ctx->h_results.
// The processor-level calls validate nothing (the reference's processors compute over the samples
// as they are): their frames take the WARN / SKIP search paths.
qmiv_status run_frame_kernels(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_params& prm,
                              const FrameWork& fw_in, Layout& L, bool& used_fast) {
  FrameWork fw = fw_in;
  fw.ipa = QMIV_PEL_SKIP;
  const FramePlan P = plan_frame(ctx, p0->g, prm, fw);
  L = P.L;
  used_fast = P.fast;
  QMIV_TRY(ensure_scratch(ctx, P.L.total));
  QMIV_TRY(ensure_results(ctx, P.L.results_bytes));
  uint16_t* scp = nullptr;
  if (fw.need_scp) {
    QMIV_TRY(ensure_scp(ctx, 6 * (size_t)p0->g.w * p0->g.h * sizeof(uint16_t)));
    scp = (uint16_t*)ctx->d_scp;
  }
  QMIV_TRY(enqueue_frame(ctx, p0, p1, prm, fw, P, (char*)ctx->d_scratch, (char*)ctx->h_results, scp, ctx->stream));
  return wait_stream(ctx);
}

// Per-row 4:4:4 SSD of component c from the per-plane rows (4:2:0 chroma rows count twice
// horizontally: each chroma row of the 4:4:4 picture is two copies of every sample).
void rows444(const Geom& g, const uint64_t* row_ssd, int c, std::vector<uint64_t>& out) {
  out.resize((size_t)g.h);
  if (c == 0) {
    for (int32_t y = 0; y < g.h; y++) out[(size_t)y] = row_ssd[y];
    return;
  }
  const uint64_t* r = row_ssd + g.h + (size_t)(c - 1) * g.ch;
  for (int32_t y = 0; y < g.h; y++) out[(size_t)y] = r[y >> g.cs] << g.cs;
}

void apply_row_semantics(uint64_t* rows, size_t n, bool int32_rows) {
  if (!int32_rows) return;
  for (size_t i = 0; i < n; i++) rows[i] = (uint64_t)(int64_t)(int32_t)(uint32_t)rows[i];
}

qmiv_status check_pair(const qmiv_pic* a, const qmiv_pic* b) {
  if (!a || !b) return fail(QMIV_ERR_INVALID_ARG, "null picture");
  if (!same_geom(a->g, b->g)) return fail(QMIV_ERR_INVALID_ARG, "pictures differ in size, bit depth or chroma format");
  if (a->g.cf == 400) return fail(QMIV_ERR_INVALID_ARG, "4:0:0 pictures serve as masks only");
  if (a->ctx != b->ctx) return fail(QMIV_ERR_INVALID_ARG, "pictures belong to different contexts");
  return QMIV_OK;
}

// Host finalisation of one frame from its results block (apps/QMIV/src/xAppQMIV.cpp:831-904).
qmiv_status finalize_frame(qmiv_ctx* ctx, const Geom& g, const qmiv_params& prm, uint32_t metrics,
                           const FrameWork& fw, const Layout& L, const char* hb, qmiv_frame_result* out) {
  std::memset(out, 0, sizeof(*out));
  // The range flags of the pictures as read; everything else from the summary the kernels used
  // (under concealment: of the clipped pictures).
  const FrameSummary* read = (const FrameSummary*)(hb + L.off_summary);
  const FrameSummary* sum = (const FrameSummary*)(hb + L.off_summary2);
  const uint64_t* h_cps = (const uint64_t*)(hb + L.off_cps_rows);
  const uint64_t* h_ssd = (const uint64_t*)(hb + L.off_row_ssd);
  const int32_t maxv = (1 << g.bd) - 1;
  out->in_range[0] = g.bd >= 16 || read->max_pel[0] <= maxv;
  out->in_range[1] = g.bd >= 16 || read->max_pel[1] <= maxv;
  for (int c = 0; c < 3; c++) out->exact[c] = sum->ssd[c] == 0;
  if (fw.need_gcd)
    for (int c = 0; c < 4; c++) out->gcd[c] = sum->gcd[c];
  if (!read->in_range && fw.ipa == QMIV_PEL_STOP)
    return fail(QMIV_ERR_PEL_RANGE, "input contains pels above the bit-depth maximum");

  const int32_t area = g.w * g.h;
  const double fake = fin::psnr_from_ssd(1.0, area, g.bd);
  const double* ws = nullptr;
  double corr = 1.0;
  if (prm.erp) {
    const qmiv_ctx::WsTable& t = ws_table(ctx, g.h, prm.lat_range_deg);
    ws = t.w.data();
    corr = t.corr;
  }
  if (metrics & QMIV_PSNR) {
    for (int c = 0; c < 3; c++) {
      const double v = fin::psnr_from_ssd((double)sum->ssd[c], area, g.bd);
      out->psnr[c] = (prm.fake_psnr_for_exact && out->exact[c]) ? fake : v;
    }
    out->psnr_pic = fin::pic_value(out->psnr, prm.cmp_weights_average);
  }
  if (metrics & QMIV_WSPSNR) {
    double v[3];
    if (prm.erp) {
      // Per-plane row SSDs onto the 4:4:4 rows: a 4:2:0 chroma row counts twice horizontally and
      // serves two 4:4:4 rows.
      const uint64_t* r[3] = {h_ssd, h_ssd + g.h, h_ssd + g.h + g.ch};
      const int sh[3] = {0, g.cs, g.cs};
      fin::wspsnr3_from_rows(r, sh, g.h, ws, corr, area, g.bd, v);
    } else {
      for (int c = 0; c < 3; c++) v[c] = fin::psnr_from_ssd((double)sum->ssd[c], area, g.bd);
    }
    for (int c = 0; c < 3; c++) out->wspsnr[c] = (prm.fake_psnr_for_exact && out->exact[c]) ? fake : v[c];
    out->wspsnr_pic = fin::pic_value(out->wspsnr, prm.cmp_weights_average);
  }
  if (metrics & QMIV_IVPSNR) {
    std::vector<uint64_t> rows(h_cps, h_cps + 6 * (size_t)g.h);
    apply_row_semantics(rows.data(), rows.size(), prm.cps_int32_rows != 0);
    for (int d = 0; d < 2; d++)
      for (int c = 0; c < 3; c++) {
        uint64_t s = 0;
        for (int32_t y = 0; y < g.h; y++) s += rows[((size_t)d * 3 + c) * g.h + y];
        out->cps_sum[d][c] = s;
      }
    const double* wsp = ws;
    out->ivpsnr_r2t = fin::qual_from_rows(rows.data(), g.h, wsp, area, g.bd, prm.cmp_weights_average);
    out->ivpsnr_t2r = fin::qual_from_rows(rows.data() + 3 * (size_t)g.h, g.h, wsp, area, g.bd, prm.cmp_weights_average);
    out->ivpsnr = fin::min2(out->ivpsnr_r2t, out->ivpsnr_t2r);
  }
  if (!L.ssim.empty()) {
    double ms[3][5] = {{0}}, t2r[3] = {0, 0, 0}, r2t[3] = {0, 0, 0};
    for (const SsimEval& e : L.ssim) {
      const double v = fin::ssim_from_rows((const double*)(hb + e.off_rows), e.h, e.w);
      if (e.kind == 0 || e.kind == 4) out->ssim[e.cmp] = v;
      else if (e.kind == 1) ms[e.cmp][e.level] = v;
      else if (e.kind == 2) t2r[e.cmp] = v;
      else r2t[e.cmp] = v;
    }
    if (metrics & QMIV_SSIM) out->ssim_pic = fin::pic_value(out->ssim, prm.cmp_weights_average);
    if (metrics & QMIV_MSSSIM) {
      for (int c = 0; c < 3; c++) {
        out->msssim[c] = fin::msssim_combine(ms[c]);
        for (int l = 0; l < 5; l++) out->msssim_scales[l][c] = ms[c][l];
      }
      out->msssim_pic = fin::pic_value(out->msssim, prm.cmp_weights_average);
    }
    if (metrics & QMIV_IVSSIM) {
      // xIVSSIM::calcPicIVSSIM (libs/pmbb-ivqm/src/xSSIM.cpp:192-212)
      out->ivssim_t2r = fin::weighted_average4(t2r, prm.cmp_weights_average);
      out->ivssim_r2t = fin::weighted_average4(r2t, prm.cmp_weights_average);
      out->ivssim = fin::min2(out->ivssim_t2r, out->ivssim_r2t);
    }
  }
  return QMIV_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
// =============================================================================================
// Batched frame sets: every view's work captured once into one HIP graph with concurrent
// branches, replayed per frame set.
// =============================================================================================
namespace qmiv {

// A few persistent host threads for a parallel loop over independent items (the host finalisation
// of a frame set's views); the calling thread takes items too.
class HostPool {
 public:
  explicit HostPool(int workers) {
    for (int i = 0; i < workers; i++) th_.emplace_back([this] { loop(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  // Runs f(0) .. f(n - 1), each once, and returns when all have finished.
  void run(int n, const std::function<void(int)>& f) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &f;
      n_ = n;
      next_.store(0);
      active_ = (int)th_.size();
      gen_++;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return active_ == 0; });
    job_ = nullptr;
  }

 private:
  void work() {
    for (int i = next_.fetch_add(1); i < n_; i = next_.fetch_add(1)) (*job_)(i);
  }
  void loop() {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      lk.unlock();
      work();
      lk.lock();
      if (--active_ == 0) done_cv_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* job_ = nullptr;
  int n_ = 0;
  std::atomic<int> next_{0};
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Pinned staging of host-to-device copies from pageable memory. A copy from pageable memory is
// split into slot-sized chunks; the pool's threads copy a chunk into the next pinned slot together
// and the slot goes to the device on the copy's stream while the host fills the following slot. A
// slot is refilled once its previous transfer has completed (its event).
struct Stager {
  static constexpr int kSlots = 4;
  static constexpr size_t kSlotBytes = (size_t)8 << 20;
  static constexpr size_t kPiece = (size_t)1 << 20;  // one thread's share of a chunk
  explicit Stager(int workers) : pool(workers) {}
  HostPool pool;
  void* slot[kSlots] = {};
  hipEvent_t ev[kSlots] = {};
  bool used[kSlots] = {};
  int next = 0;
};

void destroy_stager(Stager* s) {
  if (!s) return;
  for (int k = 0; k < Stager::kSlots; k++) {
    if (s->used[k]) (void)hipEventSynchronize(s->ev[k]);
    if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
    if (s->slot[k]) (void)hipHostFree(s->slot[k]);
  }
  delete s;
}

}  // namespace qmiv

namespace {

// Host-to-device copy of `bytes` from caller memory, ordered on stream st. Pinned memory goes to
// the copy engine directly (asynchronous: the caller keeps it unchanged until st has passed the
// copy); pageable memory through the context's pinned staging slots, and the source has been read
// in full when the call returns.
qmiv_status host_to_device(qmiv_ctx* ctx, void* dev, const void* src, size_t bytes, hipStream_t st) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, src) == hipSuccess && at.type == hipMemoryTypeHost) {
    QMIV_HIP(hipMemcpyAsync(dev, src, bytes, hipMemcpyHostToDevice, st));
    return QMIV_OK;
  }
  (void)hipGetLastError();  // an unregistered pointer leaves an error behind
  if (!ctx->stager) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    std::unique_ptr<Stager> s(new Stager((int)std::min(7u, hw - 1)));
    for (int k = 0; k < Stager::kSlots; k++) {
      if (hipHostMalloc(&s->slot[k], Stager::kSlotBytes, hipHostMallocDefault) != hipSuccess) {
        s->slot[k] = nullptr;
        destroy_stager(s.release());
        return fail(QMIV_ERR_OUT_OF_MEMORY, "pinned staging allocation failed");
      }
      if (hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming) != hipSuccess) {
        destroy_stager(s.release());
        return fail(QMIV_ERR_HIP, "staging event creation failed");
      }
    }
    ctx->stager = s.release();
  }
  Stager& S = *ctx->stager;
  for (size_t off = 0; off < bytes; off += Stager::kSlotBytes) {
    const size_t len = std::min(Stager::kSlotBytes, bytes - off);
    const int k = S.next;
    S.next = (k + 1) % Stager::kSlots;
    if (S.used[k]) QMIV_HIP(hipEventSynchronize(S.ev[k]));
    char* dst = (char*)S.slot[k];
    const char* from = (const char*)src + off;
    const int pieces = (int)((len + Stager::kPiece - 1) / Stager::kPiece);
    S.pool.run(pieces, [&](int i) {
      const size_t o = (size_t)i * Stager::kPiece;
      std::memcpy(dst + o, from + o, std::min(Stager::kPiece, len - o));
    });
    QMIV_HIP(hipMemcpyAsync((char*)dev + off, dst, len, hipMemcpyHostToDevice, st));
    QMIV_HIP(hipEventRecord(S.ev[k], st));
    S.used[k] = true;
  }
  return QMIV_OK;
}

// The masked views of a masked batch (mask pictures, parameters, plans), defined with the masked
// mode below.
struct MaskedBatchData;
qmiv_status batch_build_masked(qmiv_batch* b, int32_t branches);
void destroy_masked_batch(MaskedBatchData* m);
MaskedBatchData* copy_masked_batch(const MaskedBatchData* m);
const qmiv_params& masked_batch_params(const MaskedBatchData* m);

}  // namespace

struct qmiv_batch {
  std::unique_ptr<HostPool> pool;   // host finalisation of the views
  qmiv_ctx* ctx = nullptr;
  int32_t n = 0;
  std::vector<const qmiv_pic*> p0, p1;
  std::vector<qmiv_params> prm;
  std::vector<uint32_t> metrics;
  std::vector<FrameWork> fw;
  std::vector<FramePlan> plans;
  std::vector<size_t> doff, hoff;
  std::vector<int32_t> branch_of;   // view -> branch
  std::vector<size_t> scp_off;      // branch -> element offset of its SCP buffer in d_scp
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  void* h_results = nullptr;
  size_t results_bytes = 0;
  void* d_scp = nullptr;
  size_t scp_bytes = 0;
  std::vector<hipStream_t> streams;  // branch streams, default priority
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;  // recorded after a launched replay
  bool pending = false;       // a replay launched and not yet collected
  qmiv_batch_info info{};
  MaskedBatchData* mb = nullptr;  // a masked batch's views (null: an unmasked batch)
};

namespace {

constexpr int32_t kMaxBranches = 16;

void batch_drop_graph(qmiv_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Grows a device (or pinned host) buffer; the old contents are not kept.
qmiv_status batch_grow(void*& ptr, size_t& have, size_t need, bool host) {
  if (have >= need) return QMIV_OK;
  if (ptr) {
    if (host) QMIV_HIP(hipHostFree(ptr));
    else QMIV_HIP(hipFree(ptr));
  }
  ptr = nullptr;
  have = 0;
  const size_t sz = align_up(std::max<size_t>(need, 1), 1 << 16);
  const hipError_t e = host ? hipHostMalloc(&ptr, sz, hipHostMallocDefault) : hipMalloc(&ptr, sz);
  if (e != hipSuccess) return fail(QMIV_ERR_OUT_OF_MEMORY, "batch buffer allocation failed");
  have = sz;
  return QMIV_OK;
}

// Captures the graph of a batch whose views are planned, laid out and assigned to branches
// (b->branch_of): enqueue(i, st) enqueues view i on stream st, its branch's. The branches fork from
// and join into the context stream; the graph is instantiated.
qmiv_status batch_capture(qmiv_batch* b, int32_t branches,
                          const std::function<qmiv_status(size_t, hipStream_t)>& enqueue) {
  qmiv_ctx* ctx = b->ctx;
  while ((int32_t)b->streams.size() < branches) {
    hipStream_t s;
    QMIV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b->streams.push_back(s);
    hipEvent_t e;
    QMIV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->join.push_back(e);
  }
  if (!b->fork) QMIV_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));

  // Capture. Kernel timing is off inside the graph; enqueue errors end the capture first.
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  QMIV_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  qmiv_status st = QMIV_OK;
  if (hipEventRecord(b->fork, ctx->stream) != hipSuccess) st = fail(QMIV_ERR_HIP, "hipEventRecord (fork)");
  for (int32_t br = 0; br < branches && st == QMIV_OK; br++)
    if (hipStreamWaitEvent(b->streams[(size_t)br], b->fork, 0) != hipSuccess)
      st = fail(QMIV_ERR_HIP, "hipStreamWaitEvent (fork)");
  for (size_t i = 0; i < (size_t)b->n && st == QMIV_OK; i++) st = enqueue(i, b->streams[(size_t)b->branch_of[i]]);
  for (int32_t br = 0; br < branches && st == QMIV_OK; br++) {
    if (hipEventRecord(b->join[(size_t)br], b->streams[(size_t)br]) != hipSuccess ||
        hipStreamWaitEvent(ctx->stream, b->join[(size_t)br], 0) != hipSuccess)
      st = fail(QMIV_ERR_HIP, "graph join");
  }
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
  ctx->profiling = prof;
  if (st != QMIV_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (ee != hipSuccess) return fail(QMIV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
  b->graph = graph;
  QMIV_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->info.branches = branches;
  return QMIV_OK;
}

// Plans every view for (branches, sweep cap), assigns views to branches (longest work first onto
// the least loaded branch), lays out the buffers and captures the graph.
qmiv_status batch_build(qmiv_batch* b, int32_t branches, int32_t sweep) {
  if (b->mb) return batch_build_masked(b, branches);
  qmiv_ctx* ctx = b->ctx;
  batch_drop_graph(b);
  const size_t n = (size_t)b->n;
  b->plans.assign(n, FramePlan{});
  b->doff.assign(n, 0);
  b->hoff.assign(n, 0);
  size_t dtot = 0, htot = 0;
  std::vector<double> cost(n);
  for (size_t i = 0; i < n; i++) {
    b->plans[i] = plan_frame(ctx, b->p0[i]->g, b->prm[i], b->fw[i], sweep);
    b->doff[i] = dtot;
    b->hoff[i] = htot;
    dtot += align_up(b->plans[i].L.total, 256);
    htot += align_up(b->plans[i].L.results_bytes, 256);
    const Geom& g = b->p0[i]->g;
    cost[i] = (double)g.w * g.h * (1.0 + (b->fw[i].need_cps || b->fw[i].need_scp ? 4.0 : 0.0) +
                                   4.0 * (double)b->plans[i].L.ssim.size());
  }
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return cost[a] > cost[c]; });
  std::vector<double> load((size_t)branches, 0.0);
  std::vector<size_t> scp_need((size_t)branches, 0);
  b->branch_of.assign(n, 0);
  for (size_t i : order) {
    const size_t br = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
    b->branch_of[i] = (int32_t)br;
    load[br] += cost[i];
    if (b->fw[i].need_scp)
      scp_need[br] = std::max(scp_need[br], 6 * (size_t)b->p0[i]->g.w * b->p0[i]->g.h);
  }
  b->scp_off.assign((size_t)branches, 0);
  size_t stot = 0;
  for (int32_t br = 0; br < branches; br++) {
    b->scp_off[(size_t)br] = stot;
    stot += align_up(scp_need[(size_t)br], 128);
  }
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  QMIV_TRY(batch_grow(b->d_scratch, b->scratch_bytes, dtot, false));
  QMIV_TRY(batch_grow(b->h_results, b->results_bytes, htot, true));
  if (stot) QMIV_TRY(batch_grow(b->d_scp, b->scp_bytes, stot * sizeof(uint16_t), false));

  QMIV_TRY(batch_capture(b, branches, [&](size_t i, hipStream_t st) {
    const int32_t br = b->branch_of[i];
    uint16_t* scp = b->fw[i].need_scp ? (uint16_t*)b->d_scp + b->scp_off[(size_t)br] : nullptr;
    return enqueue_frame(ctx, b->p0[i], b->p1[i], b->prm[i], b->fw[i], b->plans[i], (char*)b->d_scratch + b->doff[i],
                         (char*)b->h_results + b->hoff[i], scp, st);
  }));
  b->info.branches = branches;
  b->info.cps_sweep = sweep;
  return QMIV_OK;
}

// Mean time of `reps` replays of the current graph (after one untimed), each launched and waited
// for on the host clock as qmiv_batch_run does: a replay's cost to a frame-set loop is its launch
// (enqueueing the graph's nodes on the host) and its run up to the join. Back-to-back replays
// between device events hide the launch behind the previous replay and rank candidates by a
// figure that such a loop never sees.
qmiv_status batch_time(qmiv_batch* b, int32_t reps, double& ms) {
  qmiv_ctx* ctx = b->ctx;
  reps = std::max(reps, 1);
  QMIV_HIP(hipGraphLaunch(b->exec, ctx->stream));
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  const auto t0 = std::chrono::steady_clock::now();
  for (int32_t r = 0; r < reps; r++) {
    QMIV_HIP(hipGraphLaunch(b->exec, ctx->stream));
    QMIV_HIP(hipStreamSynchronize(ctx->stream));
  }
  ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
  return QMIV_OK;
}

}  // namespace

// =============================================================================================
// Differentiable SSIM family on float planes.
// =============================================================================================
namespace {

qmiv_status ensure_fwork(qmiv_ctx* ctx, size_t bytes) {
  if (ctx->fwork_bytes >= bytes) return QMIV_OK;
  if (ctx->d_fwork) {
    QMIV_HIP(hipStreamSynchronize(ctx->stream));
    QMIV_HIP(hipFree(ctx->d_fwork));
    ctx->d_fwork = nullptr;
    ctx->fwork_bytes = 0;
  }
  const size_t sz = align_up(bytes, 1 << 20);
  if (hipMalloc(&ctx->d_fwork, sz) != hipSuccess) return fail(QMIV_ERR_OUT_OF_MEMORY, "float workspace allocation failed");
  ctx->fwork_bytes = sz;
  return QMIV_OK;
}

// Bump allocation inside the float workspace: sizes first (base = null), then pointers.
struct Bump {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t n) {
    T* p = base ? (T*)(base + off) : nullptr;
    off = align_up(off + n * sizeof(T), 256);
    return p;
  }
};

void ssimf_consts(double max_value, double& c1, double& c2, double f[100]) {
  const double k1 = 0.01 * max_value, k2 = 0.03 * max_value;
  c1 = k1 * k1;
  c2 = k2 * k2;
  fin::ssim_window(f);
}

// The gradient correlation (k_ssimf_grad) of the maps an enqueue_ssimf call computed for the same
// pictures, scales and max_value; g holds the outputs, their coefficients and accumulate.
qmiv_status enqueue_ssimf_grad(const void* tst, int32_t tst_f64, double scale_t, const void* ref, int32_t ref_f64,
                               double scale_r, int32_t w, int32_t h, double max_value, const double* maps,
                               SsimGArgs& g, hipStream_t st) {
  double c1, c2;
  ssimf_consts(max_value, c1, c2, g.f);
  g.tst = tst;
  g.ref = ref;
  g.tst_f64 = tst_f64;
  g.ref_f64 = ref_f64;
  g.scale_t = scale_t;
  g.scale_r = scale_r;
  g.w = w;
  g.h = h;
  g.maps = maps;
  QMIV_HIP(launch_ssimf_grad(g, st));
  return QMIV_OK;
}

// The row reduction of IV-SSIM's six evaluations in one launch: evaluation e (T2R components 0-2,
// then R2T components 0-2) has its tile partials at partials + e ssimf_tiles_x(w) h and its rows at
// rows + e h.
SsimRowsJobs ivssim_rows_jobs(int32_t w, int32_t h, const double* partials, double* rows) {
  SsimRowsJobs j{};
  j.n = 6;
  j.max_h = h;
  for (int e = 0; e < 6; e++) {
    j.tiles_x[e] = ssimf_tiles_x(w);
    j.h[e] = h;
    j.part[e] = partials + (size_t)e * ssimf_tiles_x(w) * h;
    j.rows[e] = rows + (size_t)e * h;
  }
  return j;
}

// One float SSIM evaluation on stream st (null: the context stream): row sums into rows_dev (h
// doubles; null: the tile partials are left for the caller to reduce), and when maps is not null the
// gradient maps; then, when g.grad_t or g.grad_r is set, the gradient.
qmiv_status enqueue_ssimf(qmiv_ctx* ctx, const void* tst, int32_t tst_f64, double scale_t, const void* ref,
                          int32_t ref_f64, double scale_r, int32_t w, int32_t h, int32_t calc_l, double max_value,
                          double* partials, double* rows_dev, double* maps, SsimGArgs* g, hipStream_t st = nullptr) {
  if (!st) st = ctx->stream;
  SsimFArgs a{};
  a.tst = tst;
  a.ref = ref;
  a.tst_f64 = tst_f64;
  a.ref_f64 = ref_f64;
  a.scale_t = scale_t;
  a.scale_r = scale_r;
  a.w = w;
  a.h = h;
  a.tiles_x = ssimf_tiles_x(w);
  ssimf_consts(max_value, a.c1, a.c2, a.f);
  a.partials = partials;
  a.maps = maps;
  QMIV_HIP(launch_ssimf(a, calc_l, maps ? 1 : 0, st));
  if (rows_dev) {
    SsimRowsJobs j{};
    j.n = 1;
    j.max_h = h;
    j.tiles_x[0] = a.tiles_x;
    j.h[0] = h;
    j.part[0] = partials;
    j.rows[0] = rows_dev;
    QMIV_HIP(launch_ssim_rows(j, st));
  }
  if (g && maps && (g->grad_t || g->grad_r))
    QMIV_TRY(enqueue_ssimf_grad(tst, tst_f64, scale_t, ref, ref_f64, scale_r, w, h, max_value, maps, *g, st));
  return QMIV_OK;
}

qmiv_status check_fargs(qmiv_ctx* ctx, int32_t w, int32_t h, int32_t planes, const float* tst, const float* ref,
                        double max_value) {
  if (!ctx || !tst || !ref) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (w <= 0 || h <= 0 || planes < 1 || planes > 4) return fail(QMIV_ERR_INVALID_ARG, "bad picture shape");
  if ((int64_t)w * h * planes >= ((int64_t)1 << 31)) return fail(QMIV_ERR_UNSUPPORTED, "picture too large");
  if (!(max_value > 0)) return fail(QMIV_ERR_INVALID_ARG, "max_value must be positive");
  return QMIV_OK;
}

// Quantised 4:4:4 picture of the context for the shift search (recreated on a geometry change).
qmiv_status fq_picture(qmiv_ctx* ctx, int k, int32_t w, int32_t h, int32_t bd, qmiv_pic** out) {
  if ((int)ctx->fq.size() <= k) ctx->fq.resize((size_t)k + 1, nullptr);
  qmiv_pic*& q = ctx->fq[(size_t)k];
  Geom g;
  QMIV_TRY(make_geom(w, h, bd, 444, g));
  if (q && !same_geom(q->g, g)) {
    QMIV_TRY(wait_stream(ctx));  // every stream that may read it
    (void)hipFree(q->dev);
    delete q;
    q = nullptr;
  }
  if (!q) {
    void* dev = nullptr;
    if (hipMalloc(&dev, align_up(g.frame_bytes, 256)) != hipSuccess)
      return fail(QMIV_ERR_OUT_OF_MEMORY, "quantised picture allocation failed");
    q = new qmiv_pic{ctx, g, dev, true};
  }
  *out = q;
  return QMIV_OK;
}

}  // namespace

extern "C" {

const char* qmiv_last_error(void) { return g_last_error.c_str(); }
int32_t qmiv_abi_version(void) { return QMIV_ABI_VERSION; }

void qmiv_params_default(qmiv_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->search_range = 2;
  p->cmp_weights_search[0] = 4;
  p->cmp_weights_search[1] = 1;
  p->cmp_weights_search[2] = 1;
  p->cmp_weights_average[0] = 4;
  p->cmp_weights_average[1] = 1;
  p->cmp_weights_average[2] = 1;
  p->unnoticeable_coef[0] = p->unnoticeable_coef[1] = p->unnoticeable_coef[2] = 0.01f;
  p->lon_range_deg = 360;
  p->lat_range_deg = 180;
  p->cps_int32_rows = 1;
  p->fake_psnr_for_exact = 1;
}

size_t qmiv_frame_bytes(int32_t width, int32_t height, int32_t bit_depth, int32_t chroma_format) {
  return file_frame_bytes(width, height, bit_depth, chroma_format);
}

qmiv_status qmiv_create(int32_t device, qmiv_ctx** out) {
  if (!out) return fail(QMIV_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  int n = 0;
  QMIV_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(QMIV_ERR_INVALID_ARG, "no such HIP device");
  QMIV_HIP(hipSetDevice(device));
  qmiv_ctx* ctx = new qmiv_ctx;
  ctx->device = device;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    ctx->num_cus = cus;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
  for (hipStream_t& a : ctx->aux_stream)
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming);
  // The allocation-free entry point's buffers: the device's fixed scratch (its address resolved with
  // this device current) and the pinned block its results return through.
  if (e == hipSuccess) {
    ctx->fixed_err = fixed_scratch_setup(device, &ctx->d_fixed);
    if (ctx->fixed_err == hipSuccess) ctx->fixed_bytes = fixed_scratch_bytes();
    else ctx->d_fixed = nullptr;
    e = hipHostMalloc(&ctx->h_fixed, kFixedResultsBytes, hipHostMallocDefault);
    if (e != hipSuccess) ctx->h_fixed = nullptr;
  }
  if (e != hipSuccess) {
    if (ctx->h_fixed) (void)hipHostFree(ctx->h_fixed);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    for (hipStream_t a : ctx->aux_stream)
      if (a) (void)hipStreamDestroy(a);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return fail(QMIV_ERR_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
  }
  *out = ctx;
  return QMIV_OK;
}

void qmiv_destroy(qmiv_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  for (hipStream_t a : ctx->aux_stream)
    if (a) (void)hipStreamSynchronize(a);
  for (hipEvent_t ev : ctx->copy_ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : ctx->done_ev) (void)hipEventDestroy(ev);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  if (ctx->d_scratch) (void)hipFree(ctx->d_scratch);
  if (ctx->d_scp) (void)hipFree(ctx->d_scp);
  if (ctx->d_fwork) (void)hipFree(ctx->d_fwork);
  for (qmiv_pic* q : ctx->fq)
    if (q) {
      (void)hipFree(q->dev);
      delete q;
    }
  if (ctx->h_results) (void)hipHostFree(ctx->h_results);
  if (ctx->h_fixed) (void)hipHostFree(ctx->h_fixed);
  destroy_stager(ctx->stager);
  for (auto& pr : ctx->ev_pool) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  for (hipStream_t a : ctx->aux_stream)
    if (a) (void)hipStreamDestroy(a);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

void* qmiv_stream(qmiv_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

qmiv_status qmiv_profile_enable(qmiv_ctx* ctx, int32_t on) {
  if (!ctx) return fail(QMIV_ERR_INVALID_ARG, "null context");
  ctx->profiling = on != 0;
  return QMIV_OK;
}

qmiv_status qmiv_profile_read(qmiv_ctx* ctx, qmiv_profile* out, int32_t reset) {
  if (!ctx || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  harvest_events(ctx);
  *out = ctx->prof;
  if (reset) ctx->prof = qmiv_profile{};
  return QMIV_OK;
}

qmiv_status qmiv_sync(qmiv_ctx* ctx) {
  if (!ctx) return fail(QMIV_ERR_INVALID_ARG, "null context");
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_create(qmiv_ctx* ctx, int32_t width, int32_t height, int32_t bit_depth, int32_t chroma_format,
                            qmiv_pic** out) {
  if (!ctx || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  Geom g;
  QMIV_TRY(make_geom(width, height, bit_depth, chroma_format, g));
  QMIV_HIP(hipSetDevice(ctx->device));
  void* dev = nullptr;
  if (hipMalloc(&dev, align_up(g.frame_bytes, 256)) != hipSuccess)
    return fail(QMIV_ERR_OUT_OF_MEMORY, "picture allocation failed");
  *out = new qmiv_pic{ctx, g, dev, true};
  return QMIV_OK;
}

qmiv_status qmiv_pic_wrap(qmiv_ctx* ctx, int32_t width, int32_t height, int32_t bit_depth, int32_t chroma_format,
                          void* device_ptr, qmiv_pic** out) {
  if (!ctx || !out || !device_ptr) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  Geom g;
  QMIV_TRY(make_geom(width, height, bit_depth, chroma_format, g));
  if (((uintptr_t)device_ptr) % (uintptr_t)g.bytes_per_sample)
    return fail(QMIV_ERR_INVALID_ARG, "device pointer not aligned to the sample size");
  *out = new qmiv_pic{ctx, g, device_ptr, false};
  return QMIV_OK;
}

void qmiv_pic_destroy(qmiv_pic* pic) {
  if (!pic) return;
  // hipFree waits for outstanding device work itself; the picture never touches its context
  // here, so pictures may outlive the context that created them.
  if (pic->owned && pic->dev) (void)hipFree(pic->dev);
  delete pic;
}

void* qmiv_pic_device_ptr(qmiv_pic* pic) { return pic ? pic->dev : nullptr; }

qmiv_status qmiv_pic_upload(qmiv_ctx* ctx, qmiv_pic* pic, const void* host, size_t bytes) {
  if (!ctx || !pic || !host) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (bytes != pic->g.frame_bytes) return fail(QMIV_ERR_INVALID_ARG, "byte count differs from the frame size");
  QMIV_HIP(hipSetDevice(ctx->device));
  return host_to_device(ctx, pic->dev, host, bytes, ctx->stream);
}

qmiv_status qmiv_pic_upload_unpack(qmiv_ctx* ctx, qmiv_pic* pic, const void* host, size_t bytes,
                                   int32_t file_chroma_format) {
  if (!ctx || !pic || !host) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (pic->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "picture belongs to another context");
  return upload_file_frame(ctx, pic, host, bytes, file_chroma_format);
}

}  // extern "C"

namespace {

// Checks shared by the reference-layout transfers: a 4:4:4 picture of this context, a stride of at
// least the width; 8-bit pictures hold samples up to 255 only.
qmiv_status check_ref_layout(qmiv_ctx* ctx, const qmiv_pic* pic, int32_t stride) {
  if (pic->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "picture belongs to another context");
  if (!is444(pic->g)) return fail(QMIV_ERR_INVALID_ARG, "reference-layout pictures are 4:4:4");
  if (stride < pic->g.w) return fail(QMIV_ERR_INVALID_ARG, "stride below the picture width");
  return QMIV_OK;
}

qmiv_status check_8bit_rows(const uint16_t* base, int32_t stride, int32_t w, int32_t h, int32_t step) {
  for (int32_t y = 0; y < h; y++) {
    const uint16_t* r = base + (size_t)y * stride * step;
    for (int32_t x = 0; x < w * step; x++)
      if ((step == 1 || x % 4 != 3) && r[x] > 255)
        return fail(QMIV_ERR_PEL_RANGE, "sample above 255 for an 8-bit picture");
  }
  return QMIV_OK;
}

// Per-row statistics of (a, b) into the context scratch, row maxima and SSDs copied back:
// host rows [0, n) row SSD (uint64), then row maxima of a (uint32).
qmiv_status picture_rows(qmiv_ctx* ctx, const qmiv_pic* a, const qmiv_pic* b, const uint64_t** ssd,
                         const uint32_t** maxa) {
  const Geom& g = a->g;
  const size_t n = (size_t)stat_rows(g);
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_TRY(ensure_scratch(ctx, 24 * n));
  QMIV_TRY(ensure_results(ctx, 12 * n));
  char* d = (char*)ctx->d_scratch;
  int64_t* row_sd = (int64_t*)d;
  uint64_t* row_ssd = (uint64_t*)(d + 8 * n);
  uint32_t* row_max = (uint32_t*)(d + 16 * n);
  QMIV_HIP(launch_stats(a->dev, b->dev, g, row_sd, row_ssd, row_max, ctx->stream));
  char* h = (char*)ctx->h_results;
  QMIV_HIP(hipMemcpyAsync(h, row_ssd, 8 * n, hipMemcpyDeviceToHost, ctx->stream));
  QMIV_HIP(hipMemcpyAsync(h + 8 * n, row_max, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  *ssd = (const uint64_t*)h;
  *maxa = (const uint32_t*)(h + 8 * n);
  return QMIV_OK;
}

// Rows of plane c in the statistics' row order (luma, then Cb, then Cr rows).
void plane_rows(const Geom& g, int c, size_t& r0, size_t& r1) {
  r0 = c == 0 ? 0 : (size_t)g.h + (size_t)(c - 1) * g.ch;
  r1 = c == 0 ? (size_t)g.h : r0 + (size_t)g.ch;
}

}  // namespace

extern "C" {

qmiv_status qmiv_pic_upload_planar(qmiv_ctx* ctx, qmiv_pic* pic, const uint16_t* const planes[3], int32_t stride) {
  if (!ctx || !pic || !planes || !planes[0] || !planes[1] || !planes[2]) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_ref_layout(ctx, pic, stride));
  const Geom& g = pic->g;
  QMIV_HIP(hipSetDevice(ctx->device));
  const size_t plane = (size_t)g.w * g.h;
  if (g.bytes_per_sample == 2) {
    for (int c = 0; c < 3; c++)
      QMIV_HIP(hipMemcpy2DAsync((uint16_t*)pic->dev + c * plane, 2 * (size_t)g.w, planes[c], 2 * (size_t)stride,
                                2 * (size_t)g.w, (size_t)g.h, hipMemcpyHostToDevice, ctx->stream));
    return QMIV_OK;
  }
  for (int c = 0; c < 3; c++) QMIV_TRY(check_8bit_rows(planes[c], stride, g.w, g.h, 1));
  QMIV_TRY(ensure_stage(ctx, 6 * plane));
  uint16_t* st = (uint16_t*)ctx->d_stage;
  for (int c = 0; c < 3; c++)
    QMIV_HIP(hipMemcpy2DAsync(st + c * plane, 2 * (size_t)g.w, planes[c], 2 * (size_t)stride, 2 * (size_t)g.w,
                              (size_t)g.h, hipMemcpyHostToDevice, ctx->stream));
  QMIV_HIP(launch_narrow(st, (int64_t)(3 * plane), (uint8_t*)pic->dev, ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_download_planar(qmiv_ctx* ctx, const qmiv_pic* pic, uint16_t* const planes[3], int32_t stride) {
  if (!ctx || !pic || !planes || !planes[0] || !planes[1] || !planes[2]) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_ref_layout(ctx, pic, stride));
  const Geom& g = pic->g;
  QMIV_HIP(hipSetDevice(ctx->device));
  const size_t plane = (size_t)g.w * g.h;
  const uint16_t* src = (const uint16_t*)pic->dev;
  if (g.bytes_per_sample == 1) {
    QMIV_TRY(ensure_stage(ctx, 6 * plane));
    QMIV_HIP(launch_widen((const uint8_t*)pic->dev, (int64_t)(3 * plane), (uint16_t*)ctx->d_stage, ctx->stream));
    src = (const uint16_t*)ctx->d_stage;
  }
  for (int c = 0; c < 3; c++)
    QMIV_HIP(hipMemcpy2DAsync(planes[c], 2 * (size_t)stride, src + c * plane, 2 * (size_t)g.w, 2 * (size_t)g.w,
                              (size_t)g.h, hipMemcpyDeviceToHost, ctx->stream));
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_upload_interleaved(qmiv_ctx* ctx, qmiv_pic* pic, const uint16_t* origin, int32_t stride) {
  if (!ctx || !pic || !origin) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_ref_layout(ctx, pic, stride));
  const Geom& g = pic->g;
  if (g.bytes_per_sample == 1) QMIV_TRY(check_8bit_rows(origin, stride, g.w, g.h, 4));
  QMIV_HIP(hipSetDevice(ctx->device));
  const size_t row = 8 * (size_t)g.w;  // 4 x uint16 per pel
  QMIV_TRY(ensure_stage(ctx, row * g.h));
  QMIV_HIP(hipMemcpy2DAsync(ctx->d_stage, row, origin, 8 * (size_t)stride, row, (size_t)g.h, hipMemcpyHostToDevice,
                            ctx->stream));
  QMIV_HIP(launch_interleave(pic->dev, g, (uint16_t*)ctx->d_stage, false, ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_download_interleaved(qmiv_ctx* ctx, const qmiv_pic* pic, uint16_t* origin, int32_t stride) {
  if (!ctx || !pic || !origin) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_ref_layout(ctx, pic, stride));
  const Geom& g = pic->g;
  QMIV_HIP(hipSetDevice(ctx->device));
  const size_t row = 8 * (size_t)g.w;
  QMIV_TRY(ensure_stage(ctx, row * g.h));
  QMIV_HIP(launch_interleave(pic->dev, g, (uint16_t*)ctx->d_stage, true, ctx->stream));
  QMIV_HIP(hipMemcpy2DAsync(origin, 8 * (size_t)stride, ctx->d_stage, row, row, (size_t)g.h, hipMemcpyDeviceToHost,
                            ctx->stream));
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_copy(qmiv_ctx* ctx, qmiv_pic* dst, const qmiv_pic* src) {
  if (!ctx || !dst || !src) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (!same_geom(dst->g, src->g)) return fail(QMIV_ERR_INVALID_ARG, "pictures differ in size, bit depth or chroma format");
  if (dst->ctx != ctx || src->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "pictures belong to another context");
  if (dst->dev == src->dev) return QMIV_OK;
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_HIP(hipMemcpyAsync(dst->dev, src->dev, src->g.frame_bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_check(qmiv_ctx* ctx, const qmiv_pic* pic, int32_t in_range[3]) {
  if (!ctx || !pic || !in_range) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (pic->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "picture belongs to another context");
  const Geom& g = pic->g;
  const uint64_t* ssd;
  const uint32_t* mx;
  QMIV_TRY(picture_rows(ctx, pic, pic, &ssd, &mx));
  const uint32_t maxv = (uint32_t)((1 << g.bd) - 1);
  for (int c = 0; c < 3; c++) {
    size_t r0, r1;
    plane_rows(g, c, r0, r1);
    uint32_t m = 0;
    for (size_t r = r0; r < r1; r++) m = std::max(m, mx[r]);
    // CheckIfInRange: every sample fits at 16 bits; a 4:0:0 picture has no chroma to check.
    in_range[c] = g.bd >= 16 || (c > 0 && g.cf == 400) || m <= maxv;
  }
  return QMIV_OK;
}

qmiv_status qmiv_pic_conceal(qmiv_ctx* ctx, qmiv_pic* pic) {
  if (!ctx || !pic) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (pic->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "picture belongs to another context");
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_HIP(launch_clip(pic->dev, pic->g, (1 << pic->g.bd) - 1, nullptr, ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_equal(qmiv_ctx* ctx, const qmiv_pic* a, const qmiv_pic* b, int32_t equal[3]) {
  if (!ctx || !equal) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (!a || !b) return fail(QMIV_ERR_INVALID_ARG, "null picture");
  if (!same_geom(a->g, b->g)) return fail(QMIV_ERR_INVALID_ARG, "pictures differ in size, bit depth or chroma format");
  if (a->ctx != ctx || b->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "pictures belong to another context");
  const uint64_t* ssd;
  const uint32_t* mx;
  QMIV_TRY(picture_rows(ctx, a, b, &ssd, &mx));
  for (int c = 0; c < 3; c++) {
    size_t r0, r1;
    plane_rows(a->g, c, r0, r1);
    uint64_t s = 0;
    for (size_t r = r0; r < r1; r++) s |= ssd[r];  // a sum of squares is 0 only if every term is
    equal[c] = s == 0;
  }
  return QMIV_OK;
}

qmiv_status qmiv_pic_download(qmiv_ctx* ctx, const qmiv_pic* pic, void* host, size_t bytes) {
  if (!ctx || !pic || !host) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (bytes != pic->g.frame_bytes) return fail(QMIV_ERR_INVALID_ARG, "byte count differs from the frame size");
  QMIV_HIP(hipMemcpyAsync(host, pic->dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  return QMIV_OK;
}

}  // extern "C"

qmiv_status qmiv::upload_file_frame(qmiv_ctx* ctx, qmiv_pic* pic, const void* host, size_t bytes, int32_t fcf) {
  const Geom& g = pic->g;
  bool unpack = false;
  QMIV_TRY(file_route(pic, g.w, g.h, g.bd, fcf, unpack));
  if (bytes != file_frame_bytes(g.w, g.h, g.bd, fcf))
    return fail(QMIV_ERR_INVALID_ARG, "byte count differs from the file frame size");
  QMIV_HIP(hipSetDevice(ctx->device));
  if (!unpack) return host_to_device(ctx, pic->dev, host, bytes, ctx->stream);
  QMIV_TRY(ensure_stage(ctx, bytes));
  QMIV_TRY(host_to_device(ctx, ctx->d_stage, host, bytes, ctx->stream));
  QMIV_HIP(launch_unpack(ctx->d_stage, bytes, g.w, g.h, g.bd, fcf, pic->dev, ctx->stream));
  return QMIV_OK;
}

namespace {
qmiv_status frames_impl(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                        const void* const* h0, const void* const* h1, const qmiv_params* prm_in, uint32_t metrics,
                        qmiv_frame_result* out);
}  // namespace

extern "C" {

qmiv_status qmiv_frame(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_params* prm_in,
                       uint32_t metrics, qmiv_frame_result* out) {
  const qmiv_pic* a[1] = {p0};
  const qmiv_pic* b[1] = {p1};
  return qmiv_frames(ctx, 1, a, b, prm_in, metrics, out);
}

qmiv_status qmiv_frames(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                        const qmiv_params* prm_in, uint32_t metrics, qmiv_frame_result* out) {
  if (!ctx || !out || !p0 || !p1 || n < 0) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  return frames_impl(ctx, n, p0, p1, nullptr, nullptr, prm_in, metrics, out);
}

qmiv_status qmiv_frames_upload(qmiv_ctx* ctx, int32_t n, qmiv_pic* const* p0, qmiv_pic* const* p1,
                               const void* const* h0, const void* const* h1, const qmiv_params* prm_in,
                               uint32_t metrics, qmiv_frame_result* out) {
  if (!ctx || !out || !p0 || !p1 || !h0 || !h1 || n < 0) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  // Every picture is written by its frame's upload while earlier frames compute: no picture may
  // serve two frames of the set.
  std::vector<const qmiv_pic*> all;
  all.reserve(2 * (size_t)n);
  for (int32_t i = 0; i < n; i++) {
    if (!h0[i] || !h1[i]) return fail(QMIV_ERR_INVALID_ARG, "null source frame");
    all.push_back(p0[i]);
    all.push_back(p1[i]);
  }
  std::sort(all.begin(), all.end());
  if (std::adjacent_find(all.begin(), all.end()) != all.end())
    return fail(QMIV_ERR_INVALID_ARG, "a picture appears twice in the frame set");
  std::vector<const qmiv_pic*> c0(p0, p0 + n), c1(p1, p1 + n);
  return frames_impl(ctx, n, c0.data(), c1.data(), h0, h1, prm_in, metrics, out);
}

}  // extern "C"

namespace {

// The frame-set loop of qmiv_frames and qmiv_frames_upload. h0 / h1 (nullable): packed host frames
// copied into p0[i] / p1[i] on the copy stream, each frame's kernels waiting for its own copies.
qmiv_status frames_impl(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                        const void* const* h0, const void* const* h1, const qmiv_params* prm_in, uint32_t metrics,
                        qmiv_frame_result* out) {
  if (n == 0) return QMIV_OK;
  for (int32_t i = 0; i < n; i++) QMIV_TRY(check_pair(p0[i], p1[i]));
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  if (prm.invalid_pel_action < QMIV_PEL_STOP || prm.invalid_pel_action > QMIV_PEL_CNCL)
    return fail(QMIV_ERR_INVALID_ARG, "invalid_pel_action must be 0..3");
  const FrameWork fw = frame_work(metrics, prm);
  // Plan every frame, then lay their scratch and result blocks end to end.
  std::vector<FramePlan> plans((size_t)n);
  std::vector<size_t> doff((size_t)n), hoff((size_t)n);
  size_t dtot = 0, htot = 0;
  for (int32_t i = 0; i < n; i++) {
    plans[(size_t)i] = plan_frame(ctx, p0[i]->g, prm, fw);
    doff[(size_t)i] = dtot;
    hoff[(size_t)i] = htot;
    dtot += align_up(plans[(size_t)i].L.total, 256);
    htot += align_up(plans[(size_t)i].L.results_bytes, 256);
  }
  QMIV_TRY(ensure_scratch(ctx, dtot));
  QMIV_TRY(ensure_results(ctx, htot));
  // Frames are dealt round-robin over the context stream and the aux streams; each stream's frames
  // run in order, so one set of shift-compensated pictures per stream serves them.
  const int32_t nstreams = std::min<int32_t>(n, 1 + qmiv_ctx::kAuxStreams);
  uint16_t* scp = nullptr;
  size_t scp_stride = 0;
  if (fw.need_scp) {
    size_t most = 0;
    for (int32_t i = 0; i < n; i++) most = std::max(most, (size_t)p0[i]->g.w * p0[i]->g.h);
    scp_stride = align_up(6 * most, 128);
    QMIV_TRY(ensure_scp(ctx, scp_stride * nstreams * sizeof(uint16_t)));
    scp = (uint16_t*)ctx->d_scp;
  }
  while ((int32_t)ctx->done_ev.size() < n) {
    hipEvent_t e;
    QMIV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->done_ev.push_back(e);
  }
  while (h0 && (int32_t)ctx->copy_ev.size() < n) {
    hipEvent_t e;
    QMIV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->copy_ev.push_back(e);
  }
  const hipStream_t streams[1 + qmiv_ctx::kAuxStreams] = {ctx->stream, ctx->aux_stream[0], ctx->aux_stream[1]};
  if (nstreams > 1 || h0) {  // the aux and copy streams start after everything on the context stream
    QMIV_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
    for (int32_t k = 1; k < nstreams; k++) QMIV_HIP(hipStreamWaitEvent(streams[k], ctx->fork_ev, 0));
    if (h0) QMIV_HIP(hipStreamWaitEvent(ctx->copy_stream, ctx->fork_ev, 0));
  }
  // Each frame is finalised on the host as soon as its results have arrived, while the device
  // works on the frames after it: between two enqueues, every frame whose results are in.
  qmiv_status first_err = QMIV_OK;
  int32_t done = 0;
  auto finalize_next = [&]() {
    const int32_t i = done++;
    const qmiv_status st = finalize_frame(ctx, p0[i]->g, prm, metrics, fw, plans[(size_t)i].L,
                                          (const char*)ctx->h_results + hoff[(size_t)i], &out[i]);
    if (st != QMIV_OK && first_err == QMIV_OK) first_err = st;
  };
  for (int32_t i = 0; i < n; i++) {
    const int32_t k = i % nstreams;
    qmiv_status st = QMIV_OK;
    if (h0) {  // xSeq::readFrame per frame (libs/pmbb-core/src/xSeq.cpp:155-174), here an async copy
      st = host_to_device(ctx, p0[i]->dev, h0[i], p0[i]->g.frame_bytes, ctx->copy_stream);
      if (st == QMIV_OK) st = host_to_device(ctx, p1[i]->dev, h1[i], p1[i]->g.frame_bytes, ctx->copy_stream);
      if (st == QMIV_OK) {
        hipError_t e = hipEventRecord(ctx->copy_ev[(size_t)i], ctx->copy_stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(streams[k], ctx->copy_ev[(size_t)i], 0);
        if (e != hipSuccess) st = fail(QMIV_ERR_HIP, std::string("frame upload: ") + hipGetErrorString(e));
      }
    }
    if (st == QMIV_OK)
      st = enqueue_frame(ctx, p0[i], p1[i], prm, fw, plans[(size_t)i], (char*)ctx->d_scratch + doff[(size_t)i],
                         (char*)ctx->h_results + hoff[(size_t)i], scp ? scp + scp_stride * (size_t)k : nullptr,
                         streams[k], nullptr, nullptr, ctx->done_ev[(size_t)i]);
    if (st != QMIV_OK) {
      // Drain what was enqueued: scratch and result buffers must be idle before any later reuse.
      (void)hipStreamSynchronize(ctx->stream);
      for (hipStream_t a : ctx->aux_stream) (void)hipStreamSynchronize(a);
      (void)hipStreamSynchronize(ctx->copy_stream);
      return st;
    }
    while (done < i && hipEventQuery(ctx->done_ev[(size_t)done]) == hipSuccess) finalize_next();
  }
  for (int32_t k = 1; k < nstreams; k++) {  // later work on the context stream follows the whole set
    QMIV_HIP(hipEventRecord(ctx->join_ev, streams[k]));
    QMIV_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0));
  }
  while (done < n) {
    QMIV_HIP(hipEventSynchronize(ctx->done_ev[(size_t)done]));
    finalize_next();
  }
  QMIV_TRY(wait_stream(ctx));
  return first_err;
}

}  // namespace

namespace {

// Tunes (unless the options fix the shape), builds and checks the batch's graph; *out receives b, or
// b is destroyed on failure. views_sig: the views' geometry, metrics and parameters.
qmiv_status batch_tune_finish(qmiv_batch* b, const qmiv_batch_options& opt, const std::string& views_sig,
                              qmiv_batch** out) {
  qmiv_ctx* ctx = b->ctx;
  const int32_t n = b->n;
  b->info.views = n;
  // Candidates: branch counts first (with the occupancy sweep rule), then sweep caps (unmasked
  // batches only: the masked search takes its own).
  std::vector<int32_t> br_cands, sw_cands;
  if (opt.branches > 0) br_cands = {std::min(opt.branches, n)};
  else
    for (int32_t c : {1, 2, 4, 8, 16})
      if (c <= n) br_cands.push_back(c);
  if (b->mb) sw_cands = {0};
  else if (opt.cps_sweep >= 0) sw_cands = {opt.cps_sweep};
  else sw_cands = {0, 1, 2, 4};
  int32_t best_br = br_cands[0], best_sw = sw_cands[0];
  bool tune = br_cands.size() > 1 || sw_cands.size() > 1;
  double best_ms = 0.0;
  qmiv_status st = QMIV_OK;
  // The signature: which knobs are tuned, the batch's kind and the views' own signature.
  std::string sig;
  {
    const int32_t knobs[3] = {opt.branches, opt.cps_sweep, b->mb ? 1 : 0};
    sig.append((const char*)knobs, sizeof(knobs));
    sig += views_sig;
  }
  if (tune)
    for (const auto& e : ctx->batch_tuned)
      if (e.sig == sig) {
        best_br = e.branches;
        best_sw = e.sweep;
        best_ms = e.ms;
        b->info.from_cache = 1;
        tune = false;
        break;
      }
  // A captured graph's replay time depends on its instance as well as its shape: instances of one
  // shape replay in one of a few distinct times (up to 25% apart for 24 views in 16 branches),
  // each instance always in the same one, as the runtime lays the instance's branches over the
  // device's hardware queues. Candidates are therefore ranked by their fastest instance, and the
  // instance kept at the end is checked against that figure.
  if (tune) {
    // Every candidate is built and timed once per round, in kRounds interleaved rounds, and ranked
    // by its fastest round. The candidates are built in a batch of their own (own streams and
    // scratch) that ends with the tuning: the batch kept starts from fresh streams.
    constexpr int kRounds = 8;
    b->info.tune_rounds = kRounds;
    std::unique_ptr<qmiv_batch, void (*)(qmiv_batch*)> tb(new qmiv_batch, qmiv_batch_destroy);
    tb->ctx = ctx;
    tb->n = n;
    tb->p0 = b->p0;
    tb->p1 = b->p1;
    tb->prm = b->prm;
    tb->metrics = b->metrics;
    tb->fw = b->fw;
    if (b->mb) tb->mb = copy_masked_batch(b->mb);
    auto time_one = [&](int32_t br, int32_t sw, double& fastest_t) -> qmiv_status {
      QMIV_TRY(batch_build(tb.get(), br, sw));
      double t = 0.0;
      QMIV_TRY(batch_time(tb.get(), opt.tune_reps, t));
      b->info.candidates_tried++;
      fastest_t = std::min(fastest_t, t);
      return QMIV_OK;
    };
    // Warm-up: replays of the first candidate until about 30 ms of work have run, so that the
    // device clocks have left their idle state before the first timed round.
    st = batch_build(tb.get(), br_cands[0], sw_cands[0]);
    double warm = 0.0;
    for (int k = 0; k < 64 && st == QMIV_OK && warm < 30.0; k++) {
      double t = 0.0;
      st = batch_time(tb.get(), opt.tune_reps, t);
      warm += t * opt.tune_reps;
    }
    std::vector<double> br_t(br_cands.size(), 1e300);
    for (int round = 0; round < kRounds && st == QMIV_OK; round++)
      for (size_t i = 0; i < br_cands.size() && st == QMIV_OK; i++) st = time_one(br_cands[i], sw_cands[0], br_t[i]);
    const double fastest = *std::min_element(br_t.begin(), br_t.end());
    // More branches unless fewer are more than 3% faster: the choice does not flip on noise.
    for (size_t i = 0; i < br_cands.size() && st == QMIV_OK; i++)
      if (br_t[i] <= 1.03 * fastest) {
        best_br = br_cands[i];
        best_ms = br_t[i];
      }
    if (sw_cands.size() > 1 && st == QMIV_OK) {
      std::vector<double> sw_t(sw_cands.size(), 1e300);
      for (int round = 0; round < kRounds && st == QMIV_OK; round++)
        for (size_t k = 0; k < sw_cands.size() && st == QMIV_OK; k++) st = time_one(best_br, sw_cands[k], sw_t[k]);
      if (st == QMIV_OK) {
        // The occupancy rule unless another cap is more than 3% faster.
        const double rule = std::min(sw_t[0], best_ms);
        best_ms = rule;
        best_sw = sw_cands[0];
        for (size_t k = 1; k < sw_cands.size(); k++)
          if (sw_t[k] < best_ms && sw_t[k] < rule / 1.03) {
            best_ms = sw_t[k];
            best_sw = sw_cands[k];
          }
      }
    }
    if (st == QMIV_OK) ctx->batch_tuned.push_back({sig, best_br, best_sw, best_ms});
  }
  if (st == QMIV_OK) st = batch_build(b, best_br, best_sw);
  // Instance check (tuned or cached shapes): a new instance while this one replays more than 8%
  // slower than the shape's fastest, at most kInstances builds in all.
  if (st == QMIV_OK && best_ms > 0.0) {
    constexpr int kInstances = 6;
    for (int k = 1; st == QMIV_OK; k++) {
      double t = 0.0;
      st = batch_time(b, opt.tune_reps, t);
      b->info.instances = k;
      b->info.instance_ms = t;
      if (st != QMIV_OK || t <= 1.08 * best_ms || k == kInstances) break;
      st = batch_build(b, best_br, best_sw);
    }
  }
  if (st != QMIV_OK) {
    qmiv_batch_destroy(b);
    return st;
  }
  b->info.tuned_ms = best_ms;
  *out = b;
  return QMIV_OK;
}

}  // namespace

extern "C" {

qmiv_status qmiv_host_alloc(qmiv_ctx* ctx, size_t bytes, void** ptr) {
  if (!ctx || !ptr || bytes == 0) return fail(QMIV_ERR_INVALID_ARG, "null argument or zero size");
  *ptr = nullptr;
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return QMIV_OK;
}

// Reads nothing of the context: pinned memory may outlive the context it was allocated through.
qmiv_status qmiv_host_free(qmiv_ctx* ctx, void* ptr) {
  (void)ctx;
  if (!ptr) return QMIV_OK;
  QMIV_HIP(hipHostFree(ptr));
  return QMIV_OK;
}

qmiv_status qmiv_frame_noalloc(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_params* prm_in,
                               uint32_t metrics, qmiv_frame_result* out) {
  if (!ctx || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(p0, p1));
  if (!ctx->d_fixed)
    return fail(QMIV_ERR_HIP, std::string("fixed scratch unavailable on this device: ") + hipGetErrorString(ctx->fixed_err));
  if (!ctx->h_fixed) return fail(QMIV_ERR_HIP, "no pinned results block on this context");
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  if (prm.invalid_pel_action < QMIV_PEL_STOP || prm.invalid_pel_action > QMIV_PEL_CNCL)
    return fail(QMIV_ERR_INVALID_ARG, "invalid_pel_action must be 0..3");
  const FrameWork fw = frame_work(metrics, prm);
  const Geom& g = p0->g;
  const FramePlan P = plan_frame(ctx, g, prm, fw);
  // Device block: the frame's scratch, then (IV-SSIM) the two shift-compensated pictures.
  const size_t scratch = align_up(P.L.total, 256);
  const size_t scp_bytes = fw.need_scp ? 6 * (size_t)g.w * g.h * sizeof(uint16_t) : 0;
  if (scratch + scp_bytes > ctx->fixed_bytes)
    return fail(QMIV_ERR_UNSUPPORTED, "frame needs " + std::to_string(scratch + scp_bytes) +
                                          " bytes of device scratch; the allocation-free entry point holds " +
                                          std::to_string(ctx->fixed_bytes));
  if (P.L.results_bytes > kFixedResultsBytes)
    return fail(QMIV_ERR_UNSUPPORTED, "frame needs " + std::to_string(P.L.results_bytes) +
                                          " bytes of results; the allocation-free entry point holds " +
                                          std::to_string(kFixedResultsBytes));
  char* dbase = (char*)ctx->d_fixed;
  uint16_t* scp = fw.need_scp ? (uint16_t*)(dbase + scratch) : nullptr;
  // Nothing in the block needs clearing: every kernel of the flow writes all it reads later.
  const qmiv_status st =
      enqueue_frame(ctx, p0, p1, prm, fw, P, dbase, (char*)ctx->h_fixed, scp, ctx->stream);
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  harvest_events(ctx);
  QMIV_TRY(st);
  return finalize_frame(ctx, g, prm, metrics, fw, P.L, (const char*)ctx->h_fixed, out);
}

qmiv_status qmiv_calc_gcd(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm_in,
                          int32_t gcd[4]) {
  if (!ctx || !gcd) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(tst, ref));
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  FrameWork fw;
  fw.need_gcd = true;
  Layout L;
  bool fast;
  // sum(Ref - Tst): the statistics pass computes sum(P1 - P0) with P0 = Tst, P1 = Ref.
  QMIV_TRY(run_frame_kernels(ctx, tst, ref, prm, fw, L, fast));
  const FrameSummary* sum = (const FrameSummary*)((const char*)ctx->h_results + L.off_summary);
  for (int c = 0; c < 4; c++) gcd[c] = sum->gcd[c];
  return QMIV_OK;
}

qmiv_status qmiv_calc_psnr(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, double psnr[3]) {
  if (!ctx || !psnr) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(tst, ref));
  qmiv_params prm;
  qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  FrameWork fw;
  Layout L;
  bool fast;
  QMIV_TRY(run_frame_kernels(ctx, tst, ref, prm, fw, L, fast));
  const FrameSummary* sum = (const FrameSummary*)((const char*)ctx->h_results + L.off_summary);
  const int32_t area = tst->g.w * tst->g.h;
  for (int c = 0; c < 3; c++) psnr[c] = fin::psnr_from_ssd((double)sum->ssd[c], area, tst->g.bd);
  return QMIV_OK;
}

qmiv_status qmiv_calc_wspsnr(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm_in,
                             double wspsnr[3]) {
  if (!ctx || !wspsnr) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(tst, ref));
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  FrameWork fw;
  Layout L;
  bool fast;
  QMIV_TRY(run_frame_kernels(ctx, tst, ref, prm, fw, L, fast));
  const char* hb = (const char*)ctx->h_results;
  const FrameSummary* sum = (const FrameSummary*)(hb + L.off_summary);
  const uint64_t* h_ssd = (const uint64_t*)(hb + L.off_row_ssd);
  const Geom& g = tst->g;
  const int32_t area = g.w * g.h;
  if (!prm.erp) {
    for (int c = 0; c < 3; c++) wspsnr[c] = fin::psnr_from_ssd((double)sum->ssd[c], area, g.bd);
    return QMIV_OK;
  }
  const qmiv_ctx::WsTable& wst = ws_table(ctx, g.h, prm.lat_range_deg);
  const std::vector<double>& ws = wst.w;
  const double corr = wst.corr;
  std::vector<uint64_t> r;
  for (int c = 0; c < 3; c++) {
    rows444(g, h_ssd, c, r);
    wspsnr[c] = fin::wspsnr_from_rows(r.data(), g.h, ws.data(), corr, area, g.bd);
  }
  return QMIV_OK;
}

qmiv_status qmiv_calc_ivpsnr(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm_in,
                             const int32_t gcd_ref2tst[4], double* ivpsnr, double* r2t, double* t2r,
                             uint64_t* rows_out) {
  if (!ctx || !gcd_ref2tst) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(tst, ref));
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  FrameWork fw;
  fw.gcd = gcd_ref2tst;
  fw.need_cps = true;
  Layout L;
  bool fast;
  // R2T searches Tst + gcd in Ref: the kernel's direction 0 with P0 = Tst, P1 = Ref.
  QMIV_TRY(run_frame_kernels(ctx, tst, ref, prm, fw, L, fast));
  const char* hb = (const char*)ctx->h_results;
  const Geom& g = tst->g;
  std::vector<uint64_t> rows((const uint64_t*)(hb + L.off_cps_rows), (const uint64_t*)(hb + L.off_cps_rows) + 6 * (size_t)g.h);
  apply_row_semantics(rows.data(), rows.size(), prm.cps_int32_rows != 0);
  static const std::vector<double> kNoWs;
  const qmiv_ctx::WsTable* wst = prm.erp ? &ws_table(ctx, g.h, prm.lat_range_deg) : nullptr;
  const std::vector<double>& ws = wst ? wst->w : kNoWs;  // IV-PSNR's WS sums take no H / sum(w) correction
  const double* wsp = prm.erp ? ws.data() : nullptr;
  const int32_t area = g.w * g.h;
  const double q0 = fin::qual_from_rows(rows.data(), g.h, wsp, area, g.bd, prm.cmp_weights_average);
  const double q1 = fin::qual_from_rows(rows.data() + 3 * (size_t)g.h, g.h, wsp, area, g.bd, prm.cmp_weights_average);
  if (r2t) *r2t = q0;
  if (t2r) *t2r = q1;
  if (ivpsnr) *ivpsnr = fin::min2(q0, q1);
  if (rows_out) std::memcpy(rows_out, rows.data(), rows.size() * sizeof(uint64_t));
  return QMIV_OK;
}

// Parameters of a processor-level call through the frame flow: no validation.
static qmiv_params processor_params(const qmiv_params* prm) {
  qmiv_params p;
  if (prm) p = *prm;
  else qmiv_params_default(&p);
  p.invalid_pel_action = QMIV_PEL_SKIP;
  return p;
}

qmiv_status qmiv_calc_ssim(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, int32_t calc_l, double ssim[3]) {
  if (!ssim) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  qmiv_frame_result r;
  const uint32_t m = QMIV_SSIM | (calc_l ? 0u : kSsimNoLuminance);
  const qmiv_params prm = processor_params(nullptr);
  QMIV_TRY(qmiv_frame(ctx, tst, ref, &prm, m, &r));
  for (int c = 0; c < 3; c++) ssim[c] = r.ssim[c];
  return QMIV_OK;
}

qmiv_status qmiv_calc_msssim(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, double msssim[3],
                             double* scales) {
  if (!msssim) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  qmiv_frame_result r;
  const qmiv_params prm = processor_params(nullptr);
  QMIV_TRY(qmiv_frame(ctx, tst, ref, &prm, QMIV_MSSSIM, &r));
  for (int c = 0; c < 3; c++) msssim[c] = r.msssim[c];
  if (scales) std::memcpy(scales, r.msssim_scales, sizeof(r.msssim_scales));
  return QMIV_OK;
}

qmiv_status qmiv_calc_ivssim(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm,
                             double* ivssim, double* r2t, double* t2r) {
  qmiv_frame_result r;
  const qmiv_params p = processor_params(prm);
  QMIV_TRY(qmiv_frame(ctx, tst, ref, &p, QMIV_IVSSIM, &r));
  if (ivssim) *ivssim = r.ivssim;
  if (r2t) *r2t = r.ivssim_r2t;
  if (t2r) *t2r = r.ivssim_t2r;
  return QMIV_OK;
}

qmiv_status qmiv_gen_scp(qmiv_ctx* ctx, const qmiv_pic* src_ref, const qmiv_pic* src_tst, const qmiv_params* prm_in,
                         const int32_t gcd_ref2tst[4], uint16_t* dst_ref, uint16_t* dst_tst) {
  if (!ctx || !gcd_ref2tst || !dst_ref || !dst_tst) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(src_tst, src_ref));
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  FrameWork fw;
  fw.gcd = gcd_ref2tst;
  fw.need_scp = true;
  Layout L;
  bool fast;
  // DstRef is rebuilt from SrcTst + gcd searched in SrcRef: the kernel's direction 0 with
  // P0 = SrcTst, P1 = SrcRef, i.e. its SCP of P1; DstTst is its SCP of P0.
  QMIV_TRY(run_frame_kernels(ctx, src_tst, src_ref, prm, fw, L, fast));
  const size_t plane3 = 3 * (size_t)src_tst->g.w * src_tst->g.h;
  const uint16_t* scp0 = (const uint16_t*)ctx->d_scp;
  QMIV_HIP(hipMemcpyAsync(dst_tst, scp0, plane3 * sizeof(uint16_t), hipMemcpyDeviceToHost, ctx->stream));
  QMIV_HIP(hipMemcpyAsync(dst_ref, scp0 + plane3, plane3 * sizeof(uint16_t), hipMemcpyDeviceToHost, ctx->stream));
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_gen_scp_pics(qmiv_ctx* ctx, const qmiv_pic* src_ref, const qmiv_pic* src_tst, const qmiv_params* prm_in,
                              const int32_t gcd_ref2tst[4], qmiv_pic* dst_ref, qmiv_pic* dst_tst) {
  if (!ctx || !gcd_ref2tst || !dst_ref || !dst_tst) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(src_tst, src_ref));
  const Geom& gs = src_tst->g;
  for (const qmiv_pic* d : {(const qmiv_pic*)dst_ref, (const qmiv_pic*)dst_tst})
    if (d->g.w != gs.w || d->g.h != gs.h || d->g.bd != gs.bd || !is444(d->g))
      return fail(QMIV_ERR_INVALID_ARG, "destination pictures must be 4:4:4 of the sources' size and bit depth");
  if (dst_ref == dst_tst) return fail(QMIV_ERR_INVALID_ARG, "destination pictures must differ");
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  FrameWork fw;
  fw.gcd = gcd_ref2tst;
  fw.need_scp = true;
  Layout L;
  bool fast;
  QMIV_TRY(run_frame_kernels(ctx, src_tst, src_ref, prm, fw, L, fast));
  const int64_t plane3 = 3 * (int64_t)gs.w * gs.h;
  const uint16_t* scp0 = (const uint16_t*)ctx->d_scp;
  // The frame kernels' SCP of P0 (= SrcTst) is DstTst; their SCP of P1 (= SrcRef) is DstRef.
  if (gs.bytes_per_sample == 2) {
    QMIV_HIP(hipMemcpyAsync(dst_tst->dev, scp0, plane3 * 2, hipMemcpyDeviceToDevice, ctx->stream));
    QMIV_HIP(hipMemcpyAsync(dst_ref->dev, scp0 + plane3, plane3 * 2, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    QMIV_HIP(launch_narrow(scp0, plane3, (uint8_t*)dst_tst->dev, ctx->stream));
    QMIV_HIP(launch_narrow(scp0 + plane3, plane3, (uint8_t*)dst_ref->dev, ctx->stream));
  }
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_calc_ivssim_scp(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* tst_scp,
                                 const qmiv_pic* ref_scp, const qmiv_params* prm_in, double* ivssim, double* r2t,
                                 double* t2r) {
  if (!ctx) return fail(QMIV_ERR_INVALID_ARG, "null context");
  QMIV_TRY(check_pair(tst, ref));
  QMIV_TRY(check_pair(tst, ref_scp));
  QMIV_TRY(check_pair(ref, tst_scp));
  const qmiv_params prm = processor_params(prm_in);
  qmiv_frame_result a, b;
  QMIV_TRY(qmiv_frame(ctx, tst, ref_scp, &prm, QMIV_SSIM, &a));  // SSIMs_T2R = calcPicSSIM(Tst, RefSCP)
  QMIV_TRY(qmiv_frame(ctx, ref, tst_scp, &prm, QMIV_SSIM, &b));  // SSIMs_R2T = calcPicSSIM(Ref, TstSCP)
  const double v_t2r = fin::weighted_average4(a.ssim, prm.cmp_weights_average);
  const double v_r2t = fin::weighted_average4(b.ssim, prm.cmp_weights_average);
  if (ivssim) *ivssim = fin::min2(v_t2r, v_r2t);
  if (r2t) *r2t = v_r2t;
  if (t2r) *t2r = v_t2r;
  return QMIV_OK;
}

qmiv_status qmiv_pic_upsample(qmiv_ctx* ctx, const qmiv_pic* src, qmiv_pic* dst) {
  if (!ctx || !src || !dst) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (dst->g.w != src->g.w || dst->g.h != src->g.h || dst->g.bd != src->g.bd || !is444(dst->g))
    return fail(QMIV_ERR_INVALID_ARG, "destination must be a 4:4:4 picture of the source's size and bit depth");
  if (src->g.cf == 400) return fail(QMIV_ERR_INVALID_ARG, "a 4:0:0 source has no chroma to upsample");
  if (dst == src || dst->dev == src->dev) return fail(QMIV_ERR_INVALID_ARG, "destination must differ from the source");
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_HIP(launch_upsample(src->dev, src->g, dst->dev, ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_colorspace(qmiv_ctx* ctx, const qmiv_pic* src, qmiv_pic* dst, int32_t space, int32_t to_rgb) {
  if (!ctx || !src || !dst) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (src->ctx != ctx || dst->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "pictures belong to another context");
  if (!is444(src->g) || !is444(dst->g)) return fail(QMIV_ERR_INVALID_ARG, "colour conversion needs 4:4:4 pictures");
  if (dst->g.w != src->g.w || dst->g.h != src->g.h || dst->g.bd != src->g.bd)
    return fail(QMIV_ERR_INVALID_ARG, "destination must match the source's size and bit depth");
  if (to_rgb != 0 && to_rgb != 1) return fail(QMIV_ERR_INVALID_ARG, "to_rgb must be 0 or 1");
  const int64_t bytes = (int64_t)src->g.frame_bytes;
  const char* s0 = (const char*)src->dev;
  const char* d0 = (const char*)dst->dev;
  if (d0 != s0 && d0 < s0 + bytes && s0 < d0 + bytes)
    return fail(QMIV_ERR_INVALID_ARG, "destination overlaps the source without being the same picture");
  ColourArgs k;
  if (!fin::colour_matrix(space, to_rgb != 0, k.m)) return fail(QMIV_ERR_INVALID_ARG, "colour space must be 0..3");
  k.mid = 1 << (src->g.bd - 1);
  k.maxv = (1 << src->g.bd) - 1;
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_HIP(launch_colour(src->dev, dst->dev, src->g, k, to_rgb != 0, ctx->stream));
  return QMIV_OK;
}

qmiv_status qmiv_pic_swap_planes(qmiv_ctx* ctx, qmiv_pic* pic, int32_t a, int32_t b) {
  if (!ctx || !pic) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (pic->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "picture belongs to another context");
  if (!is444(pic->g)) return fail(QMIV_ERR_INVALID_ARG, "plane exchange needs a 4:4:4 picture");
  if (a < 0 || a > 2 || b < 0 || b > 2) return fail(QMIV_ERR_INVALID_ARG, "plane index must be 0..2");
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_HIP(launch_swap_planes(pic->dev, pic->g, a, b, ctx->stream));
  return QMIV_OK;
}


void qmiv_batch_options_default(qmiv_batch_options* o) {
  if (!o) return;
  o->branches = 0;
  o->cps_sweep = -1;
  o->tune_reps = 12;
}

qmiv_status qmiv_batch_create(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                              const qmiv_params* prm, const uint32_t* metrics, const qmiv_batch_options* opt_in,
                              qmiv_batch** out) {
  if (!ctx || !p0 || !p1 || !metrics || !out || n <= 0) return fail(QMIV_ERR_INVALID_ARG, "null argument or empty batch");
  *out = nullptr;
  qmiv_batch_options opt;
  qmiv_batch_options_default(&opt);
  if (opt_in) opt = *opt_in;
  if (opt.branches < 0 || opt.branches > kMaxBranches) return fail(QMIV_ERR_INVALID_ARG, "branches must be 0..16");
  if (opt.cps_sweep < -1) return fail(QMIV_ERR_INVALID_ARG, "cps_sweep must be -1, 0 or positive");
  if (opt.tune_reps <= 0) opt.tune_reps = 12;
  for (int32_t i = 0; i < n; i++) {
    QMIV_TRY(check_pair(p0[i], p1[i]));
    if (p0[i]->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "pictures belong to another context");
    if (metrics[i] & ~(uint32_t)(QMIV_PSNR | QMIV_WSPSNR | QMIV_IVPSNR | QMIV_SSIM | QMIV_MSSSIM | QMIV_IVSSIM))
      return fail(QMIV_ERR_INVALID_ARG, "unknown metric bits");
  }
  QMIV_HIP(hipSetDevice(ctx->device));
  qmiv_batch* b = new qmiv_batch;
  b->ctx = ctx;
  b->n = n;
  b->p0.assign(p0, p0 + n);
  b->p1.assign(p1, p1 + n);
  b->prm.resize((size_t)n);
  b->metrics.assign(metrics, metrics + n);
  b->fw.resize((size_t)n);
  for (int32_t i = 0; i < n; i++) {
    if (prm) b->prm[(size_t)i] = prm[i];
    else qmiv_params_default(&b->prm[(size_t)i]);
    const qmiv_params& pi = b->prm[(size_t)i];
    if (pi.invalid_pel_action < QMIV_PEL_STOP || pi.invalid_pel_action > QMIV_PEL_CNCL) {
      delete b;
      return fail(QMIV_ERR_INVALID_ARG, "invalid_pel_action must be 0..3");
    }
    b->fw[(size_t)i] = frame_work(metrics[i] & (QMIV_PSNR | QMIV_WSPSNR | QMIV_IVPSNR | QMIV_SSIM | QMIV_MSSSIM | QMIV_IVSSIM), pi);
  }
  // The batch signature: every view's geometry, metrics and parameters.
  std::string sig;
  for (int32_t i = 0; i < n; i++) {
    sig.append((const char*)&p0[i]->g, sizeof(Geom));
    sig.append((const char*)&metrics[i], sizeof(uint32_t));
    sig.append((const char*)&b->prm[(size_t)i], sizeof(qmiv_params));
  }
  return batch_tune_finish(b, opt, sig, out);
}

qmiv_status qmiv_batch_launch(qmiv_batch* b) {
  if (!b) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (b->pending) return fail(QMIV_ERR_INVALID_ARG, "the batch has a replay in flight: collect it first");
  qmiv_ctx* ctx = b->ctx;
  QMIV_HIP(hipSetDevice(ctx->device));
  if (!b->done) QMIV_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
  QMIV_HIP(hipGraphLaunch(b->exec, ctx->stream));
  QMIV_HIP(hipEventRecord(b->done, ctx->stream));
  b->pending = true;
  // The spherical weights every view needs exist before the views are finalised in parallel (the
  // table is only read there).
  for (int32_t i = 0; i < b->n; i++) {
    const qmiv_params& pi = b->mb ? masked_batch_params(b->mb) : b->prm[(size_t)i];
    if (pi.erp) (void)ws_table(ctx, b->p0[(size_t)i]->g.h, pi.lat_range_deg);
  }
  if (!b->pool && b->n > 1) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    b->pool.reset(new HostPool((int)std::min<unsigned>(7u, std::min<unsigned>(hw, (unsigned)b->n) - 1u)));
  }
  return QMIV_OK;
}

qmiv_status qmiv_batch_collect(qmiv_batch* b, qmiv_frame_result* out) {
  if (!b || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (b->mb) return fail(QMIV_ERR_INVALID_ARG, "a masked batch: use qmiv_batch_collect_masked");
  if (!b->pending) return fail(QMIV_ERR_INVALID_ARG, "the batch has no replay in flight");
  QMIV_HIP(hipSetDevice(b->ctx->device));
  b->pending = false;
  QMIV_HIP(hipEventSynchronize(b->done));
  std::vector<qmiv_status> st((size_t)b->n, QMIV_OK);
  auto fin_view = [&](int i) {
    const size_t k = (size_t)i;
    st[k] = finalize_frame(b->ctx, b->p0[k]->g, b->prm[k], b->metrics[k], b->fw[k], b->plans[k].L,
                           (const char*)b->h_results + b->hoff[k], &out[i]);
  };
  if (b->pool) b->pool->run(b->n, fin_view);
  else for (int32_t i = 0; i < b->n; i++) fin_view(i);
  for (int32_t i = 0; i < b->n; i++)
    if (st[(size_t)i] != QMIV_OK)
      return fail(st[(size_t)i], "view " + std::to_string(i) + ": input contains pels above the bit-depth maximum");
  return QMIV_OK;
}

qmiv_status qmiv_batch_run(qmiv_batch* b, qmiv_frame_result* out) {
  if (!b || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (b->mb) return fail(QMIV_ERR_INVALID_ARG, "a masked batch: use qmiv_batch_run_masked");
  QMIV_TRY(qmiv_batch_launch(b));
  return qmiv_batch_collect(b, out);
}

qmiv_status qmiv_batch_get_info(const qmiv_batch* b, qmiv_batch_info* info) {
  if (!b || !info) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  *info = b->info;
  return QMIV_OK;
}

void qmiv_batch_destroy(qmiv_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_drop_graph(b);
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->d_scratch) (void)hipFree(b->d_scratch);
  if (b->d_scp) (void)hipFree(b->d_scp);
  if (b->h_results) (void)hipHostFree(b->h_results);
  destroy_masked_batch(b->mb);
  delete b;
}


qmiv_status qmiv_ssim_backward(qmiv_ctx* ctx, int32_t w, int32_t h, int32_t planes, const float* tst, const float* ref,
                               double max_value, int32_t calc_l, double* value, float* grad_tst, float* grad_ref) {
  QMIV_TRY(check_fargs(ctx, w, h, planes, tst, ref, max_value));
  if (!value) return fail(QMIV_ERR_INVALID_ARG, "null value");
  QMIV_HIP(hipSetDevice(ctx->device));
  const bool grad = grad_tst || grad_ref;
  const size_t plane = (size_t)w * h;
  Bump b{nullptr};
  auto layout = [&](Bump& bb, double*& partials, double*& rows, double*& maps) {
    partials = bb.take<double>(std::max<size_t>((size_t)ssimf_tiles_x(w) * h, 1));
    rows = bb.take<double>((size_t)planes * h);
    maps = grad ? bb.take<double>(4 * plane) : nullptr;
  };
  double *partials, *rows, *maps;
  layout(b, partials, rows, maps);
  QMIV_TRY(ensure_fwork(ctx, b.off));
  QMIV_TRY(ensure_results(ctx, sizeof(double) * (size_t)planes * h));
  Bump bb{(char*)ctx->d_fwork};
  layout(bb, partials, rows, maps);
  for (int32_t c = 0; c < planes; c++) {
    SsimGArgs g{};
    g.grad_t = grad_tst ? grad_tst + c * plane : nullptr;
    g.grad_r = grad_ref ? grad_ref + c * plane : nullptr;
    g.coef_t = g.coef_r = 1.0;
    QMIV_TRY(enqueue_ssimf(ctx, tst + c * plane, 0, max_value, ref + c * plane, 0, max_value, w, h, calc_l, max_value,
                           partials, rows + (size_t)c * h, maps, grad ? &g : nullptr));
  }
  QMIV_HIP(hipMemcpyAsync(ctx->h_results, rows, sizeof(double) * (size_t)planes * h, hipMemcpyDeviceToHost, ctx->stream));
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  for (int32_t c = 0; c < planes; c++) value[c] = fin::ssim_from_rows((const double*)ctx->h_results + (size_t)c * h, h, w);
  return QMIV_OK;
}

qmiv_status qmiv_msssim_backward(qmiv_ctx* ctx, int32_t w, int32_t h, int32_t planes, const float* tst,
                                 const float* ref, double max_value, double* value, double* scales, float* grad_tst,
                                 float* grad_ref) {
  QMIV_TRY(check_fargs(ctx, w, h, planes, tst, ref, max_value));
  if (!value) return fail(QMIV_ERR_INVALID_ARG, "null value");
  QMIV_HIP(hipSetDevice(ctx->device));
  const bool grad = grad_tst || grad_ref;
  const size_t plane = (size_t)w * h;
  constexpr int kScales = 5;
  // Level l >= 1 of every plane: pictures in fp64 (exact 2 x 2 means) and the level's gradients.
  struct Lv {
    const void* t;
    const void* r;
    float *gt, *gr;
  };
  auto layout = [&](Bump& bb, double*& partials, double*& rows, double*& maps, std::vector<Lv>& lv) {
    partials = bb.take<double>(std::max<size_t>((size_t)ssimf_tiles_x(w) * h, 1));
    rows = bb.take<double>((size_t)planes * kScales * h);
    maps = grad ? bb.take<double>(4 * plane) : nullptr;
    for (int32_t c = 0; c < planes; c++)
      for (int l = 1; l < kScales; l++) {
        const size_t n = std::max<size_t>((size_t)(w >> l) * (h >> l), 1);
        Lv& v = lv[(size_t)c * kScales + l];
        v.t = bb.take<double>(n);
        v.r = bb.take<double>(n);
        v.gt = grad ? bb.take<float>(n) : nullptr;
        v.gr = grad ? bb.take<float>(n) : nullptr;
      }
  };
  double *partials, *rows, *maps;
  std::vector<Lv> lv((size_t)planes * kScales);
  Bump b{nullptr};
  layout(b, partials, rows, maps, lv);
  QMIV_TRY(ensure_fwork(ctx, b.off));
  QMIV_TRY(ensure_results(ctx, sizeof(double) * (size_t)planes * kScales * h));
  Bump bb{(char*)ctx->d_fwork};
  layout(bb, partials, rows, maps, lv);
  hipStream_t st = ctx->stream;
  for (int32_t c = 0; c < planes; c++) {
    Lv* L = &lv[(size_t)c * kScales];
    L[0].t = tst + c * plane;
    L[0].r = ref + c * plane;
    L[0].gt = grad_tst ? grad_tst + c * plane : nullptr;
    L[0].gr = grad_ref ? grad_ref + c * plane : nullptr;
    for (int l = 1; l < kScales; l++) {
      QMIV_HIP(launch_down2(L[l - 1].t, l > 1, w >> (l - 1), h >> (l - 1), (double*)L[l].t, st));
      QMIV_HIP(launch_down2(L[l - 1].r, l > 1, w >> (l - 1), h >> (l - 1), (double*)L[l].r, st));
    }
    for (int l = 0; l < kScales; l++) {
      const int32_t wl = w >> l, hl = h >> l;
      double* rl = rows + ((size_t)c * kScales + l) * h;
      if (wl <= 0 || hl <= 0) continue;
      SsimGArgs g{};
      g.grad_t = L[l].gt;
      g.grad_r = L[l].gr;
      g.coef_t = g.coef_r = 1.0;
      // CS at the first and last scale, L x CS between (libs/pmbb-ivqm/src/xSSIM.cpp:74,80)
      const int32_t calc_l = (l == 0 || l == kScales - 1) ? 0 : 1;
      QMIV_TRY(enqueue_ssimf(ctx, L[l].t, l > 0, max_value, L[l].r, l > 0, max_value, wl, hl, calc_l, max_value,
                             partials, rl, maps, grad ? &g : nullptr));
    }
  }
  QMIV_HIP(hipMemcpyAsync(ctx->h_results, rows, sizeof(double) * (size_t)planes * kScales * h, hipMemcpyDeviceToHost,
                          st));
  QMIV_HIP(hipStreamSynchronize(st));
  static const double wts[kScales] = {0.0448, 0.2856, 0.3001, 0.2363, 0.1333};  // xStructSimConsts.h:27
  for (int32_t c = 0; c < planes; c++) {
    double sc[kScales];
    for (int l = 0; l < kScales; l++)
      sc[l] = fin::ssim_from_rows((const double*)ctx->h_results + ((size_t)c * kScales + l) * h, h >> l, w >> l);
    value[c] = fin::msssim_combine(sc);
    if (scales)
      for (int l = 0; l < kScales; l++) scales[l * planes + c] = sc[l];
    if (grad) {
      // d(prod s_l^w_l)/d s_l = MS w_l / s_l; the pyramid adjoint folds coarse into fine levels.
      Lv* L = &lv[(size_t)c * kScales];
      for (int l = kScales - 1; l >= 0; l--) {
        const int32_t wl = w >> l, hl = h >> l;
        if (wl <= 0 || hl <= 0) continue;
        const double coef = value[c] * wts[l] / sc[l];
        const bool top = l == kScales - 1 || (w >> (l + 1)) <= 0 || (h >> (l + 1)) <= 0;
        if (L[l].gt) QMIV_HIP(launch_up2_addf(top ? nullptr : L[l + 1].gt, coef, wl, hl, L[l].gt, st));
        if (L[l].gr) QMIV_HIP(launch_up2_addf(top ? nullptr : L[l + 1].gr, coef, wl, hl, L[l].gr, st));
      }
    }
  }
  if (grad) QMIV_HIP(hipStreamSynchronize(st));
  return QMIV_OK;
}

qmiv_status qmiv_ivssim_backward(qmiv_ctx* ctx, int32_t w, int32_t h, const float* tst, const float* ref,
                                 double max_value, const qmiv_params* prm_in, double* ivssim, double* r2t, double* t2r,
                                 float* grad_tst, float* grad_ref) {
  QMIV_TRY(check_fargs(ctx, w, h, 3, tst, ref, max_value));
  int32_t bd = 0;
  for (int b = 8; b <= 16; b++)
    if (max_value == (double)((1 << b) - 1)) bd = b;
  if (!bd) return fail(QMIV_ERR_UNSUPPORTED, "IV-SSIM needs max_value = 2^bd - 1 with bd in 8..16");
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  const bool grad = grad_tst || grad_ref;
  const size_t plane = (size_t)w * h;
  hipStream_t st = ctx->stream;
  // Shift search on the quantised pictures (P0 = tst, P1 = ref), keeping each pel's winner.
  qmiv_pic *q0, *q1;
  QMIV_TRY(fq_picture(ctx, 0, w, h, bd, &q0));
  QMIV_TRY(fq_picture(ctx, 1, w, h, bd, &q1));
  auto layout = [&](Bump& bb, int16_t*& best, float*& scp, float*& gscp, double*& partials, double*& rows,
                    double*& maps) {
    best = bb.take<int16_t>(2 * plane);
    scp = bb.take<float>(6 * plane);
    gscp = grad ? bb.take<float>(3 * plane) : nullptr;
    partials = bb.take<double>(std::max<size_t>(6 * (size_t)ssimf_tiles_x(w) * h, 1));  // one per evaluation
    rows = bb.take<double>(6 * (size_t)h);
    maps = grad ? bb.take<double>(24 * plane) : nullptr;  // 4 maps of each of the six evaluations
  };
  int16_t* best;
  float *scp, *gscp;
  double *partials, *rows, *maps;
  Bump b{nullptr};
  layout(b, best, scp, gscp, partials, rows, maps);
  QMIV_TRY(ensure_fwork(ctx, b.off));
  Bump bb{(char*)ctx->d_fwork};
  layout(bb, best, scp, gscp, partials, rows, maps);
  const int32_t bytes = bd <= 8 ? 1 : 2;
  QMIV_HIP(launch_quantf(tst, 3 * (int64_t)plane, max_value, (1 << bd) - 1, q0->dev, bytes, st));
  QMIV_HIP(launch_quantf(ref, 3 * (int64_t)plane, max_value, (1 << bd) - 1, q1->dev, bytes, st));
  FrameWork fw;
  fw.need_gcd = true;
  fw.need_cps = true;
  fw.best0 = best;
  fw.best1 = best + plane;
  Layout L;
  bool fast;
  QMIV_TRY(run_frame_kernels(ctx, q0, q1, prm, fw, L, fast));
  const FrameSummary* sum = (const FrameSummary*)((const char*)ctx->h_results + L.off_summary);
  int32_t gcd[3] = {sum->gcd[0], sum->gcd[1], sum->gcd[2]};
  const int32_t maxv = (1 << bd) - 1;
  // SCP1 = clip(P1(best0) - gcd) (R2T winners), SCP0 = clip(P0(best1) + gcd) (T2R winners), in samples.
  ScpFArgs s1{ref, best, w, h, prm.search_range, max_value, 1, {gcd[0], gcd[1], gcd[2]}, nullptr, maxv, scp + 3 * plane};
  ScpFArgs s0{tst, best + plane, w, h, prm.search_range, max_value, -1, {gcd[0], gcd[1], gcd[2]}, nullptr, maxv, scp};
  QMIV_HIP(launch_scp_gatherf(s1, st));
  QMIV_HIP(launch_scp_gatherf(s0, st));
  // T2R_c = SSIM(P0_c, SCP1_c), R2T_c = SSIM(P1_c, SCP0_c) (libs/pmbb-ivqm/src/xSSIM.cpp:197-207). With
  // gradients, every evaluation also keeps its maps (evaluation e = 3 dir + c, dir 0 = T2R), so the
  // gradient pass after the minimum's direction is known only correlates the chosen three.
  const size_t pstep = (size_t)ssimf_tiles_x(w) * h;
  for (int c = 0; c < 3; c++) {
    QMIV_TRY(enqueue_ssimf(ctx, tst + c * plane, 0, max_value, scp + 3 * plane + c * plane, 0, 1.0, w, h, 1, max_value,
                           partials + c * pstep, nullptr, maps ? maps + (size_t)c * 4 * plane : nullptr, nullptr));
    QMIV_TRY(enqueue_ssimf(ctx, ref + c * plane, 0, max_value, scp + c * plane, 0, 1.0, w, h, 1, max_value,
                           partials + (3 + c) * pstep, nullptr, maps ? maps + (size_t)(3 + c) * 4 * plane : nullptr,
                           nullptr));
  }
  QMIV_HIP(launch_ssim_rows(ivssim_rows_jobs(w, h, partials, rows), st));
  QMIV_TRY(ensure_results(ctx, sizeof(double) * 6 * (size_t)h));
  QMIV_HIP(hipMemcpyAsync(ctx->h_results, rows, sizeof(double) * 6 * (size_t)h, hipMemcpyDeviceToHost, st));
  QMIV_HIP(hipStreamSynchronize(st));
  double vt2r[3], vr2t[3];
  for (int c = 0; c < 3; c++) {
    vt2r[c] = fin::ssim_from_rows((const double*)ctx->h_results + (size_t)c * h, h, w);
    vr2t[c] = fin::ssim_from_rows((const double*)ctx->h_results + (size_t)(3 + c) * h, h, w);
  }
  const double T2R = fin::weighted_average4(vt2r, prm.cmp_weights_average);
  const double R2T = fin::weighted_average4(vr2t, prm.cmp_weights_average);
  const double IV = fin::min2(T2R, R2T);
  if (ivssim) *ivssim = IV;
  if (t2r) *t2r = T2R;
  if (r2t) *r2t = R2T;
  if (!grad) return QMIV_OK;
  // Subgradient: the shifts, the GCD and the direction of the minimum stay fixed. Every output pel
  // is written: the direct argument's by the correlation, the other's by the SCP adjoint.
  const bool use_r2t = R2T < T2R;  // min2 keeps T2R on ties
  const double wsum = (double)(prm.cmp_weights_average[0] + prm.cmp_weights_average[1] + prm.cmp_weights_average[2] +
                               prm.cmp_weights_average[3]);
  // Direct argument: P0 (T2R) or P1 (R2T); the shift-compensated argument is scattered back into
  // the other picture through its winners.
  const float* direct = use_r2t ? ref : tst;
  float* g_direct = use_r2t ? grad_ref : grad_tst;
  float* g_other = use_r2t ? grad_tst : grad_ref;
  const float* scp_sel = use_r2t ? scp : scp + 3 * plane;
  const double* maps_sel = maps + (use_r2t ? (size_t)3 * 4 * plane : 0);
  for (int c = 0; c < 3; c++) {
    SsimGArgs g{};
    g.grad_t = g_direct ? g_direct + c * plane : nullptr;
    g.grad_r = g_other ? gscp + c * plane : nullptr;
    g.coef_t = g.coef_r = (double)prm.cmp_weights_average[c] / wsum;
    g.accumulate = 0;
    QMIV_TRY(enqueue_ssimf_grad(direct + c * plane, 0, max_value, scp_sel + c * plane, 0, 1.0, w, h, max_value,
                                maps_sel + (size_t)c * 4 * plane, g, st));
  }
  if (g_other) QMIV_HIP(launch_scp_scatterf(use_r2t ? s0 : s1, gscp, 1.0, g_other, st));
  QMIV_HIP(hipStreamSynchronize(st));
  return QMIV_OK;
}

qmiv_status qmiv_ivssim_backward_batch(qmiv_ctx* ctx, int32_t n, int32_t w, int32_t h, const float* const* tst,
                                       const float* const* ref, double max_value, const qmiv_params* prm_in,
                                       double* ivssim, double* r2t, double* t2r, float* const* grad_tst,
                                       float* const* grad_ref) {
  if (!ctx || n < 0 || (n > 0 && (!tst || !ref))) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (n == 0) return QMIV_OK;
  for (int32_t i = 0; i < n; i++) QMIV_TRY(check_fargs(ctx, w, h, 3, tst[i], ref[i], max_value));
  int32_t bd = 0;
  for (int b = 8; b <= 16; b++)
    if (max_value == (double)((1 << b) - 1)) bd = b;
  if (!bd) return fail(QMIV_ERR_UNSUPPORTED, "IV-SSIM needs max_value = 2^bd - 1 with bd in 8..16");
  qmiv_params prm;
  if (prm_in) prm = *prm_in;
  else qmiv_params_default(&prm);
  QMIV_HIP(hipSetDevice(ctx->device));
  const size_t plane = (size_t)w * h;
  auto want_grad = [&](int32_t i) { return (grad_tst && grad_tst[i]) || (grad_ref && grad_ref[i]); };
  bool any_grad = false;
  for (int32_t i = 0; i < n; i++) any_grad = any_grad || want_grad(i);
  // The search of every pair on its quantised pictures, as qmiv_ivssim_backward's; one plan serves all.
  Geom g;
  QMIV_TRY(make_geom(w, h, bd, 444, g));
  FrameWork fw0;
  fw0.ipa = QMIV_PEL_SKIP;
  fw0.need_gcd = true;
  fw0.need_cps = true;
  const FramePlan P = plan_frame(ctx, g, prm, fw0);
  const Layout& L = P.L;
  const size_t dstep = align_up(L.total, 256);
  const size_t hstep = align_up(L.results_bytes + sizeof(double) * 6 * (size_t)h, 256);
  // Each pair's float workspace: winners, both SCP pictures, the SCP gradient, partials, rows, maps.
  struct Work {
    int16_t* best;
    float *scp, *gscp;
    double *partials, *rows, *maps;
  };
  auto layout = [&](Bump& bb, Work& wk) {
    wk.best = bb.take<int16_t>(2 * plane);
    wk.scp = bb.take<float>(6 * plane);
    wk.gscp = any_grad ? bb.take<float>(3 * plane) : nullptr;
    wk.partials = bb.take<double>(std::max<size_t>(6 * (size_t)ssimf_tiles_x(w) * h, 1));  // one per evaluation
    wk.rows = bb.take<double>(6 * (size_t)h);
    wk.maps = any_grad ? bb.take<double>(24 * plane) : nullptr;  // 4 maps of each of the six evaluations
  };
  Work probe;
  Bump b0{nullptr};
  layout(b0, probe);
  const size_t fstep = align_up(b0.off, 256);
  std::vector<qmiv_pic*> q((size_t)2 * n);
  for (int32_t i = 0; i < 2 * n; i++) QMIV_TRY(fq_picture(ctx, i, w, h, bd, &q[(size_t)i]));
  QMIV_TRY(wait_stream(ctx));  // the shared blocks below are reused: nothing may still read them
  QMIV_TRY(ensure_scratch(ctx, dstep * n));
  QMIV_TRY(ensure_results(ctx, hstep * n));
  QMIV_TRY(ensure_fwork(ctx, fstep * n));
  std::vector<Work> wk((size_t)n);
  std::vector<ScpFArgs> s0((size_t)n), s1((size_t)n);
  const hipStream_t streams[1 + qmiv_ctx::kAuxStreams] = {ctx->stream, ctx->aux_stream[0], ctx->aux_stream[1]};
  const int32_t nstreams = std::min<int32_t>(n, 1 + qmiv_ctx::kAuxStreams);
  if (nstreams > 1) {  // the aux streams start after everything already on the context stream
    QMIV_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
    for (int32_t k = 1; k < nstreams; k++) QMIV_HIP(hipStreamWaitEvent(streams[k], ctx->fork_ev, 0));
  }
  const int32_t maxv = (1 << bd) - 1;
  const int32_t bytes = bd <= 8 ? 1 : 2;
  // Values: quantisation, search, shift-compensated pictures and the six SSIM evaluations of every
  // pair, its pairs dealt round-robin over the streams; the GCD goes from the search's summary to the
  // gathers on the device.
  for (int32_t i = 0; i < n; i++) {
    const hipStream_t st = streams[i % nstreams];
    Bump bb{(char*)ctx->d_fwork + fstep * i};
    Work& k = wk[(size_t)i];
    layout(bb, k);
    char* dbase = (char*)ctx->d_scratch + dstep * i;
    char* hbase = (char*)ctx->h_results + hstep * i;
    QMIV_HIP(launch_quantf(tst[i], 3 * (int64_t)plane, max_value, maxv, q[2 * (size_t)i]->dev, bytes, st));
    QMIV_HIP(launch_quantf(ref[i], 3 * (int64_t)plane, max_value, maxv, q[2 * (size_t)i + 1]->dev, bytes, st));
    FrameWork fw = fw0;
    fw.best0 = k.best;
    fw.best1 = k.best + plane;
    QMIV_TRY(enqueue_frame(ctx, q[2 * (size_t)i], q[2 * (size_t)i + 1], prm, fw, P, dbase, hbase, nullptr, st));
    const int32_t* gcd_dev = ((const FrameSummary*)(dbase + L.off_summary2))->gcd;
    s1[(size_t)i] = ScpFArgs{ref[i], k.best, w, h, prm.search_range, max_value, 1, {0, 0, 0}, gcd_dev, maxv,
                             k.scp + 3 * plane};
    s0[(size_t)i] = ScpFArgs{tst[i], k.best + plane, w, h, prm.search_range, max_value, -1, {0, 0, 0}, gcd_dev, maxv,
                             k.scp};
    QMIV_HIP(launch_scp_gatherf(s1[(size_t)i], st));
    QMIV_HIP(launch_scp_gatherf(s0[(size_t)i], st));
    // With gradients the evaluations keep their maps, as in qmiv_ivssim_backward.
    double* maps = want_grad(i) ? k.maps : nullptr;
    const size_t pstep = (size_t)ssimf_tiles_x(w) * h;
    for (int c = 0; c < 3; c++) {
      QMIV_TRY(enqueue_ssimf(ctx, tst[i] + c * plane, 0, max_value, k.scp + 3 * plane + c * plane, 0, 1.0, w, h, 1,
                             max_value, k.partials + c * pstep, nullptr,
                             maps ? maps + (size_t)c * 4 * plane : nullptr, nullptr, st));
      QMIV_TRY(enqueue_ssimf(ctx, ref[i] + c * plane, 0, max_value, k.scp + c * plane, 0, 1.0, w, h, 1, max_value,
                             k.partials + (3 + c) * pstep, nullptr,
                             maps ? maps + (size_t)(3 + c) * 4 * plane : nullptr, nullptr, st));
    }
    QMIV_HIP(launch_ssim_rows(ivssim_rows_jobs(w, h, k.partials, k.rows), st));
    QMIV_HIP(hipMemcpyAsync(hbase + L.results_bytes, k.rows, sizeof(double) * 6 * (size_t)h, hipMemcpyDeviceToHost,
                            st));
  }
  for (int32_t k = 0; k < nstreams; k++) QMIV_HIP(hipStreamSynchronize(streams[k]));
  const double wsum = (double)(prm.cmp_weights_average[0] + prm.cmp_weights_average[1] +
                               prm.cmp_weights_average[2] + prm.cmp_weights_average[3]);
  std::vector<char> use_r2t((size_t)n, 0);
  for (int32_t i = 0; i < n; i++) {
    const double* hr = (const double*)((const char*)ctx->h_results + hstep * i + L.results_bytes);
    double vt2r[3], vr2t[3];
    for (int c = 0; c < 3; c++) {
      vt2r[c] = fin::ssim_from_rows(hr + (size_t)c * h, h, w);
      vr2t[c] = fin::ssim_from_rows(hr + (size_t)(3 + c) * h, h, w);
    }
    const double T2R = fin::weighted_average4(vt2r, prm.cmp_weights_average);
    const double R2T = fin::weighted_average4(vr2t, prm.cmp_weights_average);
    if (ivssim) ivssim[i] = fin::min2(T2R, R2T);
    if (t2r) t2r[i] = T2R;
    if (r2t) r2t[i] = R2T;
    use_r2t[(size_t)i] = R2T < T2R;  // min2 keeps T2R on ties
  }
  if (!any_grad) return QMIV_OK;
  // Gradients (the subgradient of qmiv_ivssim_backward: shifts, GCD and the minimum's direction fixed).
  for (int32_t i = 0; i < n; i++) {
    if (!want_grad(i)) continue;
    const hipStream_t st = streams[i % nstreams];
    const Work& k = wk[(size_t)i];
    float* gt = grad_tst ? grad_tst[i] : nullptr;
    float* gr = grad_ref ? grad_ref[i] : nullptr;
    const bool r2 = use_r2t[(size_t)i] != 0;
    const float* direct = r2 ? ref[i] : tst[i];
    float* g_direct = r2 ? gr : gt;
    float* g_other = r2 ? gt : gr;
    const float* scp_sel = r2 ? k.scp : k.scp + 3 * plane;
    const double* maps_sel = k.maps + (r2 ? (size_t)3 * 4 * plane : 0);
    for (int c = 0; c < 3; c++) {
      SsimGArgs ga{};
      ga.grad_t = g_direct ? g_direct + c * plane : nullptr;
      ga.grad_r = g_other ? k.gscp + c * plane : nullptr;
      ga.coef_t = ga.coef_r = (double)prm.cmp_weights_average[c] / wsum;
      ga.accumulate = 0;
      QMIV_TRY(enqueue_ssimf_grad(direct + c * plane, 0, max_value, scp_sel + c * plane, 0, 1.0, w, h, max_value,
                                  maps_sel + (size_t)c * 4 * plane, ga, st));
    }
    if (g_other) QMIV_HIP(launch_scp_scatterf(r2 ? s0[(size_t)i] : s1[(size_t)i], k.gscp, 1.0, g_other, st));
  }
  for (int32_t k = 0; k < nstreams; k++) QMIV_HIP(hipStreamSynchronize(streams[k]));
  return QMIV_OK;
}

// ---------------------------------------------------------------------------------------------
// Masked mode (apps/QMIV/src/xAppQMIV.cpp:319-337,695,825-889)
// ---------------------------------------------------------------------------------------------
namespace {

// Scratch of one masked evaluation.
//   device: [results block | nnz | sd | max | diff | CPS partials]
//   results block (copied back): MaskSummary | masked SSD rows [3][h] | CPS rows [2][3][h]
struct MaskedLayout {
  size_t off_sum = 0, off_ssd = 0, off_cps = 0, results_bytes = 0;
  size_t off_nnz = 0, off_sd = 0, off_max = 0, off_diff = 0, off_part = 0, total = 0;
  int32_t nseg = 0;
  // As in Layout: the summary the search and the host read (the clipped inputs' under concealment),
  // and the general search's partials when it is the fallback of an out-of-range frame.
  size_t off_sum2 = 0, off_part_gen = 0;
  int32_t nseg_gen = 0;
};

// nseg: CPS partial segments per row (of the kernel that runs the search); conceal: a second
// summary; nseg_gen > 0: partials of the general search as the out-of-range fallback.
MaskedLayout masked_layout(const Geom& g, bool cps, int32_t nseg, bool conceal = false, int32_t nseg_gen = 0) {
  MaskedLayout L;
  const size_t h = (size_t)g.h;
  size_t o = 0;
  L.off_sum = o;
  o = align_up(o + sizeof(MaskSummary), 256);
  L.off_sum2 = L.off_sum;
  if (conceal) {
    L.off_sum2 = o;
    o = align_up(o + sizeof(MaskSummary), 256);
  }
  L.off_ssd = o;
  o = align_up(o + 3 * h * sizeof(uint64_t), 256);
  L.off_cps = o;
  o = align_up(o + (cps ? 6 * h * sizeof(uint64_t) : 0), 256);
  L.results_bytes = o;
  L.off_nnz = o;
  o = align_up(o + h * sizeof(int32_t), 256);
  L.off_sd = o;
  o = align_up(o + 3 * h * sizeof(int64_t), 256);
  L.off_max = o;
  o = align_up(o + 3 * h * sizeof(uint32_t), 256);
  L.off_diff = o;
  o = align_up(o + h * sizeof(uint32_t), 256);
  L.nseg = nseg;
  L.off_part = o;
  o = align_up(o + (cps ? (size_t)L.nseg * 6 * h * sizeof(uint64_t) : 0), 256);
  if (cps && nseg_gen > 0) {
    L.nseg_gen = nseg_gen;
    L.off_part_gen = o;
    o = align_up(o + (size_t)nseg_gen * 6 * h * sizeof(uint64_t), 256);
  }
  L.total = o;
  return L;
}

qmiv_status check_mask(const qmiv_pic* p0, const qmiv_pic* mask) {
  if (!mask) return fail(QMIV_ERR_INVALID_ARG, "null mask picture");
  if (mask->ctx != p0->ctx) return fail(QMIV_ERR_INVALID_ARG, "mask belongs to another context");
  if (mask->g.w != p0->g.w || mask->g.h != p0->g.h)
    return fail(QMIV_ERR_INVALID_ARG, "mask differs from the pictures in size");
  return QMIV_OK;
}

// The masked search on the fast kernels (4:2:0 and 4:4:4): besides the unmasked kernels'
// conditions, a masked candidate's record must lose to every unmasked one, which the kernels ensure
// when every real key (and, in the 4:2:0 kernel, every chroma term T added to it) stays below 2^29
// in magnitude. gcd_bound: |GCD| per component, at most.
bool fast_masked_cps_ok(const Geom& g, const qmiv_params& p, const int32_t* gcd_bound) {
  if (!fast_cps_ok(g, p, gcd_bound)) return false;
  const int64_t maxv = (1 << g.bd) - 1;
  int64_t m = 0;
  for (int c = 0; c < 3; c++) m = std::max<int64_t>(m, maxv + std::abs((int64_t)gcd_bound[c]));
  const int64_t w0 = p.cmp_weights_search[0], w1 = p.cmp_weights_search[1], w2 = p.cmp_weights_search[2];
  // key = 32 sum_c w_c Y_c (Y_c - 2 X_c) + h, with Y in [0, maxv] and |X| <= m: each term lies in
  // [-m^2, m^2]; T = -64 (w1 X1 U + w2 X2 V).
  const int64_t lim = (int64_t)1 << 29;
  return 32 * (w0 + w1 + w2) * m * m + 32 < lim && 64 * (w1 + w2) * m * maxv < lim;
}

// Everything decided on the host before a masked frame's kernels are enqueued.
struct MaskedPlan {
  MaskedLayout L;
  bool cps = false;
  bool fast = false;  // the search runs on a fast kernel (the 4:2:0 one when g.cs == 1)
  bool s16 = false;   // ... the 16-bit 4:2:0 kernel
  CpsArgs fa{};       // its tiling (tiles_x, sweep, nseg) when fast
  int32_t md[4];
  int32_t ipa = QMIV_PEL_STOP;
  bool conceal = false;   // CNCL: clip out-of-range inputs, then the statistics again
  bool fallback = false;  // WARN / SKIP: the general search, for a frame with out-of-range pels
};

// gcd_given: null to compute the masked GCD; gm: the mask's geometry; ipa: the invalid-pel action
// (processor-level calls pass QMIV_PEL_SKIP: no validation).
qmiv_status plan_masked(qmiv_ctx* ctx, const Geom& g, const Geom& gm, const qmiv_params& prm, int32_t ipa, bool cps,
                        const int32_t* gcd_given, MaskedPlan& P) {
  if (cps && (prm.search_range < 0 || prm.search_range > kMaxMaskedSearchRange))
    return fail(QMIV_ERR_UNSUPPORTED, "masked search supports search ranges 0..12");
  P = MaskedPlan{};
  P.cps = cps;
  // The computed GCD is clamped to the unnoticeable-difference bound, and its fourth entry is 0.
  int32_t gbound[4];
  for (int c = 0; c < 4; c++) P.md[c] = fin::gcd_max_diff(prm.unnoticeable_coef[c], g.bd);
  for (int c = 0; c < 4; c++) gbound[c] = gcd_given ? gcd_given[c] : (c < 3 ? P.md[c] : 0);
  P.fast = cps && prm.search_range == 2 && fast_masked_cps_ok(g, prm, gbound);
  if (P.fast) {
    const bool f420 = g.cs == 1;
    // The 16-bit kernel keeps each lane's masked sums of a sweep in 32 bits: 2 pels per row and tile,
    // each square (|X - Y| <= maxv + |gcd|) times a mask value of at most maxm.
    int64_t s16_sweep = 0;
    if (f420 && cps420s_ok(g, prm, gbound)) {
      const int64_t maxv = (1 << g.bd) - 1, maxm = (1 << gm.bd) - 1;
      int64_t m = 0;
      for (int c = 0; c < 3; c++) m = std::max<int64_t>(m, maxv + std::abs((int64_t)gbound[c]));
      s16_sweep = (int64_t)0xffffffffLL / (2 * m * m * maxm);
    }
    P.s16 = s16_sweep >= 1;
    const int kind = P.s16 ? kCps420s : f420 ? kCps420 : kCps444;
    int32_t tw, th;
    cps_tile(kind, &tw, &th);
    P.fa.tiles_x = (g.w + tw - 1) / tw;
    P.fa.sweep = pick_sweep(ctx, g, kind, gbound);
    if (P.s16) P.fa.sweep = (int32_t)std::min<int64_t>(P.fa.sweep, s16_sweep);
    P.fa.nseg = (P.fa.tiles_x + P.fa.sweep - 1) / P.fa.sweep;
  }
  P.ipa = ipa;
  const bool exceed = may_exceed(g) || may_exceed(gm);
  P.conceal = ipa == QMIV_PEL_CNCL && exceed;
  P.fallback = (ipa == QMIV_PEL_WARN || ipa == QMIV_PEL_SKIP) && exceed && P.fast;
  P.L = masked_layout(g, cps, P.fast ? P.fa.nseg : masked_cps_segments(g), P.conceal,
                      P.fallback ? masked_cps_segments(g) : 0);
  return QMIV_OK;
}

// Enqueues the masked statistics (and, with P.cps, the masked search of both directions) of
// P0 / P1 under the mask on stream st, over scratch d (P.L.total bytes), then the copy of the
// results block into the pinned host block h.
qmiv_status enqueue_masked(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_pic* mask,
                           const qmiv_params& prm, MaskedPlan& P, int32_t nnm_given, const int32_t* gcd_given,
                           char* d, char* h, hipStream_t st) {
  const Geom& g = p0->g;
  const MaskedLayout& L = P.L;
  MaskRows r;
  r.nnz = (int32_t*)(d + L.off_nnz);
  r.sd = (int64_t*)(d + L.off_sd);
  r.ssd = (uint64_t*)(d + L.off_ssd);
  r.max = (uint32_t*)(d + L.off_max);
  r.diff = (uint32_t*)(d + L.off_diff);
  MaskSummary* sum = (MaskSummary*)(d + L.off_sum);
  MaskSummary* sum2 = (MaskSummary*)(d + L.off_sum2);
  const int32_t maxm = (1 << mask->g.bd) - 1;
  {
    KernelTimer t(ctx, QMIV_K_STATS, st);
    QMIV_HIP(launch_mstats(p0->dev, p1->dev, g, mask->dev, mask->g, r, st));
  }
  {
    KernelTimer t(ctx, QMIV_K_SUMMARY, st);
    QMIV_HIP(launch_msummary(g, maxm, r, P.md, nnm_given, gcd_given, sum, st));
  }
  if (P.conceal) {
    // CNCL (xAppQMIV.cpp:753-762) over the three inputs: clip those holding a pel above their
    // maximum, then the masked statistics again; each step is a no-op for an in-range frame.
    const int32_t maxv = (1 << g.bd) - 1;
    QMIV_HIP(launch_clip(p0->dev, g, maxv, &sum->max_pel[0], st));
    QMIV_HIP(launch_clip(p1->dev, g, maxv, &sum->max_pel[1], st));
    QMIV_HIP(launch_clip(mask->dev, mask->g, maxm, &sum->max_pel[2], st));
    {
      KernelTimer t(ctx, QMIV_K_STATS, st);
      QMIV_HIP(launch_mstats(p0->dev, p1->dev, g, mask->dev, mask->g, r, st, sum));
    }
    KernelTimer t(ctx, QMIV_K_SUMMARY, st);
    QMIV_HIP(launch_msummary(g, maxm, r, P.md, nnm_given, gcd_given, sum2, st, sum));
  }
  if (P.cps) {
    uint64_t* part = (uint64_t*)(d + L.off_part);
    const bool any_range = P.ipa == QMIV_PEL_WARN || P.ipa == QMIV_PEL_SKIP;
    {
      KernelTimer t(ctx, P.fast ? QMIV_K_CPS : QMIV_K_CPS_GENERAL, st);
      if (P.fast) {
        CpsArgs& fa = P.fa;
        fa.p0 = p0->dev;
        fa.p1 = p1->dev;
        fa.g = g;
        for (int c = 0; c < 4; c++) fa.ws[c] = prm.cmp_weights_search[c];
        fa.B = prm.search_range;
        fa.partials = part;
        fa.maxv = (1 << g.bd) - 1;
        fa.mask = mask->dev;
        fa.mask_bytes = (uint32_t)((size_t)g.w * g.h * mask->g.bytes_per_sample);  // its luma plane
        fa.mask_bytes_per_sample = mask->g.bytes_per_sample;
        fa.msum = sum2;
        if (P.s16) QMIV_HIP(launch_cps420s(fa, st));
        else if (g.cs == 1) QMIV_HIP(launch_cps_fast420(fa, st));
        else QMIV_HIP(launch_cps_fast(fa, st));
      } else {
        QMIV_HIP(launch_cps_masked(p0->dev, p1->dev, g, mask->dev, mask->g, sum2, prm.search_range,
                                   prm.cmp_weights_search, part, st, any_range ? kGateAlways : kGateInRange));
      }
    }
    {
      KernelTimer t(ctx, QMIV_K_CPS_ROWS, st);
      QMIV_HIP(launch_cps_rows(part, L.nseg, g.h, (uint64_t*)(d + L.off_cps), st, &sum2->all_in_range,
                               P.fallback ? 1 : 0));
    }
    if (P.fallback) {
      uint64_t* pg = (uint64_t*)(d + L.off_part_gen);
      {
        KernelTimer t(ctx, QMIV_K_CPS_GENERAL, st);
        QMIV_HIP(launch_cps_masked(p0->dev, p1->dev, g, mask->dev, mask->g, sum2, prm.search_range,
                                   prm.cmp_weights_search, pg, st, kGateOutOfRange));
      }
      KernelTimer t(ctx, QMIV_K_CPS_ROWS, st);
      QMIV_HIP(launch_cps_rows(pg, L.nseg_gen, g.h, (uint64_t*)(d + L.off_cps), st, &sum2->all_in_range, 2));
    }
  }
  QMIV_HIP(hipMemcpyAsync(h, d, L.results_bytes, hipMemcpyDeviceToHost, st));
  return QMIV_OK;
}

// One masked evaluation on the context stream, waited for; results at the start of ctx->h_results.
qmiv_status run_masked(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_pic* mask,
                       const qmiv_params& prm, bool cps, int32_t nnm_given, const int32_t* gcd_given,
                       MaskedLayout& L) {
  MaskedPlan P;
  QMIV_TRY(plan_masked(ctx, p0->g, mask->g, prm, QMIV_PEL_SKIP, cps, gcd_given, P));
  L = P.L;
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_TRY(ensure_scratch(ctx, L.total));
  QMIV_TRY(ensure_results(ctx, L.results_bytes));
  QMIV_TRY(enqueue_masked(ctx, p0, p1, mask, prm, P, nnm_given, gcd_given, (char*)ctx->d_scratch,
                          (char*)ctx->h_results, ctx->stream));
  return wait_stream(ctx);
}

const MaskSummary* masked_summary(const qmiv_ctx* ctx, const MaskedLayout& L) {
  return (const MaskSummary*)((const char*)ctx->h_results + L.off_sum2);
}

qmiv_params masked_params(const qmiv_params* prm_in, bool processor) {
  qmiv_params p;
  if (prm_in) {
    p = *prm_in;
  } else {
    qmiv_params_default(&p);
    if (processor) p.fake_psnr_for_exact = 0;  // xPSNR's default: no fake values
  }
  return p;
}

}  // namespace

namespace {

// Host finalisation of one masked frame from its results block hb (the application's masked
// flow, xAppQMIV.cpp:825-889).
qmiv_status finalize_masked(qmiv_ctx* ctx, const Geom& g, int32_t bdm, const qmiv_params& prm, uint32_t metrics,
                            const MaskedLayout& L, const char* hb, qmiv_masked_result* out) {
  const bool iv = (metrics & QMIV_IVPSNR) != 0;
  // Range flags of the inputs as read; the rest from the summary the search used.
  const MaskSummary* read = (const MaskSummary*)(hb + L.off_sum);
  const MaskSummary* sum = (const MaskSummary*)(hb + L.off_sum2);
  const int32_t nnm = sum->nnm;
  for (int i = 0; i < 3; i++) out->in_range[i] = read->in_range[i];
  for (int c = 0; c < 3; c++) out->exact[c] = sum->exact[c];
  for (int c = 0; c < 4; c++) out->gcd[c] = sum->gcd[c];
  out->num_non_masked = nnm;
  if (!read->all_in_range && prm.invalid_pel_action == QMIV_PEL_STOP)
    return fail(QMIV_ERR_PEL_RANGE, "input or mask contains pels above the bit-depth maximum");
  const double fake = fin::psnr_from_ssd(1.0, g.w * g.h, g.bd);
  double plain[3];
  for (int c = 0; c < 3; c++) plain[c] = fin::psnr_from_masked_ssd((double)sum->ssd[c], nnm, g.bd, bdm);
  static const std::vector<double> kNoWs;
  const qmiv_ctx::WsTable* wst = prm.erp ? &ws_table(ctx, g.h, prm.lat_range_deg) : nullptr;
  const std::vector<double>& ws = wst ? wst->w : kNoWs;
  const double corr = wst ? wst->corr : 1.0;
  if (metrics & QMIV_PSNR) {
    for (int c = 0; c < 3; c++) out->psnr[c] = (prm.fake_psnr_for_exact && out->exact[c]) ? fake : plain[c];
    out->psnr_pic = fin::pic_value(out->psnr, prm.cmp_weights_average);
  }
  if (metrics & QMIV_WSPSNR) {
    const uint64_t* rows = (const uint64_t*)(hb + L.off_ssd);
    for (int c = 0; c < 3; c++) {
      const double v = prm.erp ? fin::wspsnr_from_rows_masked(rows + (size_t)c * g.h, g.h, ws.data(), corr, nnm, g.bd, bdm)
                               : plain[c];
      out->wspsnr[c] = (prm.fake_psnr_for_exact && out->exact[c]) ? fake : v;
    }
    out->wspsnr_pic = fin::pic_value(out->wspsnr, prm.cmp_weights_average);
  }
  if (iv) {
    const uint64_t* rows = (const uint64_t*)(hb + L.off_cps);
    const double* w = prm.erp ? ws.data() : nullptr;
    out->ivpsnr_r2t = fin::qual_from_rows_masked(rows, g.h, w, nnm, g.bd, bdm, prm.cmp_weights_average);
    out->ivpsnr_t2r = fin::qual_from_rows_masked(rows + 3 * (size_t)g.h, g.h, w, nnm, g.bd, bdm, prm.cmp_weights_average);
    out->ivpsnr = fin::min2(out->ivpsnr_r2t, out->ivpsnr_t2r);
    for (int d = 0; d < 2; d++)
      for (int c = 0; c < 3; c++) {
        uint64_t s = 0;
        for (int32_t y = 0; y < g.h; y++) s += rows[((size_t)d * 3 + c) * g.h + y];
        out->cps_sum[d][c] = s;
      }
  }
  return QMIV_OK;
}

}  // namespace

qmiv_status qmiv_frame_masked(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_pic* mask,
                              const qmiv_params* prm_in, uint32_t metrics, qmiv_masked_result* out) {
  if (!ctx || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  const qmiv_pic* a[1] = {p0};
  const qmiv_pic* b[1] = {p1};
  const qmiv_pic* m[1] = {mask};
  return qmiv_frames_masked(ctx, 1, a, b, m, prm_in, metrics, out);
}

qmiv_status qmiv_frames_masked(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                               const qmiv_pic* const* mask, const qmiv_params* prm_in, uint32_t metrics,
                               qmiv_masked_result* out) {
  if (!ctx || !out || !p0 || !p1 || !mask || n < 0) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (n == 0) return QMIV_OK;
  for (int32_t i = 0; i < n; i++) {
    QMIV_TRY(check_pair(p0[i], p1[i]));
    QMIV_TRY(check_mask(p0[i], mask[i]));
    if (p0[i]->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "pictures belong to another context");
  }
  if (metrics & (QMIV_SSIM | QMIV_MSSSIM | QMIV_IVSSIM))
    return fail(QMIV_ERR_INVALID_ARG, "structural similarity metrics cannot be combined with mask mode");
  if (metrics & ~(uint32_t)(QMIV_PSNR | QMIV_WSPSNR | QMIV_IVPSNR)) return fail(QMIV_ERR_INVALID_ARG, "unknown metric bits");
  const qmiv_params prm = masked_params(prm_in, false);
  for (int32_t i = 0; i < n; i++) std::memset(&out[i], 0, sizeof(out[i]));
  const bool iv = (metrics & QMIV_IVPSNR) != 0;
  // Without IV-PSNR the application computes no colour difference and reports zeros.
  const int32_t zero[4] = {0, 0, 0, 0};
  const int32_t* gcd_given = iv ? nullptr : zero;
  // Plan every frame, then lay their scratch and result blocks end to end.
  std::vector<MaskedPlan> plans((size_t)n);
  std::vector<size_t> doff((size_t)n), hoff((size_t)n);
  size_t dtot = 0, htot = 0;
  if (prm.invalid_pel_action < QMIV_PEL_STOP || prm.invalid_pel_action > QMIV_PEL_CNCL)
    return fail(QMIV_ERR_INVALID_ARG, "invalid_pel_action must be 0..3");
  for (int32_t i = 0; i < n; i++) {
    QMIV_TRY(plan_masked(ctx, p0[i]->g, mask[i]->g, prm, prm.invalid_pel_action, iv, gcd_given, plans[(size_t)i]));
    doff[(size_t)i] = dtot;
    hoff[(size_t)i] = htot;
    dtot += align_up(plans[(size_t)i].L.total, 256);
    htot += align_up(plans[(size_t)i].L.results_bytes, 256);
  }
  QMIV_HIP(hipSetDevice(ctx->device));
  QMIV_TRY(ensure_scratch(ctx, dtot));
  QMIV_TRY(ensure_results(ctx, htot));
  while ((int32_t)ctx->done_ev.size() < n) {
    hipEvent_t e;
    QMIV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->done_ev.push_back(e);
  }
  // Frames are dealt round-robin over the context stream and the aux streams, so one frame's
  // statistics overlap the others' searches.
  const int32_t nstreams = std::min<int32_t>(n, 1 + qmiv_ctx::kAuxStreams);
  const hipStream_t streams[1 + qmiv_ctx::kAuxStreams] = {ctx->stream, ctx->aux_stream[0], ctx->aux_stream[1]};
  if (nstreams > 1) {  // the aux streams start after everything already on the context stream
    QMIV_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
    for (int32_t k = 1; k < nstreams; k++) QMIV_HIP(hipStreamWaitEvent(streams[k], ctx->fork_ev, 0));
  }
  // Each frame is finalised on the host as soon as its results have arrived, between enqueues too.
  qmiv_status first_err = QMIV_OK;
  int32_t done = 0;
  auto finalize_next = [&]() {
    const int32_t i = done++;
    const qmiv_status st = finalize_masked(ctx, p0[i]->g, mask[i]->g.bd, prm, metrics, plans[(size_t)i].L,
                                           (const char*)ctx->h_results + hoff[(size_t)i], &out[i]);
    if (st != QMIV_OK && first_err == QMIV_OK) first_err = st;
  };
  for (int32_t i = 0; i < n; i++) {
    const hipStream_t st = streams[i % nstreams];
    qmiv_status e = enqueue_masked(ctx, p0[i], p1[i], mask[i], prm, plans[(size_t)i], -1, gcd_given,
                                   (char*)ctx->d_scratch + doff[(size_t)i], (char*)ctx->h_results + hoff[(size_t)i], st);
    if (e == QMIV_OK && hipEventRecord(ctx->done_ev[(size_t)i], st) != hipSuccess)
      e = fail(QMIV_ERR_HIP, "event record failed");
    if (e != QMIV_OK) {
      // Drain what was enqueued: scratch and result buffers must be idle before any later reuse.
      (void)hipStreamSynchronize(ctx->stream);
      for (hipStream_t a : ctx->aux_stream) (void)hipStreamSynchronize(a);
      return e;
    }
    while (done < i && hipEventQuery(ctx->done_ev[(size_t)done]) == hipSuccess) finalize_next();
  }
  for (int32_t k = 1; k < nstreams; k++) {  // later work on the context stream follows the whole set
    QMIV_HIP(hipEventRecord(ctx->join_ev, streams[k]));
    QMIV_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0));
  }
  while (done < n) {
    QMIV_HIP(hipEventSynchronize(ctx->done_ev[(size_t)done]));
    finalize_next();
  }
  QMIV_TRY(wait_stream(ctx));  // every stream idle; the set's kernel timings are harvested
  return first_err;
}

qmiv_status qmiv_count_nonzero(qmiv_ctx* ctx, const qmiv_pic* mask, int32_t* out) {
  if (!ctx || !mask || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (mask->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "mask belongs to another context");
  const qmiv_params prm = masked_params(nullptr, true);
  const int32_t zero[4] = {0, 0, 0, 0};
  MaskedLayout L;
  QMIV_TRY(run_masked(ctx, mask, mask, mask, prm, false, -1, zero, L));
  *out = masked_summary(ctx, L)->nnm;
  return QMIV_OK;
}

qmiv_status qmiv_calc_gcd_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                 const qmiv_params* prm_in, int32_t num_non_masked, int32_t gcd[4]) {
  if (!ctx || !gcd) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(tst, ref));
  QMIV_TRY(check_mask(tst, mask));
  const qmiv_params prm = masked_params(prm_in, true);
  MaskedLayout L;
  // Processor form (xGlobClrDiff.h:42): the static function runs on (Ref, Tst): sum((Ref - Tst) * m).
  QMIV_TRY(run_masked(ctx, tst, ref, mask, prm, false, num_non_masked, nullptr, L));
  const MaskSummary* sum = masked_summary(ctx, L);
  for (int c = 0; c < 4; c++) gcd[c] = sum->gcd[c];
  return QMIV_OK;
}

qmiv_status qmiv_calc_psnr_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                  const qmiv_params* prm_in, int32_t num_non_masked, double psnr[3]) {
  if (!ctx || !psnr) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(tst, ref));
  QMIV_TRY(check_mask(tst, mask));
  const qmiv_params prm = masked_params(prm_in, true);
  const int32_t zero[4] = {0, 0, 0, 0};
  MaskedLayout L;
  QMIV_TRY(run_masked(ctx, tst, ref, mask, prm, false, -1, zero, L));
  const MaskSummary* sum = masked_summary(ctx, L);
  const Geom& g = tst->g;
  const int32_t nnm = num_non_masked >= 0 ? num_non_masked : sum->nnm;
  for (int c = 0; c < 3; c++) {
    psnr[c] = fin::psnr_from_masked_ssd((double)sum->ssd[c], nnm, g.bd, mask->g.bd);
    // xPSNR::xCalcCmpPSNRM (libs/pmbb-ivqm/src/xPSNR.cpp:84-92)
    if (prm.fake_psnr_for_exact && sum->ssd[c] == 0) psnr[c] = fin::psnr_from_ssd(1.0, g.w * g.h, g.bd);
  }
  return QMIV_OK;
}

qmiv_status qmiv_calc_wspsnr_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                    const qmiv_params* prm_in, int32_t num_non_masked, double wspsnr[3]) {
  if (!ctx || !wspsnr) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  const qmiv_params prm = masked_params(prm_in, true);
  // Without WS the reference recounts the mask (calcPicPSNRM without a count, xWSPSNR.cpp:71-74).
  if (!prm.erp) return qmiv_calc_psnr_masked(ctx, tst, ref, mask, &prm, -1, wspsnr);
  QMIV_TRY(check_pair(tst, ref));
  QMIV_TRY(check_mask(tst, mask));
  const int32_t zero[4] = {0, 0, 0, 0};
  MaskedLayout L;
  QMIV_TRY(run_masked(ctx, tst, ref, mask, prm, false, -1, zero, L));
  const MaskSummary* sum = masked_summary(ctx, L);
  const uint64_t* rows = (const uint64_t*)((const char*)ctx->h_results + L.off_ssd);
  const Geom& g = tst->g;
  const int32_t nnm = num_non_masked >= 0 ? num_non_masked : sum->nnm;
  const qmiv_ctx::WsTable& wst = ws_table(ctx, g.h, prm.lat_range_deg);
  const std::vector<double>& ws = wst.w;
  const double corr = wst.corr;
  for (int c = 0; c < 3; c++) {
    const uint64_t* r = rows + (size_t)c * g.h;
    wspsnr[c] = fin::wspsnr_from_rows_masked(r, g.h, ws.data(), corr, nnm, g.bd, mask->g.bd);
    // xWSPSNR::xCalcCmpWSPSNRM (libs/pmbb-ivqm/src/xWSPSNR.cpp:148-152): fake where the weighted
    // error is 0, the one case in which the masked PSNR is DBL_MAX.
    if (prm.fake_psnr_for_exact && wspsnr[c] == DBL_MAX) wspsnr[c] = fin::psnr_from_ssd(1.0, g.w * g.h, g.bd);
  }
  return QMIV_OK;
}

qmiv_status qmiv_calc_ivpsnr_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                    const qmiv_params* prm_in, int32_t num_non_masked, const int32_t gcd_ref2tst[4],
                                    double* ivpsnr, double* r2t, double* t2r) {
  if (!ctx || !ivpsnr) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  QMIV_TRY(check_pair(tst, ref));
  QMIV_TRY(check_mask(tst, mask));
  const qmiv_params prm = masked_params(prm_in, true);
  MaskedLayout L;
  // gcd_ref2tst null: xIVPSNRM::calcPicIVPSNRM(Tst, Ref, Msk, TstI, RefI) counts the mask and
  // computes the masked GCD itself (libs/pmbb-ivqm/src/xIVPSNR.cpp:138-150).
  const int32_t nnm_arg = gcd_ref2tst ? num_non_masked : -1;
  QMIV_TRY(run_masked(ctx, tst, ref, mask, prm, true, -1, gcd_ref2tst, L));
  const MaskSummary* sum = masked_summary(ctx, L);
  const Geom& g = tst->g;
  const int32_t nnm = nnm_arg >= 0 ? nnm_arg : sum->nnm;
  static const std::vector<double> kNoWs;
  const qmiv_ctx::WsTable* wst = prm.erp ? &ws_table(ctx, g.h, prm.lat_range_deg) : nullptr;
  const std::vector<double>& ws = wst ? wst->w : kNoWs;  // IV-PSNR's WS sums take no H / sum(w) correction
  const uint64_t* rows = (const uint64_t*)((const char*)ctx->h_results + L.off_cps);
  const double* w = prm.erp ? ws.data() : nullptr;
  const double a = fin::qual_from_rows_masked(rows, g.h, w, nnm, g.bd, mask->g.bd, prm.cmp_weights_average);
  const double b = fin::qual_from_rows_masked(rows + 3 * (size_t)g.h, g.h, w, nnm, g.bd, mask->g.bd,
                                              prm.cmp_weights_average);
  if (r2t) *r2t = a;
  if (t2r) *t2r = b;
  *ivpsnr = fin::min2(a, b);
  return QMIV_OK;
}

}  // extern "C"

// =============================================================================================
// Masked batches: a masked frame set (as qmiv_frames_masked) captured once into a graph.
// =============================================================================================
namespace {

struct MaskedBatchData {
  std::vector<const qmiv_pic*> mask;
  qmiv_params prm{};
  uint32_t metrics = 0;
  std::vector<MaskedPlan> plans;
};

void destroy_masked_batch(MaskedBatchData* m) { delete m; }

MaskedBatchData* copy_masked_batch(const MaskedBatchData* m) { return m ? new MaskedBatchData(*m) : nullptr; }

const qmiv_params& masked_batch_params(const MaskedBatchData* m) { return m->prm; }

// The masked counterpart of batch_build: plans every view as qmiv_frames_masked does, assigns views
// to branches (longest work first onto the least loaded branch), lays out the buffers and captures
// each view's masked flow (statistics, masked GCD, search, results copy) on its branch.
qmiv_status batch_build_masked(qmiv_batch* b, int32_t branches) {
  qmiv_ctx* ctx = b->ctx;
  MaskedBatchData& m = *b->mb;
  batch_drop_graph(b);
  const size_t n = (size_t)b->n;
  const bool iv = (m.metrics & QMIV_IVPSNR) != 0;
  // Without IV-PSNR the application computes no colour difference and reports zeros.
  static const int32_t kZero[4] = {0, 0, 0, 0};
  const int32_t* gcd_given = iv ? nullptr : kZero;
  m.plans.assign(n, MaskedPlan{});
  b->doff.assign(n, 0);
  b->hoff.assign(n, 0);
  size_t dtot = 0, htot = 0;
  std::vector<double> cost(n);
  for (size_t i = 0; i < n; i++) {
    QMIV_TRY(plan_masked(ctx, b->p0[i]->g, m.mask[i]->g, m.prm, m.prm.invalid_pel_action, iv, gcd_given, m.plans[i]));
    b->doff[i] = dtot;
    b->hoff[i] = htot;
    dtot += align_up(m.plans[i].L.total, 256);
    htot += align_up(m.plans[i].L.results_bytes, 256);
    const Geom& g = b->p0[i]->g;
    cost[i] = (double)g.w * g.h * (1.0 + (iv ? 4.0 : 0.0));
  }
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return cost[a] > cost[c]; });
  std::vector<double> load((size_t)branches, 0.0);
  b->branch_of.assign(n, 0);
  for (size_t i : order) {
    const size_t br = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
    b->branch_of[i] = (int32_t)br;
    load[br] += cost[i];
  }
  QMIV_HIP(hipStreamSynchronize(ctx->stream));
  QMIV_TRY(batch_grow(b->d_scratch, b->scratch_bytes, dtot, false));
  QMIV_TRY(batch_grow(b->h_results, b->results_bytes, htot, true));
  QMIV_TRY(batch_capture(b, branches, [&](size_t i, hipStream_t st) {
    return enqueue_masked(ctx, b->p0[i], b->p1[i], m.mask[i], m.prm, m.plans[i], -1, gcd_given,
                          (char*)b->d_scratch + b->doff[i], (char*)b->h_results + b->hoff[i], st);
  }));
  b->info.cps_sweep = 0;
  return QMIV_OK;
}

}  // namespace

extern "C" {

qmiv_status qmiv_batch_create_masked(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                                     const qmiv_pic* const* mask, const qmiv_params* prm_in, uint32_t metrics,
                                     const qmiv_batch_options* opt_in, qmiv_batch** out) {
  if (!ctx || !p0 || !p1 || !mask || !out || n <= 0) return fail(QMIV_ERR_INVALID_ARG, "null argument or empty batch");
  *out = nullptr;
  qmiv_batch_options opt;
  qmiv_batch_options_default(&opt);
  if (opt_in) opt = *opt_in;
  if (opt.branches < 0 || opt.branches > kMaxBranches) return fail(QMIV_ERR_INVALID_ARG, "branches must be 0..16");
  if (opt.tune_reps <= 0) opt.tune_reps = 12;
  for (int32_t i = 0; i < n; i++) {
    QMIV_TRY(check_pair(p0[i], p1[i]));
    QMIV_TRY(check_mask(p0[i], mask[i]));
    if (p0[i]->ctx != ctx || mask[i]->ctx != ctx) return fail(QMIV_ERR_INVALID_ARG, "pictures belong to another context");
  }
  if (metrics & (QMIV_SSIM | QMIV_MSSSIM | QMIV_IVSSIM))
    return fail(QMIV_ERR_INVALID_ARG, "structural similarity metrics cannot be combined with mask mode");
  if (metrics & ~(uint32_t)(QMIV_PSNR | QMIV_WSPSNR | QMIV_IVPSNR)) return fail(QMIV_ERR_INVALID_ARG, "unknown metric bits");
  const qmiv_params prm = masked_params(prm_in, false);
  if (prm.invalid_pel_action < QMIV_PEL_STOP || prm.invalid_pel_action > QMIV_PEL_CNCL)
    return fail(QMIV_ERR_INVALID_ARG, "invalid_pel_action must be 0..3");
  QMIV_HIP(hipSetDevice(ctx->device));
  qmiv_batch* b = new qmiv_batch;
  b->ctx = ctx;
  b->n = n;
  b->p0.assign(p0, p0 + n);
  b->p1.assign(p1, p1 + n);
  b->mb = new MaskedBatchData;
  b->mb->mask.assign(mask, mask + n);
  b->mb->prm = prm;
  b->mb->metrics = metrics;
  std::string sig;
  sig.append((const char*)&metrics, sizeof(metrics));
  sig.append((const char*)&prm, sizeof(prm));
  for (int32_t i = 0; i < n; i++) {
    sig.append((const char*)&p0[i]->g, sizeof(Geom));
    sig.append((const char*)&mask[i]->g, sizeof(Geom));
  }
  return batch_tune_finish(b, opt, sig, out);
}

qmiv_status qmiv_batch_collect_masked(qmiv_batch* b, qmiv_masked_result* out) {
  if (!b || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (!b->mb) return fail(QMIV_ERR_INVALID_ARG, "not a masked batch: use qmiv_batch_collect");
  if (!b->pending) return fail(QMIV_ERR_INVALID_ARG, "the batch has no replay in flight");
  QMIV_HIP(hipSetDevice(b->ctx->device));
  b->pending = false;
  QMIV_HIP(hipEventSynchronize(b->done));
  const MaskedBatchData& m = *b->mb;
  std::vector<qmiv_status> st((size_t)b->n, QMIV_OK);
  auto fin_view = [&](int i) {
    const size_t k = (size_t)i;
    std::memset(&out[i], 0, sizeof(out[i]));
    st[k] = finalize_masked(b->ctx, b->p0[k]->g, m.mask[k]->g.bd, m.prm, m.metrics, m.plans[k].L,
                            (const char*)b->h_results + b->hoff[k], &out[i]);
  };
  if (b->pool) b->pool->run(b->n, fin_view);
  else for (int32_t i = 0; i < b->n; i++) fin_view(i);
  for (int32_t i = 0; i < b->n; i++)
    if (st[(size_t)i] != QMIV_OK)
      return fail(st[(size_t)i], "view " + std::to_string(i) + ": input or mask contains pels above the bit-depth maximum");
  return QMIV_OK;
}

qmiv_status qmiv_batch_run_masked(qmiv_batch* b, qmiv_masked_result* out) {
  if (!b || !out) return fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (!b->mb) return fail(QMIV_ERR_INVALID_ARG, "not a masked batch: use qmiv_batch_run");
  QMIV_TRY(qmiv_batch_launch(b));
  return qmiv_batch_collect_masked(b, out);
}

}  // extern "C"
