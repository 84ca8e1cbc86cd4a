/*
  qmiv.h - C ABI of the MI355X (gfx950) per-frame metric engine.

  The engine computes IV-PSNR (corresponding-pixel-shift search with global
  colour difference, both directions), PSNR, WS-PSNR, SSIM, MS-SSIM and IV-SSIM
  for pairs of pictures resident in device memory, with results identical to
  the reference CPU implementation (integer sums and the PSNR family bit-exact;
  SSIM family within 1e-9 relative).

  The reference exposes this functionality as C++ processor classes with no
  FFI of its own; each entry point below names the reference interface it
  replaces (file:line, relative to the reference repository root).

  Conventions
    - Plain C types only. Every call returns a qmiv_status; no exception or
      C++ type crosses this boundary. qmiv_last_error() describes the last
      failure on the calling thread.
    - A context owns one HIP stream on one device and the device scratch its
      calls need. Calls on one context are not thread-safe; use one context
      per host thread.
    - Pictures hold one frame in the raw planar file layout (Y plane, then Cb,
      then Cr; 4:2:0 chroma planes are W/2 x H/2; one byte per sample at 8 bits,
      little-endian uint16 above). "P0" and "P1" are the reference
      application's InputFile0 and InputFile1.
    - Metric calls are synchronous: they enqueue the kernels on the context's
      stream, wait for the small result block and finish the double-precision
      arithmetic on the host.
*/
#ifndef QMIV_H
#define QMIV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMIV_ABI_VERSION 2

typedef enum qmiv_status {
  QMIV_OK = 0,
  QMIV_ERR_INVALID_ARG = 1,   /* null pointer, bad size, incompatible pictures */
  QMIV_ERR_UNSUPPORTED = 2,   /* parameter combination this engine does not implement */
  QMIV_ERR_HIP = 3,           /* HIP runtime error (message in qmiv_last_error) */
  QMIV_ERR_OUT_OF_MEMORY = 4,
  QMIV_ERR_PEL_RANGE = 5,     /* a pel exceeds 2^bit_depth - 1 (reference's InvalidPelActn=STOP) */
} qmiv_status;

/* Metric selection bits (order of the reference's metric list, apps/QMIV/src/xAppQMIV.h:119-131). */
enum {
  QMIV_PSNR = 1,
  QMIV_WSPSNR = 2,
  QMIV_IVPSNR = 4,
  QMIV_SSIM = 8,
  QMIV_MSSSIM = 16,
  QMIV_IVSSIM = 32,
};

typedef struct qmiv_ctx qmiv_ctx;
typedef struct qmiv_pic qmiv_pic;

/* Processor parameters. Defaults (qmiv_params_default) are the reference's:
   libs/pmbb-ivqm/src/xCorrespPixelShift.h:30-32 (B=2, weights 4:1:1:0),
   libs/pmbb-ivqm/src/xGlobClrDiff.h:18 (unnoticeable coefficient 0.01),
   libs/pmbb-ivqm/src/xWeightedSpherically.h:20 (WS off; 360 x 180 degrees). */
typedef struct qmiv_params {
  int32_t search_range;            /* B: the search window is (2B+1) x (2B+1) */
  int32_t cmp_weights_search[4];   /* component weights inside the search */
  int32_t cmp_weights_average[4];  /* component weights of the final average */
  float unnoticeable_coef[4];      /* GCD clip: +-rint(coef * max) */
  int32_t erp;                     /* 1: equirectangular, WS weights on (xWeightedSpherically::initWS) */
  int32_t lon_range_deg, lat_range_deg;
  /* 1 (default): per-row CPS sums wrap in int32 and are sign-extended, as the reference's
     SSE row kernel does (xCorrespPixelShiftSSE.cpp:27-38, the application default);
     0: exact uint64 row sums (xCorrespPixelShiftSTD.cpp:26-36). Identical unless a
     row's sum exceeds 2^31 - 1. */
  int32_t cps_int32_rows;
  /* 1 (default): qmiv_frame replaces the PSNR / WS-PSNR of identical components by the
     PSNR of SSD=1, as the application does (apps/QMIV/src/xAppQMIV.cpp:837-844,864-871). */
  int32_t fake_psnr_for_exact;
  /* What the frame flow does with a pel above 2^bd - 1, the application's InvalidPelActn
     (apps/QMIV/src/xAppQMIV.cpp:681-683,753-770): QMIV_PEL_STOP (default) returns
     QMIV_ERR_PEL_RANGE; QMIV_PEL_WARN and QMIV_PEL_SKIP compute every metric over the samples as
     they are (the in_range flags still report the check); QMIV_PEL_CNCL clips the offending
     pictures to 2^bd - 1 in place (xPicP::conceal, ClipToRange, libs/pmbb-core/src/xPixelOpsSTD.cpp:198-209)
     and computes the metrics over the clipped pictures (in_range reports the pictures as read).
     Processor-level calls (qmiv_calc_*, qmiv_gen_scp*) never validate, like the reference's
     processors. */
  int32_t invalid_pel_action;
} qmiv_params;

enum { QMIV_PEL_STOP = 0, QMIV_PEL_WARN = 1, QMIV_PEL_SKIP = 2, QMIV_PEL_CNCL = 3 };

typedef struct qmiv_frame_result {
  int32_t in_range[2];             /* per input: all pels <= 2^bd - 1 (xPicP::check) */
  int32_t exact[3];                /* per component: pictures identical (xPicP::equalCmp) */
  int32_t gcd[4];                  /* GCD R2T (xAppQMIV.cpp:693-698) */
  double psnr[3], psnr_pic;
  double wspsnr[3], wspsnr_pic;
  double ivpsnr, ivpsnr_r2t, ivpsnr_t2r;
  uint64_t cps_sum[2][3];          /* [R2T, T2R][component]: sum over rows of the per-row values */
  double ssim[3], ssim_pic;
  double msssim[3], msssim_pic;
  double msssim_scales[5][3];
  double ivssim, ivssim_r2t, ivssim_t2r;
} qmiv_frame_result;

/* ------------------------------------------------------------------------------------------ */
/* Context and pictures                                                                        */
/* ------------------------------------------------------------------------------------------ */

/* Creates a context on HIP device `device` with its own stream. Replaces the reference's
   thread pool set-up (apps/QMIV/src/xAppQMIV.cpp:458-471, libs/pmbb-core/src/xThreadPool.h:20-134). */
qmiv_status qmiv_create(int32_t device, qmiv_ctx** out);
void qmiv_destroy(qmiv_ctx* ctx);
/* Message of the last failed call on this thread ("" if none). */
const char* qmiv_last_error(void);
int32_t qmiv_abi_version(void);
/* The context's HIP stream (hipStream_t), for callers that order their own work with it. */
void* qmiv_stream(qmiv_ctx* ctx);
/* Waits until all work enqueued on the context's stream has finished (uploads included). Replaces
   xThreadPoolInterface::waitUntilTasksFinished (libs/pmbb-core/src/xThreadPool.cpp:147-170). */
qmiv_status qmiv_sync(qmiv_ctx* ctx);

void qmiv_params_default(qmiv_params* p);

/* Bytes of one packed frame of a raw file in chroma format 400, 420, 422 or 444 (any width and
   height): xSeq::create (libs/pmbb-core/src/xSeq.cpp:408-428). 0 for an invalid format. */
size_t qmiv_frame_bytes(int32_t width, int32_t height, int32_t bit_depth, int32_t chroma_format);

/* Allocates a device picture for one packed frame (bit_depth 8..16; chroma_format 420 with even
   width and height, 444, or 400: a luma plane only, for masks). Replaces xPicP::create
   (libs/pmbb-core/src/xPic.h:31) plus the margin (the engine addresses edges by clamping instead,
   identical to ExtendMargin, libs/pmbb-core/src/xPixelOpsSTD.cpp:210-236). Frames of other file
   formats (4:2:2, 4:2:0 of odd width) go into 4:4:4 pictures through qmiv_pic_upload_unpack. */
qmiv_status qmiv_pic_create(qmiv_ctx* ctx, int32_t width, int32_t height, int32_t bit_depth,
                            int32_t chroma_format, qmiv_pic** out);
/* Wraps caller-owned device memory holding one packed frame (not freed by the engine). */
qmiv_status qmiv_pic_wrap(qmiv_ctx* ctx, int32_t width, int32_t height, int32_t bit_depth,
                          int32_t chroma_format, void* device_ptr, qmiv_pic** out);
/* Frees an engine-owned picture (waits for outstanding device work); valid after qmiv_destroy. */
void qmiv_pic_destroy(qmiv_pic* pic);
void* qmiv_pic_device_ptr(qmiv_pic* pic);
/* Copies one packed frame from host memory (qmiv_frame_bytes bytes) onto the device with one
   hipMemcpyAsync ordered on the context's stream (work enqueued afterwards sees the frame). From
   pinned host memory the copy is asynchronous: keep the memory unchanged until the stream has
   passed it (qmiv_sync). From pageable memory the data goes through the context's pinned staging
   slots (four of 8 MiB, created on first use; host threads fill a slot while the previous one is
   copied) and the call returns once the source has been read in full. Replaces
   xSeqBase::readFrame's unpack step (libs/pmbb-core/src/xSeq.cpp:179-249): the engine keeps the
   packed layout. */
qmiv_status qmiv_pic_upload(qmiv_ctx* ctx, qmiv_pic* pic, const void* host, size_t bytes);
/* Copies a device picture back to host memory (packed layout). */
qmiv_status qmiv_pic_download(qmiv_ctx* ctx, const qmiv_pic* pic, void* host, size_t bytes);
/* xSeqBase::xUnpackFrame (libs/pmbb-core/src/xSeq.cpp:179-249) on the device: one packed raw-file
   frame (qmiv_frame_bytes of file_chroma_format bytes, of the picture's size and bit depth) becomes
   the picture. The same format is a plain copy (qmiv_pic_upload); 400, 420 (any width) and 422
   frames unpack into a 4:4:4 picture: chroma nearest-neighbour upsampled from rows of width >> 1
   samples (libs/pmbb-core/src/xSeq.cpp:197-229; for an odd width the last column reads the next
   chroma row's first sample, as the reference does; a read past the frame's end gives 0), 4:0:0
   chroma set to 2^(bd - 1). Asynchronous on the context's stream as qmiv_pic_upload; the frame goes
   through a context-owned device staging buffer. */
qmiv_status qmiv_pic_upload_unpack(qmiv_ctx* ctx, qmiv_pic* pic, const void* host, size_t bytes,
                                   int32_t file_chroma_format);
/* The reference's own picture layouts, for callers that hold them. xPicP
   (libs/pmbb-core/src/xPicCommon.cpp:16-33, xPic.h:18-74): three uint16 planes whose first interior
   sample planes[c] points at, rows `stride` samples apart (W + 2 margin); the margin is neither
   read nor written. xPicI (xPic.h:79-113, xPic.cpp:169-175,201-207): 4 x uint16 per pel (Y, Cb,
   Cr, A), origin at the first interior pel, rows `stride` pels apart; A is ignored on upload and
   written 0. The device picture is 4:4:4 of the same size and bit depth; an 8-bit picture takes
   samples up to 255 (else QMIV_ERR_PEL_RANGE). Uploads are ordered on the context's stream and
   return once the host memory may be reused; downloads wait for the data. */
qmiv_status qmiv_pic_upload_planar(qmiv_ctx* ctx, qmiv_pic* pic, const uint16_t* const planes[3], int32_t stride);
qmiv_status qmiv_pic_download_planar(qmiv_ctx* ctx, const qmiv_pic* pic, uint16_t* const planes[3], int32_t stride);
/* xPicI::rearrangeFromPlanar's input side (libs/pmbb-core/src/xPic.cpp:201-207) from host memory. */
qmiv_status qmiv_pic_upload_interleaved(qmiv_ctx* ctx, qmiv_pic* pic, const uint16_t* origin, int32_t stride);
qmiv_status qmiv_pic_download_interleaved(qmiv_ctx* ctx, const qmiv_pic* pic, uint16_t* origin, int32_t stride);
/* xPicP::copy (libs/pmbb-core/src/xPic.cpp:46-51) between device pictures of one geometry;
   asynchronous. */
qmiv_status qmiv_pic_copy(qmiv_ctx* ctx, qmiv_pic* dst, const qmiv_pic* src);
/* xPicP::check (libs/pmbb-core/src/xPic.cpp:62-82, xPixelOpsSTD::CheckIfInRange, :163-174): per
   component, every sample <= 2^bd - 1. Synchronous. */
qmiv_status qmiv_pic_check(qmiv_ctx* ctx, const qmiv_pic* pic, int32_t in_range[3]);
/* xPicP::conceal (libs/pmbb-core/src/xPic.cpp:83-90, xPixelOpsSTD::ClipToRange, :198-209): samples
   above 2^bd - 1 become 2^bd - 1, in place. Asynchronous on the context's stream. */
qmiv_status qmiv_pic_conceal(qmiv_ctx* ctx, qmiv_pic* pic);
/* xPicP::equalCmp per component (libs/pmbb-core/src/xPic.cpp:103-107, CompareEqual,
   xPixelOpsSTD.cpp:285-295). Synchronous. */
qmiv_status qmiv_pic_equal(qmiv_ctx* ctx, const qmiv_pic* a, const qmiv_pic* b, int32_t equal[3]);

/* ------------------------------------------------------------------------------------------ */
/* Sequences: raw planar YUV files (the read side of xSeq, libs/pmbb-core/src/xSeq.cpp) and     */
/* PNG file lists (xSeqPNG, libs/pmbb-spng/src/xSeqPNG.cpp)                                   */
/* ------------------------------------------------------------------------------------------ */
typedef struct qmiv_seq qmiv_seq;

/* Opens a raw planar file of packed frames (chroma format 400, 420, 422 or 444, any size) for
   reading. The frame count is floor(file size / frame bytes): a partial last frame is ignored
   (xSeq::xBackendOpen, xSeq.cpp:451-470; frame bytes as xSeq::create, :408-428). Host-only: no
   device is touched. */
qmiv_status qmiv_seq_open(const char* path, int32_t width, int32_t height, int32_t bit_depth,
                          int32_t chroma_format, qmiv_seq** out);
/* Opens a list of PNG files as a sequence: xSeqPNG (libs/pmbb-spng/src/xSeqPNG.cpp:55-148). The
   frame file names are the pattern formatted with the frame's file index the way fmt::format does
   for the fields {}, {:Nd} and {:0Nd} (e.g. "view_{:03d}.png"). The first index is the last of 0
   and 1 whose file exists; the frames are the files from there on, up to the first missing index
   (at most 65535). Frames are 8-bit 4:4:4 pictures with the planes R, G, B: each file must be
   width x height at bit depth 8 (any colour type; grey is replicated, palettes looked up, alpha
   dropped, as libspng's RGB8 output without transparency or gamma processing). Decoding happens on
   the host at read time (qmiv_seq_read_host / qmiv_seq_read); a file that does not decode makes
   that read fail. Host-only at open: no device is touched. */
qmiv_status qmiv_seq_open_png(const char* pattern, int32_t width, int32_t height, qmiv_seq** out);
int32_t qmiv_seq_num_frames(const qmiv_seq* seq);
size_t qmiv_seq_frame_bytes(const qmiv_seq* seq);
/* Reads frame `index` (seek by frame, xSeq.cpp:519-525) into host memory of the frame's size. */
qmiv_status qmiv_seq_read_host(qmiv_seq* seq, int32_t index, void* dst, size_t bytes);
/* Reads frame `index` into a device picture of the sequence's size and bit depth, ordered on the
   context's stream (work enqueued afterwards sees the frame): the file is read into a pinned buffer
   (by a background thread when the frame was prefetched: reading frame i prefetches frame i + 1)
   and copied asynchronously. Replaces xSeqBase::readFrame (xSeq.cpp:155-174): a picture of the
   file's own format takes the frame as it is; 4:2:2, 4:0:0 and odd-width 4:2:0 frames unpack into
   a 4:4:4 picture as qmiv_pic_upload_unpack does. A sequence serves one device. */
qmiv_status qmiv_seq_read(qmiv_ctx* ctx, qmiv_seq* seq, int32_t index, qmiv_pic* pic);
/* Waits for outstanding copies out of the sequence's buffers, then closes the file. */
void qmiv_seq_close(qmiv_seq* seq);

/* ------------------------------------------------------------------------------------------ */
/* Timing (replaces the reference's per-stage RDTSC timing, apps/QMIV/src/xAppQMIV.cpp:670-744) */
/* ------------------------------------------------------------------------------------------ */

/* Kernel kinds reported by qmiv_profile_read. */
enum {
  QMIV_K_STATS = 0,     /* per-row SD / SSD / range statistics */
  QMIV_K_SUMMARY = 1,   /* plane totals and GCD */
  QMIV_K_CPS = 2,       /* corresponding-pixel-shift search (IV-PSNR, SCP): the fast kernels */
  QMIV_K_CPS_ROWS = 3,  /* per-row reduction of the CPS partial sums */
  QMIV_K_SSIM = 4,      /* SSIM windowed moments */
  QMIV_K_PYRAMID = 5,   /* MS-SSIM 2x2 pyramid */
  QMIV_K_SSIM_ROWS = 6, /* per-row reduction of the SSIM tile partial sums */
  QMIV_K_CPS_GENERAL = 7, /* the search's general kernels (any search range, weights, bit depth) */
  QMIV_K_COUNT = 8
};

typedef struct qmiv_profile {
  double kernel_ms[QMIV_K_COUNT];   /* accumulated device time, measured with HIP events */
  int64_t launches[QMIV_K_COUNT];
} qmiv_profile;

/* When on, every kernel launch of the context is bracketed by HIP events on its stream. */
qmiv_status qmiv_profile_enable(qmiv_ctx* ctx, int32_t on);
/* Accumulated kernel times since the last reset; reset != 0 clears them. */
qmiv_status qmiv_profile_read(qmiv_ctx* ctx, qmiv_profile* out, int32_t reset);

/* ------------------------------------------------------------------------------------------ */
/* Frame flow                                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* One frame through the reference application's per-frame flow: validation, exact-component
   detection, GCD, SCP, then the selected metrics (apps/QMIV/src/xAppQMIV.cpp:661-935).
   `metrics` is a mask of QMIV_* bits. Returns QMIV_ERR_PEL_RANGE (with `out` filled up to the
   flags) when a pel exceeds the bit depth, as InvalidPelActn=STOP does (xAppQMIV.cpp:753-770). */
qmiv_status qmiv_frame(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_params* prm,
                       uint32_t metrics, qmiv_frame_result* out);

/* qmiv_frame without allocation: no counterpart in the reference. For callers that cannot
   allocate memory while they run (serving loops that forbid hipMalloc, harnesses that pre-reserve
   all memory). Same flow, arguments, results and errors as qmiv_frame, on the context's stream,
   but the call allocates no device or pinned host memory: the frame's device scratch (and, with
   QMIV_IVSSIM, its two shift-compensated pictures) lies in one fixed 64 MiB block per device,
   reserved when the library is loaded onto the device and held until the process ends, and its
   results return through a 2 MiB pinned block allocated by qmiv_create. A frame needing more of
   either returns QMIV_ERR_UNSUPPORTED with the sizes in qmiv_last_error (e.g. 4096 x 2048 with
   IV-SSIM; every metric but IV-SSIM fits 4096 x 2048, all six fit 1920 x 1080). The block keeps
   the last call's contents and nothing in it needs clearing. Every context on one device shares
   that device's block: calls on different contexts of one device must not overlap (they race). */
qmiv_status qmiv_frame_noalloc(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_params* prm,
                               uint32_t metrics, qmiv_frame_result* out);

/* The same flow for n frame pairs (sizes may differ between pairs): every frame's kernels are
   enqueued back to back on the context's stream, then one wait, then the host finalisation of
   each. Replaces the reference application's frame loop (apps/QMIV/src/xAppQMIV.cpp:666-748).
   out[i] receives frame i; the first failing frame's status is returned. */
qmiv_status qmiv_frames(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                        const qmiv_params* prm, uint32_t metrics, qmiv_frame_result* out);

/* qmiv_frames with each frame's upload: the packed frames h0[i], h1[i] (qmiv_frame_bytes each, in
   the file layout qmiv_pic_upload takes) are copied into p0[i], p1[i] on the context's copy stream,
   and frame i's kernels wait for its own two copies only, so the copies of frame i + 1 run while
   frame i is computed. Replaces the reference application's read-then-process frame loop
   (libs/pmbb-core/src/xSeq.cpp:155-174 per frame, apps/QMIV/src/xAppQMIV.cpp:666-748). The copies
   run at the link's rate from pinned memory (qmiv_host_alloc); from ordinary memory through the
   context's pinned staging slots, as qmiv_pic_upload. No picture may appear twice among the 2n pictures
   (QMIV_ERR_INVALID_ARG). The sources are no longer read once the call has returned. */
qmiv_status qmiv_frames_upload(qmiv_ctx* ctx, int32_t n, qmiv_pic* const* p0, qmiv_pic* const* p1,
                               const void* const* h0, const void* const* h1, const qmiv_params* prm,
                               uint32_t metrics, qmiv_frame_result* out);

/* Pinned (page-locked) host memory of the context's device, for frames handed to the engine:
   host-to-device copies from it run asynchronously at the link's rate. No counterpart in the
   reference, whose reader fills ordinary memory (libs/pmbb-core/src/xSeq.cpp:39-56).
   qmiv_host_free reads nothing of ctx and may run after qmiv_destroy. */
qmiv_status qmiv_host_alloc(qmiv_ctx* ctx, size_t bytes, void** ptr);
qmiv_status qmiv_host_free(qmiv_ctx* ctx, void* ptr);

/* ------------------------------------------------------------------------------------------ */
/* Processor-level calls (same argument meaning as the reference's processor methods)          */
/* ------------------------------------------------------------------------------------------ */

/* xGlobClrDiffProc::CalcGlobalColorDiff(Tst, Ref) (libs/pmbb-ivqm/src/xGlobClrDiff.h:41):
   gcd = clip(rint(sum(Ref - Tst) / area), +-rint(coef * max)). */
qmiv_status qmiv_calc_gcd(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm,
                          int32_t gcd[4]);

/* xPSNR::calcPicPSNR (libs/pmbb-ivqm/src/xPSNR.h:33): per component; SSD=0 gives DBL_MAX. */
qmiv_status qmiv_calc_psnr(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, double psnr[3]);

/* xWSPSNR::calcPicWSPSNR (libs/pmbb-ivqm/src/xWSPSNR.h:22); plain PSNR unless prm->erp. */
qmiv_status qmiv_calc_wspsnr(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm,
                             double wspsnr[3]);

/* xIVPSNR::calcPicIVPSNR(Tst, Ref, GlobalColorDiffRef2Tst) (libs/pmbb-ivqm/src/xIVPSNR.h:34):
   R2T = Q(Tst + gcd searched in Ref), T2R = Q(Ref - gcd searched in Tst), result min(R2T, T2R).
   r2t, t2r and rows may be NULL. rows (if given) receives 2*3*height uint64 per-row values
   [direction][component][row] with the prm->cps_int32_rows semantics. */
qmiv_status qmiv_calc_ivpsnr(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm,
                             const int32_t gcd_ref2tst[4], double* ivpsnr, double* r2t, double* t2r,
                             uint64_t* rows);

/* xSSIM::calcPicSSIM(Tst, Ref, CalcL) (libs/pmbb-ivqm/src/xSSIM.h:37): per component; calc_l = 0
   gives the contrast-structure term only. */
qmiv_status qmiv_calc_ssim(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, int32_t calc_l, double ssim[3]);

/* xSSIM::calcPicMSSSIM(Tst, Ref) (libs/pmbb-ivqm/src/xSSIM.h:38): per component; scales (nullable)
   receives the five per-scale scores [scale][component]. */
qmiv_status qmiv_calc_msssim(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, double msssim[3],
                             double* scales);

/* xIVSSIM::calcPicIVSSIM(Tst, Ref) (libs/pmbb-ivqm/src/xSSIM.h:74): GCD, shift-compensated pictures,
   then min(SSIM(Tst, RefSCP), SSIM(Ref, TstSCP)) of the component-weighted averages. */
qmiv_status qmiv_calc_ivssim(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_params* prm,
                             double* ivssim, double* r2t, double* t2r);

/* xShftCompPicProc::GenShftCompPics(DstRef, DstTst, SrcRef, SrcTst, GlobalColorDiffRef2Tst)
   (libs/pmbb-ivqm/src/xShftCompPic.h:36): DstRef(p) = clip(SrcRef(p + best) - gcd), best searched
   for SrcTst(p) + gcd; DstTst(p) = clip(SrcTst(p + best) + gcd) for SrcRef(p) - gcd. dst_ref and
   dst_tst are host buffers of 3 * width * height uint16 (4:4:4 planes Y, Cb, Cr). */
qmiv_status qmiv_gen_scp(qmiv_ctx* ctx, const qmiv_pic* src_ref, const qmiv_pic* src_tst, const qmiv_params* prm,
                         const int32_t gcd_ref2tst[4], uint16_t* dst_ref, uint16_t* dst_tst);

/* GenShftCompPics into device pictures (libs/pmbb-ivqm/src/xShftCompPic.h:36, the form the
   application calls, apps/QMIV/src/xAppQMIV.cpp:703): dst_ref / dst_tst are 4:4:4 pictures of the
   sources' size and bit depth (the reference's xPicP layout). */
qmiv_status qmiv_gen_scp_pics(qmiv_ctx* ctx, const qmiv_pic* src_ref, const qmiv_pic* src_tst, const qmiv_params* prm,
                              const int32_t gcd_ref2tst[4], qmiv_pic* dst_ref, qmiv_pic* dst_tst);

/* xIVSSIM::calcPicIVSSIM(Tst, Ref, TstSCP, RefSCP) (libs/pmbb-ivqm/src/xSSIM.cpp:192-212): the
   shift-compensated pictures come from the caller; T2R = SSIM(Tst, RefSCP), R2T = SSIM(Ref, TstSCP),
   each averaged with prm->cmp_weights_average, then the minimum. All four pictures share one
   geometry. */
qmiv_status qmiv_calc_ivssim_scp(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* tst_scp,
                                 const qmiv_pic* ref_scp, const qmiv_params* prm, double* ivssim, double* r2t,
                                 double* t2r);

/* Nearest-neighbour 4:2:0 -> 4:4:4 conversion on the device (a copy for 4:4:4 sources) into a
   4:4:4 picture of the same size and bit depth: the unpack step of xSeqBase::readFrame
   (libs/pmbb-core/src/xSeq.cpp:197-213, libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49,68-87).
   Asynchronous on the context stream. */
qmiv_status qmiv_pic_upsample(qmiv_ctx* ctx, const qmiv_pic* src, qmiv_pic* dst);

/* ------------------------------------------------------------------------------------------ */
/* Masked mode: metrics over the pels whose mask is non-zero, weighted by the mask value       */
/* (apps/QMIV/src/xAppQMIV.cpp:319-337,695,825-889). The mask is a picture of the same size    */
/* (any bit depth 8..16, 4:2:0 or 4:4:4) of which the luma plane is used; its margin is        */
/* clamp-to-edge like the pictures'.                                                           */
/* ------------------------------------------------------------------------------------------ */

typedef struct qmiv_masked_result {
  int32_t in_range[3];     /* P0, P1, mask: all pels <= 2^bd - 1 (xPicP::check of the three inputs) */
  int32_t exact[3];        /* per component: P0 and P1 identical over the whole picture */
  int32_t num_non_masked;  /* xPixelOps::CountNonZero of the mask's luma plane (xAppQMIV.cpp:825-829) */
  int32_t gcd[4];          /* masked GCD R2T (xAppQMIV.cpp:695), zeros without IV-PSNR */
  double psnr[3], psnr_pic;
  double wspsnr[3], wspsnr_pic;
  /* The masked processor's labels (libs/pmbb-ivqm/src/xIVPSNR.cpp:152-165): R2T searches P1 - g
     inside P0, T2R searches P0 + g inside P1 (the unmasked processor names them the other way). */
  double ivpsnr, ivpsnr_r2t, ivpsnr_t2r;
  uint64_t cps_sum[2][3];  /* [R2T, T2R][component]: sum over rows of the mask-weighted values */
} qmiv_masked_result;

/* One frame through the application's masked flow: validation of the three inputs, exact
   components, the masked GCD (with IV-PSNR), then PSNR / WS-PSNR / IV-PSNR in their masked forms
   (xPSNR::calcPicPSNRM, xPSNR.cpp:42-71; xWSPSNR::calcPicWSPSNRM, xWSPSNR.cpp:61-97;
   xIVPSNRM::calcPicIVPSNRM, xIVPSNR.cpp:152-209). metrics may hold only QMIV_PSNR, QMIV_WSPSNR and
   QMIV_IVPSNR (the application rejects SSIM metrics with a mask, xAppQMIV.cpp:337). Identical
   components get the PSNR of SSD=1 when prm->fake_psnr_for_exact is set (xAppQMIV.cpp:837-871).
   Returns QMIV_ERR_PEL_RANGE (with the flags filled) when a pel of any input exceeds its bit depth. */
qmiv_status qmiv_frame_masked(qmiv_ctx* ctx, const qmiv_pic* p0, const qmiv_pic* p1, const qmiv_pic* mask,
                              const qmiv_params* prm, uint32_t metrics, qmiv_masked_result* out);

/* The masked flow for n frames (sizes may differ between frames, each with its own mask): every
   frame's kernels are enqueued back to back (dealt round-robin over three streams), then the host
   finalisation of each as its results arrive. Replaces the reference application's frame loop in
   mask mode (apps/QMIV/src/xAppQMIV.cpp:666-748 with the masked calls of :825-889). out[i]
   receives frame i; the first failing frame's status is returned. */
qmiv_status qmiv_frames_masked(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                               const qmiv_pic* const* mask, const qmiv_params* prm, uint32_t metrics,
                               qmiv_masked_result* out);

/* xPixelOps::CountNonZero over the mask's luma plane (libs/pmbb-core/src/xPixelOpsSTD.cpp:273-284). */
qmiv_status qmiv_count_nonzero(qmiv_ctx* ctx, const qmiv_pic* mask, int32_t* out);

/* Processor-level masked calls. num_non_masked < 0 means "count it" (the reference's NOT_VALID).
   prm->fake_psnr_for_exact selects xPSNR::setFakeValsForExact (PSNR of SSD=1 where the masked
   error is 0); a null prm gives the defaults with that flag off, as the processors start. */

/* xGlobClrDiffProc::CalcGlobalColorDiffM(Tst, Ref, Msk, NNM) (libs/pmbb-ivqm/src/xGlobClrDiff.h:42,
   xGlobClrDiff.cpp:52-83): clip(rint(sum((Ref - Tst) * m) / (NNM * maxm)), +-rint(coef * max)). */
qmiv_status qmiv_calc_gcd_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                 const qmiv_params* prm, int32_t num_non_masked, int32_t gcd[4]);
/* xPSNR::calcPicPSNRM (libs/pmbb-ivqm/src/xPSNR.cpp:42-71,84-136,148-155). */
qmiv_status qmiv_calc_psnr_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                  const qmiv_params* prm, int32_t num_non_masked, double psnr[3]);
/* xWSPSNR::calcPicWSPSNRM (libs/pmbb-ivqm/src/xWSPSNR.cpp:61-97,127-155); without prm->erp the
   plain masked PSNR with the mask counted again, as the reference does. */
qmiv_status qmiv_calc_wspsnr_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                    const qmiv_params* prm, int32_t num_non_masked, double wspsnr[3]);
/* xIVPSNRM::calcPicIVPSNRM(Tst, Ref, Msk, NNM, GCD) (libs/pmbb-ivqm/src/xIVPSNR.cpp:152-209). With
   gcd_ref2tst null: calcPicIVPSNRM(Tst, Ref, Msk, TstI, RefI) (:138-150), which counts the mask
   and computes the masked GCD itself (num_non_masked is then ignored). Search ranges 0..12. */
qmiv_status qmiv_calc_ivpsnr_masked(qmiv_ctx* ctx, const qmiv_pic* tst, const qmiv_pic* ref, const qmiv_pic* mask,
                                    const qmiv_params* prm, int32_t num_non_masked, const int32_t gcd_ref2tst[4],
                                    double* ivpsnr, double* r2t, double* t2r);

/* Colour spaces of the RGB <-> YCbCr conversions: eClrSpcLC (libs/pmbb-core/src/xCommonDefCORE.h:356-364;
   SMPTE 170M shares BT.601's matrix). */
enum { QMIV_CLR_BT601 = 0, QMIV_CLR_BT709 = 1, QMIV_CLR_SMPTE240M = 2, QMIV_CLR_BT2020 = 3 };

/* Converts a 4:4:4 picture between R,G,B planes and Y,Cb,Cr planes (to_rgb 0: RGB -> YCbCr,
   1: YCbCr -> RGB) in the reference's Q16 integer arithmetic, bit-exact with
   xColorSpaceSTD::ConvertRGB2YCbCr_I32 / ConvertYCbCr2RGB_I32 (libs/pmbb-core/src/xColorSpaceSTD.cpp:195-272)
   and the matrices of xColorSpaceCoeff (libs/pmbb-core/src/xColorSpaceCoeff.cpp). dst may be src
   (the reference app converts in place: apps/QMIV/src/xAppQMIV.cpp:778-798); otherwise it must be
   a distinct 4:4:4 picture of the same size and bit depth. Asynchronous on the context stream. */
qmiv_status qmiv_pic_colorspace(qmiv_ctx* ctx, const qmiv_pic* src, qmiv_pic* dst, int32_t space, int32_t to_rgb);

/* Exchanges planes a and b (0..2) of a picture with equal plane sizes (4:4:4): xPicP::swapComponents
   (libs/pmbb-core/src/xPic.cpp:158-164), which the reference app uses to reorder BGR and GBR
   input to RGB (apps/QMIV/src/xAppQMIV.cpp:800-813). Asynchronous on the context stream. */
qmiv_status qmiv_pic_swap_planes(qmiv_ctx* ctx, qmiv_pic* pic, int32_t a, int32_t b);

/* ---------------------------------------------------------------------------------------------
   Batched frame sets (no counterpart in the reference, which runs its frame loop,
   apps/QMIV/src/xAppQMIV.cpp:661-935, once per view and process). A batch binds n picture pairs
   of any sizes and bit depths, each with its own parameters and metrics, and captures all their
   work into one HIP graph whose branches run views concurrently. qmiv_batch_run replays it once
   per frame set: upload the next frames into the same pictures, then run. An autotuner picks
   the branch count and the cap on CPS tiles per workgroup by timing replays at creation (the
   pictures' contents at that time do not matter). Destroy batches before their context.
   --------------------------------------------------------------------------------------------- */
typedef struct qmiv_batch qmiv_batch;

typedef struct qmiv_batch_options {
  int32_t branches;   /* concurrent graph branches, 1..16; 0: autotuned over {1, 2, 4, 8, 16} */
  int32_t cps_sweep;  /* cap on CPS tiles per workgroup; 0: occupancy rule; -1: autotuned over {rule, 1, 2, 4} */
  int32_t tune_reps;  /* replays per autotuner candidate and round (default 12), each launched and waited for */
} qmiv_batch_options;
/* The autotuner first replays the first candidate for about 30 ms (device clocks out of idle), then
   builds and times every candidate once per round in 8 interleaved rounds (tune_reps replays per
   round, each launched and waited for on the host clock as qmiv_batch_run does; their mean) and
   ranks candidates by their fastest round. A captured graph's replay time depends on its instance
   as well as its shape (the runtime lays each instance's branches over the device's hardware
   queues; instances of one shape replay up to 25% apart), hence the fastest round. A candidate with
   more branches (or the occupancy rule for the sweep) is kept unless another one is more than 3%
   faster. The choice is remembered per context and batch signature (the views' sizes, formats,
   metrics and parameters): a later qmiv_batch_create with the same signature and autotuned options
   takes it without timing again (from_cache = 1). Tuned or cached, the instance kept is timed and
   built again while it replays more than 8% slower than the shape's fastest, at most 6 builds. */

typedef struct qmiv_batch_info {
  int32_t views;
  int32_t branches;          /* chosen branch count */
  int32_t cps_sweep;         /* chosen sweep cap (0 = occupancy rule) */
  int32_t candidates_tried;  /* autotuner candidate timings (rounds included) */
  double tuned_ms;           /* replay time of the chosen shape's fastest instance while tuning (0: shape given) */
  int32_t tune_rounds;       /* rounds per candidate (0 when not tuned) */
  int32_t from_cache;        /* 1: the choice came from an earlier batch of the same signature */
  int32_t instances;         /* graph instances built by the instance check (0: no check, shape given) */
  double instance_ms;        /* replay time of the instance kept, as the check measured it */
} qmiv_batch_info;

void qmiv_batch_options_default(qmiv_batch_options* opt);
/* prm: n parameter sets (NULL: defaults); metrics: n metric masks. */
qmiv_status qmiv_batch_create(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                              const qmiv_params* prm, const uint32_t* metrics, const qmiv_batch_options* opt,
                              qmiv_batch** out);
/* One replay of the captured frame set; out receives n results (as qmiv_frame's). Equals
   qmiv_batch_launch followed by qmiv_batch_collect. */
qmiv_status qmiv_batch_run(qmiv_batch* batch, qmiv_frame_result* out);
/* The two halves of qmiv_batch_run (no counterpart in the reference; for callers that keep the
   device busy across frame sets). qmiv_batch_launch enqueues one replay on the context stream and
   returns; qmiv_batch_collect waits for that replay and finalises its views on the host. Replays of
   different batches run one after the other in launch order, so a caller that launches frame set
   k + 1 before collecting set k finalises set k while the device computes set k + 1. A batch holds
   at most one replay in flight (its scratch and results are reused): launching it again before the
   collect, or collecting with nothing launched, fails with QMIV_ERR_INVALID_ARG. Uploads into a
   batch's pictures go on the context stream, after the replays launched before them. */
qmiv_status qmiv_batch_launch(qmiv_batch* batch);
qmiv_status qmiv_batch_collect(qmiv_batch* batch, qmiv_frame_result* out);
/* A masked frame set captured once into a HIP graph (no counterpart in the reference, whose
   application runs masked frames one by one, apps/QMIV/src/xAppQMIV.cpp:666-748 with :825-889): the
   n frames (p0[i], p1[i], mask[i]) with one parameter set and metric mask, each view's flow as
   qmiv_frames_masked enqueues it, its branch count tuned as qmiv_batch_create tunes it (the sweep
   option does not apply: the masked search takes its own). Replayed by qmiv_batch_run_masked, or
   qmiv_batch_launch followed by qmiv_batch_collect_masked; out receives n results, as
   qmiv_frame_masked's. The unmasked run and collect refuse a masked batch and the masked ones an
   unmasked batch (QMIV_ERR_INVALID_ARG); qmiv_batch_get_info and qmiv_batch_destroy serve both. */
qmiv_status qmiv_batch_create_masked(qmiv_ctx* ctx, int32_t n, const qmiv_pic* const* p0, const qmiv_pic* const* p1,
                                     const qmiv_pic* const* mask, const qmiv_params* prm, uint32_t metrics,
                                     const qmiv_batch_options* opt, qmiv_batch** out);
qmiv_status qmiv_batch_run_masked(qmiv_batch* batch, qmiv_masked_result* out);
qmiv_status qmiv_batch_collect_masked(qmiv_batch* batch, qmiv_masked_result* out);
qmiv_status qmiv_batch_get_info(const qmiv_batch* batch, qmiv_batch_info* info);
void qmiv_batch_destroy(qmiv_batch* batch);

/* ---------------------------------------------------------------------------------------------
   Differentiable SSIM family (no counterpart in the reference, whose SSIM family is integer and
   forward-only, libs/pmbb-ivqm/src/xSSIM.cpp; for training and tuning callers that need
   gradients). Pictures are `planes` planar fp32 planes of width x height in device memory, values
   x in [0, 1] standing for samples x * max_value; the forward is the reference's SSIM map on those
   real-valued samples (10 x 10 window, C1 = (0.01 max)^2, C2 = (0.03 max)^2), so pictures holding
   k / max_value give the reference's values. Gradients are with respect to x, written to the
   non-null grad buffers (same shape as the inputs). Synchronous.
   --------------------------------------------------------------------------------------------- */
/* SSIM per plane: value[p] and d value[p] / d x of both pictures. */
qmiv_status qmiv_ssim_backward(qmiv_ctx* ctx, int32_t width, int32_t height, int32_t planes, const float* tst,
                               const float* ref, double max_value, int32_t calc_l, double* value, float* grad_tst,
                               float* grad_ref);
/* MS-SSIM per plane with a float 2 x 2 mean pyramid (the reference's integer pyramid rounds, so
   this equals it at the first scale only); scales (nullable) receives [5][planes]. */
qmiv_status qmiv_msssim_backward(qmiv_ctx* ctx, int32_t width, int32_t height, int32_t planes, const float* tst,
                                 const float* ref, double max_value, double* value, double* scales, float* grad_tst,
                                 float* grad_ref);
/* IV-SSIM of two 3-plane pictures (P0 = tst, P1 = ref): the shift search and GCD run on the
   pictures quantised to max_value = 2^bd - 1; the shift-compensated pictures gather the real
   samples. Gradient: a subgradient that holds the shifts, the GCD and the minimum's direction. */
qmiv_status qmiv_ivssim_backward(qmiv_ctx* ctx, int32_t width, int32_t height, const float* tst, const float* ref,
                                 double max_value, const qmiv_params* prm, double* ivssim, double* r2t, double* t2r,
                                 float* grad_tst, float* grad_ref);
/* qmiv_ivssim_backward over n picture pairs of one size (no counterpart in the reference; for
   training batches): pair i is (tst[i], ref[i]), its results go to ivssim[i], r2t[i], t2r[i] and the
   gradients to grad_tst[i] / grad_ref[i] (each array nullable, and so is each gradient pointer).
   The pairs' searches and SSIM evaluations run on the context's streams together, with one wait
   for all values and one for all gradients; each pair's results equal those of
   qmiv_ivssim_backward on it alone. */
qmiv_status qmiv_ivssim_backward_batch(qmiv_ctx* ctx, int32_t n, int32_t width, int32_t height,
                                       const float* const* tst, const float* const* ref, double max_value,
                                       const qmiv_params* prm, double* ivssim, double* r2t, double* t2r,
                                       float* const* grad_tst, float* const* grad_ref);

#ifdef __cplusplus
}
#endif
#endif
