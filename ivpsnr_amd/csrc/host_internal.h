// Host-side state shared by the engine's host translation units: contexts and device pictures
// (opaque types of the C ABI), and the error reporting of the C entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/qmiv.h"
#include "qmiv_internal.h"

struct qmiv_pic;
struct qmiv_ctx;

namespace qmiv {
// Records msg as the calling thread's last error (qmiv_last_error) and returns st.
qmiv_status host_fail(qmiv_status st, const std::string& msg);
// Bytes of one packed raw-file frame (xSeq::create), 0 for an invalid format.
size_t file_frame_bytes(int32_t w, int32_t h, int32_t bd, int32_t cf);
// Whether frames of file chroma format fcf go into pic by a plain copy (same format) or by the
// device unpack into a 4:4:4 picture; QMIV_ERR_INVALID_ARG when neither.
qmiv_status file_route(const qmiv_pic* pic, int32_t w, int32_t h, int32_t bd, int32_t fcf, bool& unpack);
// Enqueues on the context's stream the upload of one packed raw-file frame (host memory) into pic:
// one copy, then the unpack when the formats differ (through the context's device staging buffer).
qmiv_status upload_file_frame(qmiv_ctx* ctx, qmiv_pic* pic, const void* host, size_t bytes, int32_t fcf);
// Pinned staging slots and host threads of the context's uploads from pageable memory.
struct Stager;
void destroy_stager(Stager* s);
}  // namespace qmiv

struct qmiv_pic {
  qmiv_ctx* ctx;
  qmiv::Geom g;
  void* dev;
  bool owned;
};

struct qmiv_ctx {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  // Results copies of a frame set run here, each after its frame's kernels (an event per frame),
  // so that the next frame's kernels do not wait for the copy engine.
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> copy_ev;
  // Recorded after each frame's results copy: the host finalises frame i while later frames run.
  std::vector<hipEvent_t> done_ev;
  // Spherical weights and correction per (height, latitude range): the host finalisation's cosines
  // are computed once per size.
  struct WsTable {
    int32_t h, lat;
    std::vector<double> w;
    double corr;
  };
  std::vector<WsTable> ws_cache;
  // A frame set deals its frames round-robin over `stream` and the aux streams (independent scratch),
  // so up to three frames are in flight: one frame's short kernels and the start and end of its
  // search overlap the others' searches. Three compute streams and the copy stream fit the device's
  // four hardware queues per process (a fourth compute stream shares a queue and measured slower).
  // Fork/join events order them with the rest.
  static constexpr int kAuxStreams = 2;
  hipStream_t aux_stream[kAuxStreams] = {nullptr, nullptr};
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // Differentiable SSIM family: float workspace and the quantised pictures of the shift search.
  void* d_fwork = nullptr;
  size_t fwork_bytes = 0;
  std::vector<qmiv_pic*> fq;  // quantised pictures of the float SSIM family's shift searches, by slot
  // Device scratch (grown on demand; one frame evaluation at a time per context).
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  // Pinned host block receiving the per-frame results.
  void* h_results = nullptr;
  size_t results_bytes = 0;
  // Device staging of raw-file frames that unpack into 4:4:4 pictures, and of reference-layout
  // (planar or interleaved uint16) uploads and downloads.
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  // Host-to-device copies from pageable memory go through pinned slots filled by host threads
  // (created on the first such copy).
  qmiv::Stager* stager = nullptr;
  // Shift-compensated pictures (IV-SSIM), 2 x [3][h][w] uint16.
  void* d_scp = nullptr;
  size_t scp_bytes = 0;
  // Allocation-free frame entry point (qmiv_frame_noalloc): the device's fixed scratch (shared by
  // every context on the device; null when its set-up failed, with the HIP error kept) and a pinned
  // results block of kFixedResultsBytes allocated with the context.
  void* d_fixed = nullptr;
  size_t fixed_bytes = 0;
  hipError_t fixed_err = hipSuccess;
  void* h_fixed = nullptr;
  int32_t num_cus = 256;
  // [4:4:4 kernel, 4:2:0 kernel, 4:2:0 16-bit kernel][u8, u16], queried on first use
  int32_t cps_blocks_per_cu[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  // The batch autotuner's choice per batch signature (the views' geometry, metrics and parameters):
  // branches, sweep cap and the replay time of the fastest graph instance of that shape. A later
  // batch of the same signature takes it without tuning again and checks its instance against it.
  struct BatchChoice {
    std::string sig;
    int32_t branches, sweep;
    double ms;
  };
  std::vector<BatchChoice> batch_tuned;
  // Kernel timing: event pairs recorded around launches since the last harvest.
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  std::vector<int> ev_kind;
  size_t ev_used = 0;
  qmiv_profile prof{};
};
