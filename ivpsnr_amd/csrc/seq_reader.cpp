// Sequence readers feeding device pictures: raw planar YUV files, the read path of xSeq
// (libs/pmbb-core/src/xSeq.cpp:408-428 frame size, :451-470 frame count, :519-525 seek by frame),
// and lists of PNG files, the read path of xSeqPNG (libs/pmbb-spng/src/xSeqPNG.cpp:55-148). Packed
// frames are read (and, for PNG, decoded) into pinned host buffers by a background thread (the
// frame after the last one requested is prefetched) and copied into the pictures asynchronously on
// the context's stream, so disk reads, decoding, host-to-device copies and the metric kernels
// overlap.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>

#include <vector>

#include "host_internal.h"
#include "png_reader.h"

using namespace qmiv;

namespace {

constexpr int kSlots = 3;

struct Slot {
  void* host = nullptr;            // pinned, one packed frame
  int32_t frame = -1;              // frame held (valid when ready)
  bool ready = false;
  bool reading = false;            // the worker is filling it
  hipEvent_t copied = nullptr;     // recorded after the last device copy out of this slot
  bool copy_pending = false;
  uint64_t last_use = 0;
};

}  // namespace

struct qmiv_seq {
  int fd = -1;                     // raw sequence file
  std::vector<std::string> png;    // PNG sequence: one file per frame
  int32_t w = 0, h = 0, bd = 0, cf = 0;
  size_t frame_bytes = 0;
  int32_t num_frames = 0;
  int32_t device = -1;             // device of the pinned slots' events (set on the first read)
  Slot slot[kSlots];
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool stop = false;
  int32_t want = -1;               // frame the worker should prefetch next
  uint64_t tick = 0;
};

namespace {

// One whole frame into dst: pread of a raw frame (EINTR and short reads retried), or the decoded
// PNG file of the frame.
bool read_frame(const qmiv_seq* s, int32_t index, void* dst, std::string* err = nullptr) {
  if (!s->png.empty()) {
    std::string e;
    const bool ok = png::decode_file_rgb8_planar(s->png[(size_t)index], s->w, s->h, (uint8_t*)dst, e);
    if (!ok && err) *err = e;
    return ok;
  }
  size_t done = 0;
  const off_t base = (off_t)s->frame_bytes * (off_t)index;
  while (done < s->frame_bytes) {
    const ssize_t r = pread(s->fd, (char*)dst + done, s->frame_bytes - done, base + (off_t)done);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) {
      if (err) *err = "short read";
      return false;
    }
    done += (size_t)r;
  }
  return true;
}

// A slot that holds neither `keep` nor a read in progress, least recently used first. Called with
// the lock held.
int pick_slot(qmiv_seq* s, int32_t keep) {
  int best = -1;
  for (int i = 0; i < kSlots; i++) {
    const Slot& sl = s->slot[i];
    if (sl.reading || (sl.ready && sl.frame == keep)) continue;
    if (best < 0 || sl.last_use < s->slot[best].last_use) best = i;
  }
  return best;
}

// Waits until the device copy out of slot i has finished (the slot may then be overwritten).
// Called without the lock.
hipError_t drain_copy(qmiv_seq* s, int i) {
  Slot& sl = s->slot[i];
  if (!sl.copy_pending) return hipSuccess;
  const hipError_t e = hipEventSynchronize(sl.copied);
  sl.copy_pending = false;
  return e;
}

void worker_main(qmiv_seq* s) {
  if (s->device >= 0) (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv.wait(lk, [&] { return s->stop || s->want >= 0; });
    if (s->stop) return;
    const int32_t f = s->want;
    s->want = -1;
    bool held = false;
    for (const Slot& sl : s->slot) held = held || ((sl.ready || sl.reading) && sl.frame == f);
    if (held) continue;
    const int i = pick_slot(s, f - 1);
    if (i < 0) continue;
    Slot& sl = s->slot[i];
    sl.reading = true;
    sl.ready = false;
    sl.frame = f;
    lk.unlock();
    const bool ok = drain_copy(s, i) == hipSuccess && read_frame(s, f, sl.host);
    lk.lock();
    sl.reading = false;
    sl.ready = ok;
    sl.last_use = ++s->tick;
    s->cv.notify_all();
  }
}

qmiv_status ensure_slots(qmiv_seq* s, qmiv_ctx* ctx) {
  if (s->device >= 0) {
    if (s->device != ctx->device) return host_fail(QMIV_ERR_INVALID_ARG, "sequence already bound to another device");
    return QMIV_OK;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return host_fail(QMIV_ERR_HIP, "hipSetDevice failed");
  for (Slot& sl : s->slot) {
    if (hipHostMalloc(&sl.host, s->frame_bytes, hipHostMallocDefault) != hipSuccess)
      return host_fail(QMIV_ERR_OUT_OF_MEMORY, "pinned frame buffer allocation failed");
    if (hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming) != hipSuccess)
      return host_fail(QMIV_ERR_HIP, "hipEventCreate failed");
  }
  s->device = ctx->device;
  s->worker = std::thread(worker_main, s);
  return QMIV_OK;
}

}  // namespace

extern "C" {

qmiv_status qmiv_seq_open(const char* path, int32_t width, int32_t height, int32_t bit_depth, int32_t chroma_format,
                          qmiv_seq** out) {
  if (!path || !out) return host_fail(QMIV_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  const size_t fb = qmiv_frame_bytes(width, height, bit_depth, chroma_format);
  if (fb == 0) return host_fail(QMIV_ERR_INVALID_ARG, "unsupported frame geometry");
  const int fd = open(path, O_RDONLY | O_CLOEXEC);
  if (fd < 0) return host_fail(QMIV_ERR_INVALID_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    return host_fail(QMIV_ERR_INVALID_ARG, std::string("cannot stat ") + path);
  }
  qmiv_seq* s = new qmiv_seq;
  s->fd = fd;
  s->w = width;
  s->h = height;
  s->bd = bit_depth;
  s->cf = chroma_format;
  s->frame_bytes = fb;
  // xSeq::xBackendOpen (xSeq.cpp:451-470): whole frames only, a partial last frame is ignored.
  const long long n = (long long)st.st_size / (long long)fb;
  s->num_frames = n > INT32_MAX ? INT32_MAX : (int32_t)n;
  *out = s;
  return QMIV_OK;
}

qmiv_status qmiv_seq_open_png(const char* pattern, int32_t width, int32_t height, qmiv_seq** out) {
  if (!pattern || !out) return host_fail(QMIV_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (width <= 0 || height <= 0 || (int64_t)width * height > INT32_MAX / 3)
    return host_fail(QMIV_ERR_INVALID_ARG, "unsupported frame geometry");
  // xSeqPNG::xBackendOpen (xSeqPNG.cpp:55-78): the first index is the last of 0 and 1 whose file
  // exists; frames are the files from there on, up to the first missing index or 65535.
  auto exists = [](const std::string& f) {
    struct stat st;
    return stat(f.c_str(), &st) == 0;
  };
  std::string name, err;
  int32_t first = -1;
  for (int32_t i = 0; i <= 1; i++) {
    if (!png::format_index(pattern, i, name, err)) return host_fail(QMIV_ERR_INVALID_ARG, err);
    if (exists(name)) first = i;
  }
  if (first < 0) return host_fail(QMIV_ERR_INVALID_ARG, std::string("First Idx=(0 or 1) not found for ") + pattern);
  std::vector<std::string> files;
  for (int32_t i = first; i < 65535; i++) {
    if (!png::format_index(pattern, i, name, err)) return host_fail(QMIV_ERR_INVALID_ARG, err);
    if (!exists(name)) break;
    files.push_back(name);
  }
  qmiv_seq* s = new qmiv_seq;
  s->png = std::move(files);
  s->w = width;
  s->h = height;
  s->bd = 8;
  s->cf = 444;
  s->frame_bytes = (size_t)3 * (size_t)width * (size_t)height;
  s->num_frames = (int32_t)s->png.size();
  *out = s;
  return QMIV_OK;
}

int32_t qmiv_seq_num_frames(const qmiv_seq* seq) { return seq ? seq->num_frames : -1; }

size_t qmiv_seq_frame_bytes(const qmiv_seq* seq) { return seq ? seq->frame_bytes : 0; }

qmiv_status qmiv_seq_read_host(qmiv_seq* seq, int32_t index, void* dst, size_t bytes) {
  if (!seq || !dst) return host_fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (index < 0 || index >= seq->num_frames) return host_fail(QMIV_ERR_INVALID_ARG, "frame index out of range");
  if (bytes != seq->frame_bytes) return host_fail(QMIV_ERR_INVALID_ARG, "buffer size differs from the frame size");
  std::string err;
  if (!read_frame(seq, index, dst, &err)) return host_fail(QMIV_ERR_INVALID_ARG, err);
  return QMIV_OK;
}

qmiv_status qmiv_seq_read(qmiv_ctx* ctx, qmiv_seq* seq, int32_t index, qmiv_pic* pic) {
  if (!ctx || !seq || !pic) return host_fail(QMIV_ERR_INVALID_ARG, "null argument");
  if (pic->ctx != ctx) return host_fail(QMIV_ERR_INVALID_ARG, "picture belongs to another context");
  if (index < 0 || index >= seq->num_frames) return host_fail(QMIV_ERR_INVALID_ARG, "frame index out of range");
  bool unpack = false;
  const qmiv_status rs = file_route(pic, seq->w, seq->h, seq->bd, seq->cf, unpack);
  if (rs != QMIV_OK) return rs;
  const qmiv_status es = ensure_slots(seq, ctx);
  if (es != QMIV_OK) return es;
  std::unique_lock<std::mutex> lk(seq->mu);
  int i = -1;
  for (;;) {
    i = -1;
    bool pending = false;
    for (int k = 0; k < kSlots; k++) {
      if (seq->slot[k].frame != index) continue;
      if (seq->slot[k].ready) i = k;
      if (seq->slot[k].reading) pending = true;
    }
    if (i >= 0 || !pending) break;
    seq->cv.wait(lk);  // the worker is reading this very frame
  }
  if (i < 0) {  // not prefetched: read it here
    i = pick_slot(seq, -1);
    if (i < 0) {
      seq->cv.wait(lk, [&] { return (i = pick_slot(seq, -1)) >= 0; });
    }
    Slot& sl = seq->slot[i];
    sl.reading = true;
    sl.ready = false;
    sl.frame = index;
    lk.unlock();
    std::string err = "device copy out of the frame buffer failed";
    const bool ok = drain_copy(seq, i) == hipSuccess && read_frame(seq, index, sl.host, &err);
    lk.lock();
    sl.reading = false;
    sl.ready = ok;
    seq->cv.notify_all();
    if (!ok) return host_fail(QMIV_ERR_INVALID_ARG, err);
  }
  Slot& sl = seq->slot[i];
  sl.last_use = ++seq->tick;
  if (hipSetDevice(ctx->device) != hipSuccess) return host_fail(QMIV_ERR_HIP, "hipSetDevice failed");
  // The pinned frame goes straight into the picture, or (another chroma format) through the
  // context's staging buffer and the device unpack into the 4:4:4 picture.
  const qmiv_status us = upload_file_frame(ctx, pic, sl.host, seq->frame_bytes, seq->cf);
  if (us != QMIV_OK) return us;
  const hipError_t e = hipEventRecord(sl.copied, ctx->stream);
  if (e != hipSuccess) return host_fail(QMIV_ERR_HIP, std::string("frame copy: ") + hipGetErrorString(e));
  sl.copy_pending = true;
  if (index + 1 < seq->num_frames) {
    seq->want = index + 1;
    seq->cv.notify_all();
  }
  return QMIV_OK;
}

void qmiv_seq_close(qmiv_seq* seq) {
  if (!seq) return;
  {
    std::lock_guard<std::mutex> lk(seq->mu);
    seq->stop = true;
  }
  seq->cv.notify_all();
  if (seq->worker.joinable()) seq->worker.join();
  for (Slot& sl : seq->slot) {
    if (sl.copy_pending) (void)hipEventSynchronize(sl.copied);
    if (sl.copied) (void)hipEventDestroy(sl.copied);
    if (sl.host) (void)hipHostFree(sl.host);
  }
  if (seq->fd >= 0) close(seq->fd);
  delete seq;
}

}  // extern "C"
