/*
  qmiv.hpp - C++ processor classes over the C ABI in qmiv.h, with the method names, argument
  order and semantics of the reference's metric processors, so that a frame loop written
  against libs/pmbb-ivqm can switch to the MI355X engine by changing types, not calls.

  Header-only; link with -lqmiv. Failures throw qmiv::xError (the reference asserts instead).

    reference (namespace PMBB)                       here (namespace qmiv)
    xPicP                (libs/pmbb-core/src/xPic.h)   xPicP: planar 4:4:4 picture in device memory
    xPicI                (libs/pmbb-core/src/xPic.h)   xPicI: the interleaved picture's role (a device picture)
    xGlobClrDiffProc     (xGlobClrDiff.h:38-43)        xGlobClrDiffProc
    xShftCompPicProc     (xShftCompPic.h:33-44)        xShftCompPicProc
    xPSNR / xWSPSNR / xIVPSNR (xPSNR.h, xWSPSNR.h, xIVPSNR.h)   xIVPSNR (PSNR, WS-PSNR, IV-PSNR)
    xSSIM / xIVSSIM      (xSSIM.h:25-79)               xIVSSIM (SSIM, MS-SSIM, IV-SSIM)
    xSeq / xSeqPNG       (xSeq.h, xSeqPNG.h)           xSeq / xSeqPNG (read side)
*/
#ifndef QMIV_HPP
#define QMIV_HPP

#include <array>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "qmiv.h"

namespace qmiv {

using int32 = int32_t;
using flt32 = float;
using flt64 = double;
using int32V2 = std::array<int32, 2>;
using int32V4 = std::array<int32, 4>;
using flt32V4 = std::array<flt32, 4>;
using flt64V4 = std::array<flt64, 4>;

class xError : public std::runtime_error {
 public:
  xError(qmiv_status Status, const std::string& Msg) : std::runtime_error(Msg), m_Status(Status) {}
  qmiv_status status() const { return m_Status; }

 private:
  qmiv_status m_Status;
};

inline void xCheck(qmiv_status Status, const char* Where) {
  if (Status != QMIV_OK) throw xError(Status, std::string(Where) + ": " + qmiv_last_error());
}

// One HIP device and stream. Takes the place of the reference's thread pool (xThreadPool):
// every picture and processor below runs its work on one context.
class xContext {
 public:
  explicit xContext(int32 Device = 0) { xCheck(qmiv_create(Device, &m_Ctx), "qmiv_create"); }
  ~xContext() { qmiv_destroy(m_Ctx); }
  xContext(const xContext&) = delete;
  xContext& operator=(const xContext&) = delete;

  qmiv_ctx* get() const { return m_Ctx; }
  void waitUntilTasksFinished() { xCheck(qmiv_sync(m_Ctx), "qmiv_sync"); }

 private:
  qmiv_ctx* m_Ctx = nullptr;
};

// Picture components (eCmp, libs/pmbb-core/src/xCommonDefCORE.h): Y, Cb, Cr or R, G, B planes.
enum class eCmp : int32 { LM = 0, CB = 1, CR = 2, C0 = 0, C1 = 1, C2 = 2, R = 0, G = 1, B = 2 };
using boolV4 = std::array<bool, 4>;

// Planar 4:4:4 picture in device memory (the reference's xPicP, xPic.h:18-74). The device copy has
// no margin: the kernels clamp at the edges, which reads what ExtendMargin replicates
// (xPixelOpsSTD.cpp:210-236), so extend() has nothing to do. Margin and getStride() describe the
// reference's host layout (uint16 planes, rows W + 2 Margin apart) that fromPlanar / toPlanar
// exchange. Device samples are 8-bit at bit depth 8, else 16-bit.
class xPicP {
 public:
  xPicP(xContext& Ctx, int32V2 Size, int32 BitDepth, int32 Margin = 4) : m_Ctx(&Ctx) { create(Size, BitDepth, Margin); }
  ~xPicP() { destroy(); }
  xPicP(const xPicP&) = delete;
  xPicP& operator=(const xPicP&) = delete;

  void create(int32V2 Size, int32 BitDepth, int32 Margin = 4) {
    destroy();
    m_Size = Size;
    m_BitDepth = BitDepth;
    m_Margin = Margin;
    xCheck(qmiv_pic_create(m_Ctx->get(), Size[0], Size[1], BitDepth, 444, &m_Pic), "qmiv_pic_create");
  }
  void destroy() {
    qmiv_pic_destroy(m_Pic);
    m_Pic = nullptr;
  }

  int32V2 getSize() const { return m_Size; }
  int32 getWidth() const { return m_Size[0]; }
  int32 getHeight() const { return m_Size[1]; }
  int32 getArea() const { return m_Size[0] * m_Size[1]; }
  int32 getBitDepth() const { return m_BitDepth; }
  int32 getMaxPelValue() const { return (1 << m_BitDepth) - 1; }
  int32 getMargin() const { return m_Margin; }
  int32 getStride() const { return m_Size[0] + 2 * m_Margin; }
  int32 getNumCmps() const { return 3; }
  bool isCompatible(const xPicP* Ref) const { return Ref && Ref->m_Size == m_Size && Ref->m_BitDepth == m_BitDepth; }
  qmiv_pic* handle() const { return m_Pic; }
  qmiv_ctx* context() const { return m_Ctx->get(); }

  // xPicP::swapComponents (libs/pmbb-core/src/xPic.cpp:158-164): exchanges two planes (0..2); the
  // data moves on the device, the handle stays.
  bool swapComponents(int32 CmpIdA, int32 CmpIdB) {
    return qmiv_pic_swap_planes(m_Ctx->get(), m_Pic, CmpIdA, CmpIdB) == QMIV_OK;
  }
  bool swapComponents(eCmp CmpIdA, eCmp CmpIdB) { return swapComponents((int32)CmpIdA, (int32)CmpIdB); }

  // xPicP::copy (xPic.cpp:46-51).
  void copy(const xPicP* Src) { xCheck(qmiv_pic_copy(m_Ctx->get(), m_Pic, Src->m_Pic), "qmiv_pic_copy"); }
  // xPicP::extend (xPic.cpp:91-95): the device kernels clamp at the edges instead.
  void extend() {}
  // xPicP::check (xPic.cpp:62-82): false when a component holds a sample above 2^BitDepth - 1, after
  // printing, as the reference does, "FILE BROKEN <Name> (CMP=c)" for the first such component and
  // one "DETECTED out-of-range value" line per offending sample (xPixelOpsSTD::FindOutOfRange,
  // xPixelOpsSTD.cpp:175-197) to stdout.
  bool check(const std::string& Name) const {
    int32 Ok[3] = {1, 1, 1};
    xCheck(qmiv_pic_check(m_Ctx->get(), m_Pic, Ok), "qmiv_pic_check");
    if (Ok[0] && Ok[1] && Ok[2]) return true;
    std::vector<uint16_t> P = toPlanarBuffer();
    const int32 W = getWidth(), H = getHeight(), Max = getMaxPelValue();
    for (int32 c = 0; c < 3; c++) {
      if (Ok[c]) continue;
      std::printf("FILE BROKEN %s (CMP=%d)\n", Name.c_str(), c);
      const uint16_t* S = P.data() + (size_t)c * W * H;
      for (int32 y = 0; y < H; y++)
        for (int32 x = 0; x < W; x++)
          if (S[(size_t)y * W + x] > Max)
            std::printf("DETECTED out-of-range value (y=%d, x=%d, VALUE=%d, Expected=[0-%d])\n", y, x,
                        (int)S[(size_t)y * W + x], Max);
      return false;
    }
    return false;
  }
  // xPicP::conceal (xPic.cpp:83-90): samples above 2^BitDepth - 1 clipped to it, on the device.
  void conceal() { xCheck(qmiv_pic_conceal(m_Ctx->get(), m_Pic), "qmiv_pic_conceal"); }
  // xPicP::equalCmp / equalCmps (xPic.cpp:103-114).
  bool equalCmp(const xPicP* Src, eCmp CmpId) const { return equalCmps(Src)[(int32)CmpId]; }
  boolV4 equalCmps(const xPicP* Src) const {
    int32 E[3] = {0, 0, 0};
    xCheck(qmiv_pic_equal(m_Ctx->get(), m_Pic, Src->m_Pic, E), "qmiv_pic_equal");
    return {E[0] != 0, E[1] != 0, E[2] != 0, true};
  }
  bool equalPic(const xPicP* Src) const {
    const boolV4 E = equalCmps(Src);
    return E[0] && E[1] && E[2];
  }

  // The reference's host layout (xPicCommon.cpp:16-33): Origins[c] points at plane c's first
  // interior sample, rows Stride samples apart (getStride() for this picture's margin); margins are
  // neither read nor written.
  void fromPlanar(const uint16_t* const Origins[3], int32 Stride) {
    xCheck(qmiv_pic_upload_planar(m_Ctx->get(), m_Pic, Origins, Stride), "qmiv_pic_upload_planar");
  }
  void toPlanar(uint16_t* const Origins[3], int32 Stride) const {
    xCheck(qmiv_pic_download_planar(m_Ctx->get(), m_Pic, Origins, Stride), "qmiv_pic_download_planar");
  }

  // xSeqBase::readFrame's unpack (libs/pmbb-core/src/xSeq.cpp:179-249): one packed planar frame
  // as read from a raw file (ChromaFormat 400, 420, 422 or 444, any size) becomes this 4:4:4
  // picture; chroma is upsampled nearest-neighbour on the device.
  void fromPacked(const void* Packed, int32 ChromaFormat) {
    xCheck(qmiv_pic_upload_unpack(m_Ctx->get(), m_Pic, Packed,
                                  qmiv_frame_bytes(getWidth(), getHeight(), m_BitDepth, ChromaFormat), ChromaFormat),
           "qmiv_pic_upload_unpack");
  }
  // Planar 4:4:4 samples (Y, Cb, Cr; W x H each; 1 or 2 bytes per sample) in and out.
  void upload(const void* Planes) {
    xCheck(qmiv_pic_upload(m_Ctx->get(), m_Pic, Planes, qmiv_frame_bytes(getWidth(), getHeight(), m_BitDepth, 444)),
           "qmiv_pic_upload");
  }
  void download(void* Planes) const {
    xCheck(qmiv_pic_download(m_Ctx->get(), m_Pic, Planes, qmiv_frame_bytes(getWidth(), getHeight(), m_BitDepth, 444)),
           "qmiv_pic_download");
  }

 private:
  std::vector<uint16_t> toPlanarBuffer() const {
    const size_t A = (size_t)getArea();
    std::vector<uint16_t> P(3 * A);
    uint16_t* const O[3] = {P.data(), P.data() + A, P.data() + 2 * A};
    toPlanar(O, getWidth());
    return P;
  }

  xContext* m_Ctx;
  qmiv_pic* m_Pic = nullptr;
  int32V2 m_Size = {0, 0};
  int32 m_BitDepth = 0;
  int32 m_Margin = 4;
};

// The reference's interleaved picture (xPicI, xPic.h:79-113: 4 x uint16 per pel), the form its
// interleaved IV-PSNR processors take (xIVPSNR.h:35,49; the application's default,
// xAppQMIV.cpp:314,892). Interleaving is a CPU memory layout; on the device the picture is the
// same planar 4:4:4 picture, so rearrangeFromPlanar is a device copy. fromInterleaved /
// toInterleaved exchange the reference's host layout (origin at the first interior pel, rows Stride
// pels apart, A written 0).
class xPicI {
 public:
  xPicI(xContext& Ctx, int32V2 Size, int32 BitDepth, int32 Margin = 4)
      : m_Ctx(&Ctx), m_Size(Size), m_BitDepth(BitDepth), m_Margin(Margin) {
    xCheck(qmiv_pic_create(Ctx.get(), Size[0], Size[1], BitDepth, 444, &m_Pic), "qmiv_pic_create");
  }
  ~xPicI() { qmiv_pic_destroy(m_Pic); }
  xPicI(const xPicI&) = delete;
  xPicI& operator=(const xPicI&) = delete;

  int32V2 getSize() const { return m_Size; }
  int32 getWidth() const { return m_Size[0]; }
  int32 getHeight() const { return m_Size[1]; }
  int32 getBitDepth() const { return m_BitDepth; }
  int32 getMargin() const { return m_Margin; }
  int32 getStride() const { return m_Size[0] + 2 * m_Margin; }
  bool isCompatible(const xPicI* Ref) const { return Ref && Ref->m_Size == m_Size && Ref->m_BitDepth == m_BitDepth; }
  qmiv_pic* handle() const { return m_Pic; }

  // xPicI::rearrangeFromPlanar (xPic.cpp:201-207).
  void rearrangeFromPlanar(const xPicP* Src) { xCheck(qmiv_pic_copy(m_Ctx->get(), m_Pic, Src->handle()), "qmiv_pic_copy"); }
  void fromInterleaved(const uint16_t* Origin, int32 Stride) {
    xCheck(qmiv_pic_upload_interleaved(m_Ctx->get(), m_Pic, Origin, Stride), "qmiv_pic_upload_interleaved");
  }
  void toInterleaved(uint16_t* Origin, int32 Stride) const {
    xCheck(qmiv_pic_download_interleaved(m_Ctx->get(), m_Pic, Origin, Stride), "qmiv_pic_download_interleaved");
  }

 private:
  xContext* m_Ctx;
  qmiv_pic* m_Pic = nullptr;
  int32V2 m_Size;
  int32 m_BitDepth;
  int32 m_Margin;
};

// Read side of the reference's sequences: xSeqBase (libs/pmbb-core/src/xSeq.h:17-150,
// xSeq.cpp:28-174) with the raw backend xSeq and the PNG-list backend xSeqPNG
// (libs/pmbb-spng/src/xSeqPNG.h:14-38). Frames are read by a background thread into pinned buffers
// and copied to the device asynchronously on the picture's context (qmiv_seq_*); readFrame gives
// the 4:4:4 picture the reference's unpack gives (chroma of 4:2:0, 4:2:2 and 4:0:0 files unpacked
// on the device). Only eMode::Read is implemented.
class xSeqBase {
 public:
  enum class eMode : int32 { Unknown, Read, Write, Append };
  enum class eRetv : int32 { Success, EndOfFile, Error, WrongArg, NotImplemented };
  class tResult {
   public:
    tResult(eRetv Result, const std::string& Message = std::string()) : m_Result(Result), m_Message(Message) {}
    explicit operator bool() const { return m_Result == eRetv::Success; }
    bool operator==(eRetv Res) const { return m_Result == Res; }
    bool operator!=(eRetv Res) const { return m_Result != Res; }
    eRetv getResult() const { return m_Result; }
    const std::string& getMessage() const { return m_Message; }

   private:
    eRetv m_Result;
    std::string m_Message;
  };

  virtual ~xSeqBase() { (void)closeFile(); }
  xSeqBase(const xSeqBase&) = delete;
  xSeqBase& operator=(const xSeqBase&) = delete;

  tResult openFile(const std::string& FileName, eMode OpMode) {
    if (OpMode != eMode::Read) return eRetv::NotImplemented;
    (void)closeFile();
    const qmiv_status St = xBackendOpen(FileName);
    if (St != QMIV_OK) return tResult(eRetv::Error, qmiv_last_error());
    m_NumOfFrames = qmiv_seq_num_frames(m_Seq);
    m_CurrFrameIdx = 0;
    return eRetv::Success;
  }
  tResult closeFile() {
    qmiv_seq_close(m_Seq);
    m_Seq = nullptr;
    m_NumOfFrames = 0;
    m_CurrFrameIdx = 0;
    return eRetv::Success;
  }
  int32 getNumOfFrames() const { return m_NumOfFrames; }
  int32 getCurrFrameIdx() const { return m_CurrFrameIdx; }
  tResult seekFrame(int32 FrameNumber) {
    if (!m_Seq) return eRetv::Error;
    if (FrameNumber >= m_NumOfFrames) return eRetv::WrongArg;
    m_CurrFrameIdx = FrameNumber;
    return eRetv::Success;
  }
  tResult skipFrame(int32 NumFrames) { return seekFrame(m_CurrFrameIdx + NumFrames); }
  // The current frame into Pic (of the sequence's size and bit depth), then the next frame.
  tResult readFrame(xPicP* Pic) {
    if (!m_Seq) return tResult(eRetv::Error, "OpMode does not allow Read");
    if (m_CurrFrameIdx >= m_NumOfFrames) return eRetv::EndOfFile;
    if (!Pic || Pic->getSize() != m_Size || Pic->getBitDepth() != m_BitDepth) return eRetv::WrongArg;
    // The engine unpacks 4:2:0, 4:2:2 and 4:0:0 frames into the 4:4:4 picture on the device.
    if (qmiv_seq_read(Pic->context(), m_Seq, m_CurrFrameIdx, Pic->handle()) != QMIV_OK)
      return tResult(eRetv::Error, qmiv_last_error());
    m_CurrFrameIdx++;
    return eRetv::Success;
  }

 protected:
  xSeqBase(int32V2 Size, int32 BitDepth, int32 ChromaFormat) : m_Size(Size), m_BitDepth(BitDepth), m_ChromaFormat(ChromaFormat) {}
  virtual qmiv_status xBackendOpen(const std::string& FileName) = 0;

  qmiv_seq* m_Seq = nullptr;
  int32V2 m_Size;
  int32 m_BitDepth;
  int32 m_ChromaFormat;
  int32 m_NumOfFrames = 0;
  int32 m_CurrFrameIdx = 0;
};

// Raw planar file of packed frames (ChromaFormat 400, 420, 422 or 444), xSeq(Size, BitDepth, ChromaFormat).
class xSeq : public xSeqBase {
 public:
  xSeq(int32V2 Size, int32 BitDepth, int32 ChromaFormat) : xSeqBase(Size, BitDepth, ChromaFormat) {}

 protected:
  qmiv_status xBackendOpen(const std::string& FileName) override {
    return qmiv_seq_open(FileName.c_str(), m_Size[0], m_Size[1], m_BitDepth, m_ChromaFormat, &m_Seq);
  }
};

// List of 8-bit PNG files named by an fmt-style pattern ("frame_{:03d}.png"), read as R, G, B
// planes: xSeqPNG(Size, MaxNumFiles). The engine stops at 65535 files, the bound the reference
// application passes.
class xSeqPNG : public xSeqBase {
 public:
  explicit xSeqPNG(int32V2 Size, int32 /*MaxNumFiles*/ = 65535) : xSeqBase(Size, 8, 444) {}

 protected:
  qmiv_status xBackendOpen(const std::string& FileName) override {
    return qmiv_seq_open_png(FileName.c_str(), m_Size[0], m_Size[1], &m_Seq);
  }
};

// eClrSpcLC (libs/pmbb-core/src/xCommonDefCORE.h:356-364), the standardised YCbCr spaces.
enum class eClrSpcLC : int32 { BT601 = 0, SMPTE170M = 0, BT709 = 1, SMPTE240M = 2, BT2020 = 3 };

// xColorSpace::ConvertRGB2YCbCr / ConvertYCbCr2RGB over whole 4:4:4 pictures, as the reference
// app calls them (apps/QMIV/src/xAppQMIV.cpp:778-798): planes (R, G, B) <-> (Y, Cb, Cr), bit-exact
// Q16 integer arithmetic, Dst may be Src.
class xColorSpace {
 public:
  static void ConvertRGB2YCbCr(xPicP* Dst, const xPicP* Src, eClrSpcLC ClrSpc) {
    xCheck(qmiv_pic_colorspace(Src->context(), Src->handle(), Dst->handle(), (int32)ClrSpc, 0), "qmiv_pic_colorspace");
  }
  static void ConvertYCbCr2RGB(xPicP* Dst, const xPicP* Src, eClrSpcLC ClrSpc) {
    xCheck(qmiv_pic_colorspace(Src->context(), Src->handle(), Dst->handle(), (int32)ClrSpc, 1), "qmiv_pic_colorspace");
  }
};

// Parameters shared by the processors: xGlobClrDiffPrms (xGlobClrDiff.h:17-26),
// xCorrespPixelShiftPrms (xCorrespPixelShift.h:26-47) and xWeightedSpherically::initWS
// (xWeightedSpherically.h:21).
class xProcPrms {
 public:
  explicit xProcPrms(xContext& Ctx) : m_Ctx(&Ctx) { qmiv_params_default(&m_Prm); }
  void setUnntcbCoef(const flt32V4& UnntcbCoef) {
    for (int c = 0; c < 4; c++) m_Prm.unnoticeable_coef[c] = UnntcbCoef[c];
  }
  void setSearchRange(int32 SearchRange) { m_Prm.search_range = SearchRange; }
  void setCmpWeightsSearch(const int32V4& CmpWeightsSearch) {
    for (int c = 0; c < 4; c++) m_Prm.cmp_weights_search[c] = CmpWeightsSearch[c];
  }
  void setCmpWeightsAverage(const int32V4& CmpWeightsAverage) {
    for (int c = 0; c < 4; c++) m_Prm.cmp_weights_average[c] = CmpWeightsAverage[c];
  }
  // Width, height and bit depth come with each picture here; the arguments are accepted for
  // signature compatibility.
  void initWS(bool UseWS, int32 /*Width*/, int32 /*Height*/, int32 /*BitDepth*/, int32 LonRangeDeg = 360,
              int32 LatRangeDeg = 180) {
    m_Prm.erp = UseWS ? 1 : 0;
    m_Prm.lon_range_deg = LonRangeDeg;
    m_Prm.lat_range_deg = LatRangeDeg;
  }
  // Row buffers live in the engine's device scratch; kept for call compatibility.
  void initRowBuffers(int32 /*Height*/) {}
  const qmiv_params& params() const { return m_Prm; }

 protected:
  xContext* m_Ctx;
  qmiv_params m_Prm;
};

// xGlobClrDiffProc (libs/pmbb-ivqm/src/xGlobClrDiff.h:38-43).
class xGlobClrDiffProc : public xProcPrms {
 public:
  using xProcPrms::xProcPrms;
  // rint(sum(Ref - Tst) / Area), clipped to the unnoticeable range (the R2T colour difference).
  int32V4 CalcGlobalColorDiff(const xPicP* Tst, const xPicP* Ref) {
    int32V4 G{};
    xCheck(qmiv_calc_gcd(m_Ctx->get(), Tst->handle(), Ref->handle(), &m_Prm, G.data()), "qmiv_calc_gcd");
    return G;
  }
  // xGlobClrDiffProc::CalcGlobalColorDiffM (xGlobClrDiff.h:42, xGlobClrDiff.cpp:52-83):
  // rint(sum((Ref - Tst) * Msk) / (NumNonMasked * MaxMsk)), clipped.
  int32V4 CalcGlobalColorDiffM(const xPicP* Tst, const xPicP* Ref, const xPicP* Msk, int32 NumNonMasked) {
    int32V4 G{};
    xCheck(qmiv_calc_gcd_masked(m_Ctx->get(), Tst->handle(), Ref->handle(), Msk->handle(), &m_Prm, NumNonMasked,
                                G.data()),
           "qmiv_calc_gcd_masked");
    return G;
  }
};

// xPixelOps::CountNonZero over a mask picture's first plane (libs/pmbb-core/src/xPixelOpsSTD.cpp:273-284).
inline int32 countNonZero(xContext& Ctx, const xPicP* Msk) {
  int32 N = 0;
  xCheck(qmiv_count_nonzero(Ctx.get(), Msk->handle(), &N), "qmiv_count_nonzero");
  return N;
}

// xShftCompPicProc (libs/pmbb-ivqm/src/xShftCompPic.h:33-44).
class xShftCompPicProc : public xProcPrms {
 public:
  using xProcPrms::xProcPrms;
  void GenShftCompPics(xPicP* DstRef, xPicP* DstTst, const xPicP* SrcRef, const xPicP* SrcTst,
                       const int32V4& GlobalColorDiffRef2Tst) {
    xCheck(qmiv_gen_scp_pics(m_Ctx->get(), SrcRef->handle(), SrcTst->handle(), &m_Prm, GlobalColorDiffRef2Tst.data(),
                             DstRef->handle(), DstTst->handle()),
           "qmiv_gen_scp_pics");
  }
};

// xIVPSNR with its bases xWSPSNR and xPSNR (libs/pmbb-ivqm/src/xIVPSNR.h:15-43, xWSPSNR.h,
// xPSNR.h:15-50).
class xIVPSNR : public xProcPrms {
 public:
  using tDCfGCS = std::function<void(const int32V4&)>;
  using tDCfQAP = std::function<void(flt64, flt64)>;
  using xProcPrms::xProcPrms;

  void setDebugCallbackGCS(tDCfGCS Callback) { m_DebugCallbackGCS = std::move(Callback); }
  void setDebugCallbackQAP(tDCfQAP Callback) { m_DebugCallbackQAP = std::move(Callback); }
  // xPSNR::setFakeValsForExact: PSNR(SSD = 1) instead of flt64_max for identical components.
  void setFakeValsForExact(bool FakeValsForExact) { m_FakeValsForExact = FakeValsForExact; }

  flt64V4 calcPicPSNR(const xPicP* Tst, const xPicP* Ref) { return xPerCmp(Tst, Ref, QMIV_PSNR); }
  // Plain PSNR unless initWS enabled the spherical weights (xWSPSNR.cpp:17-27).
  flt64V4 calcPicWSPSNR(const xPicP* Tst, const xPicP* Ref) { return xPerCmp(Tst, Ref, QMIV_WSPSNR); }

  // Masked forms (xPSNR::calcPicPSNRM, xPSNR.cpp:42-71; xWSPSNR::calcPicWSPSNRM, xWSPSNR.cpp:61-97):
  // NumNonMasked < 0 (NOT_VALID) counts the mask's non-zero pels.
  flt64V4 calcPicPSNRM(const xPicP* Tst, const xPicP* Ref, const xPicP* Msk, int32 NumNonMasked = -1) {
    flt64V4 V{};
    const qmiv_params P = xMaskedPrm();
    xCheck(qmiv_calc_psnr_masked(m_Ctx->get(), Tst->handle(), Ref->handle(), Msk->handle(), &P, NumNonMasked,
                                 V.data()),
           "qmiv_calc_psnr_masked");
    xReportMsk(Msk, NumNonMasked);
    return {V[0], V[1], V[2], 0.0};
  }
  flt64V4 calcPicWSPSNRM(const xPicP* Tst, const xPicP* Ref, const xPicP* Msk, int32 NumNonMasked = -1) {
    flt64V4 V{};
    const qmiv_params P = xMaskedPrm();
    xCheck(qmiv_calc_wspsnr_masked(m_Ctx->get(), Tst->handle(), Ref->handle(), Msk->handle(), &P, NumNonMasked,
                                   V.data()),
           "qmiv_calc_wspsnr_masked");
    xReportMsk(Msk, NumNonMasked);
    return {V[0], V[1], V[2], 0.0};
  }
  // xIVPSNRM::calcPicIVPSNRM(Tst, Ref, Msk, NumNonMasked, GCD) (xIVPSNR.cpp:152-166).
  flt64 calcPicIVPSNRM(const xPicP* Tst, const xPicP* Ref, const xPicP* Msk, int32 NumNonMasked,
                       const int32V4& GlobalColorDiffRef2Tst) {
    return xIVM(Tst->handle(), Ref->handle(), Msk, NumNonMasked, GlobalColorDiffRef2Tst);
  }
  // xIVPSNRM::calcPicIVPSNRM(Tst, Ref, Msk, TstI, RefI) (xIVPSNR.cpp:138-150): counts the mask and
  // computes the masked colour difference itself.
  flt64 calcPicIVPSNRM(const xPicP* Tst, const xPicP* Ref, const xPicP* Msk) {
    const int32 NumNonMasked = countNonZero(*m_Ctx, Msk);
    if (m_DebugCallbackMSK) m_DebugCallbackMSK(NumNonMasked);
    int32V4 G{};
    const qmiv_params P = xMaskedPrm();
    xCheck(qmiv_calc_gcd_masked(m_Ctx->get(), Tst->handle(), Ref->handle(), Msk->handle(), &P, NumNonMasked,
                                G.data()),
           "qmiv_calc_gcd_masked");
    if (m_DebugCallbackGCS) m_DebugCallbackGCS(G);
    return calcPicIVPSNRM(Tst, Ref, Msk, NumNonMasked, G);
  }
  void setDebugCallbackMSK(std::function<void(int32)> Callback) { m_DebugCallbackMSK = std::move(Callback); }

  // xIVPSNR::calcPicIVPSNR(Tst, Ref) (xIVPSNR.cpp:18-31): the colour difference is
  // -CalcGlobalColorDiff(Tst, Ref) of the static form, i.e. minus the rounded sum(Tst - Ref) / Area.
  flt64 calcPicIVPSNR(const xPicP* Tst, const xPicP* Ref) {
    int32V4 G{};
    xCheck(qmiv_calc_gcd(m_Ctx->get(), Ref->handle(), Tst->handle(), &m_Prm, G.data()), "qmiv_calc_gcd");
    for (int32& g : G) g = -g;
    if (m_DebugCallbackGCS) m_DebugCallbackGCS(G);
    return calcPicIVPSNR(Tst, Ref, G);
  }
  // xIVPSNR::calcPicIVPSNR(Tst, Ref, GlobalColorDiffRef2Tst) (xIVPSNR.cpp:33-45).
  flt64 calcPicIVPSNR(const xPicP* Tst, const xPicP* Ref, const int32V4& GlobalColorDiffRef2Tst) {
    return xIV(Tst->handle(), Ref->handle(), GlobalColorDiffRef2Tst);
  }
  // The interleaved form the application calls (xIVPSNR.h:35, xIVPSNR.cpp:45-57; xAppQMIV.cpp:892).
  flt64 calcPicIVPSNR(const xPicI* Tst, const xPicI* Ref, const int32V4& GlobalColorDiffRef2Tst) {
    return xIV(Tst->handle(), Ref->handle(), GlobalColorDiffRef2Tst);
  }
  // xIVPSNRM::calcPicIVPSNRM(const xPicI* Tst, const xPicI* Ref, const xPicP* Msk, NNM, GCD)
  // (xIVPSNR.h:49, xIVPSNR.cpp:152-166), the application's masked call (xAppQMIV.cpp:889).
  flt64 calcPicIVPSNRM(const xPicI* Tst, const xPicI* Ref, const xPicP* Msk, int32 NumNonMasked,
                       const int32V4& GlobalColorDiffRef2Tst) {
    return xIVM(Tst->handle(), Ref->handle(), Msk, NumNonMasked, GlobalColorDiffRef2Tst);
  }

 private:
  flt64V4 xPerCmp(const xPicP* Tst, const xPicP* Ref, uint32_t Metric) {
    qmiv_params P = m_Prm;
    P.fake_psnr_for_exact = m_FakeValsForExact ? 1 : 0;
    P.invalid_pel_action = QMIV_PEL_SKIP;  // processors validate nothing; xPicP::check is the caller's
    qmiv_frame_result R;
    xCheck(qmiv_frame(m_Ctx->get(), Tst->handle(), Ref->handle(), &P, Metric, &R), "qmiv_frame");
    const double* V = Metric == QMIV_PSNR ? R.psnr : R.wspsnr;
    return {V[0], V[1], V[2], 0.0};
  }
  flt64 xIV(qmiv_pic* Tst, qmiv_pic* Ref, const int32V4& GlobalColorDiffRef2Tst) {
    flt64 IV = 0, R2T = 0, T2R = 0;
    xCheck(qmiv_calc_ivpsnr(m_Ctx->get(), Tst, Ref, &m_Prm, GlobalColorDiffRef2Tst.data(), &IV, &R2T, &T2R, nullptr),
           "qmiv_calc_ivpsnr");
    if (m_DebugCallbackQAP) m_DebugCallbackQAP(R2T, T2R);
    return IV;
  }
  flt64 xIVM(qmiv_pic* Tst, qmiv_pic* Ref, const xPicP* Msk, int32 NumNonMasked, const int32V4& GlobalColorDiffRef2Tst) {
    flt64 IV = 0, R2T = 0, T2R = 0;
    const qmiv_params P = xMaskedPrm();
    xCheck(qmiv_calc_ivpsnr_masked(m_Ctx->get(), Tst, Ref, Msk->handle(), &P, NumNonMasked,
                                   GlobalColorDiffRef2Tst.data(), &IV, &R2T, &T2R),
           "qmiv_calc_ivpsnr_masked");
    if (m_DebugCallbackQAP) m_DebugCallbackQAP(R2T, T2R);
    return IV;
  }
  qmiv_params xMaskedPrm() const {
    qmiv_params P = m_Prm;
    P.fake_psnr_for_exact = m_FakeValsForExact ? 1 : 0;
    return P;
  }
  void xReportMsk(const xPicP* Msk, int32 NumNonMasked) {
    if (m_DebugCallbackMSK) m_DebugCallbackMSK(NumNonMasked >= 0 ? NumNonMasked : countNonZero(*m_Ctx, Msk));
  }

  tDCfGCS m_DebugCallbackGCS;
  tDCfQAP m_DebugCallbackQAP;
  std::function<void(int32)> m_DebugCallbackMSK;
  bool m_FakeValsForExact = false;
};

// xIVSSIM with its base xSSIM (libs/pmbb-ivqm/src/xSSIM.h:25-79).
class xIVSSIM : public xProcPrms {
 public:
  using tDCfGCS = std::function<void(const int32V4&)>;
  using tDCfQAP = std::function<void(flt64, flt64)>;
  using xProcPrms::xProcPrms;

  // Pyramid and shift-compensated buffers live in the engine's device scratch; kept for call
  // compatibility (xSSIM::create, xIVSSIM::create).
  void create(int32V2 /*Size*/, int32 /*BitDepth*/, int32 /*Margin*/, bool /*EnableMS*/) {}
  void destroy() {}
  void setDebugCallbackGCS(tDCfGCS Callback) { m_DebugCallbackGCS = std::move(Callback); }
  void setDebugCallbackQAP(tDCfQAP Callback) { m_DebugCallbackQAP = std::move(Callback); }

  flt64V4 calcPicSSIM(const xPicP* Tst, const xPicP* Ref, bool CalcL = true) {
    flt64V4 V{};
    xCheck(qmiv_calc_ssim(m_Ctx->get(), Tst->handle(), Ref->handle(), CalcL ? 1 : 0, V.data()), "qmiv_calc_ssim");
    V[3] = 0.0;
    return V;
  }
  flt64V4 calcPicMSSSIM(const xPicP* Tst, const xPicP* Ref) {
    flt64V4 V{};
    xCheck(qmiv_calc_msssim(m_Ctx->get(), Tst->handle(), Ref->handle(), V.data(), nullptr), "qmiv_calc_msssim");
    V[3] = 0.0;
    return V;
  }
  // xIVSSIM::calcPicIVSSIM(Tst, Ref) (xSSIM.cpp:166-190): GCD, shift-compensated pictures, SSIMs.
  flt64 calcPicIVSSIM(const xPicP* Tst, const xPicP* Ref) {
    flt64 IV = 0, R2T = 0, T2R = 0;
    if (m_DebugCallbackGCS) {
      int32V4 G{};
      xCheck(qmiv_calc_gcd(m_Ctx->get(), Tst->handle(), Ref->handle(), &m_Prm, G.data()), "qmiv_calc_gcd");
      m_DebugCallbackGCS(G);
    }
    xCheck(qmiv_calc_ivssim(m_Ctx->get(), Tst->handle(), Ref->handle(), &m_Prm, &IV, &R2T, &T2R), "qmiv_calc_ivssim");
    if (m_DebugCallbackQAP) m_DebugCallbackQAP(R2T, T2R);
    return IV;
  }
  // xIVSSIM::calcPicIVSSIM(Tst, Ref, TstSCP, RefSCP) (xSSIM.cpp:191-212), the application's form.
  flt64 calcPicIVSSIM(const xPicP* Tst, const xPicP* Ref, const xPicP* TstSCP, const xPicP* RefSCP) {
    flt64 IV = 0, R2T = 0, T2R = 0;
    xCheck(qmiv_calc_ivssim_scp(m_Ctx->get(), Tst->handle(), Ref->handle(), TstSCP->handle(), RefSCP->handle(), &m_Prm,
                                &IV, &R2T, &T2R),
           "qmiv_calc_ivssim_scp");
    if (m_DebugCallbackQAP) m_DebugCallbackQAP(R2T, T2R);
    return IV;
  }

 private:
  tDCfGCS m_DebugCallbackGCS;
  tDCfQAP m_DebugCallbackQAP;
};

}  // namespace qmiv

#endif  // QMIV_HPP
