/*
  Reference harness: drives the reference's own metric processors (pmbb-ivqm on
  pmbb-core, compiled in place from the reference tree by oracle/Makefile) the
  way the reference application's frame loop does, and prints every
  intermediate the GPU engine is checked against as one JSON document.

  TEST INFRASTRUCTURE ONLY. This program is the checker used to produce golden
  fixtures and to time the reference CPU path; it is never linked into, loaded
  by or called from the product library.

  Frame flow follows apps/QMIV/src/xAppQMIV.cpp:661-751 (read, validate,
  preprocess, GCD, SCP, PSNR, WS-PSNR, IV-PSNR, SSIM, MS-SSIM, IV-SSIM) with the
  preprocessing of :771-830 and the fake-PSNR override of :837-844,864-871.
*/
#include "xSeq.h"
#include "xPic.h"
#include "xIVPSNR.h"
#include "xSSIM.h"
#include "xGlobClrDiff.h"
#include "xShftCompPic.h"
#include "xThreadPool.h"
#include "xMathUtils.h"
#include "xPixelOps.h"
#include "xColorSpaceSTD.h"
#include "xColorSpaceCoeff.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace PMBB_NAMESPACE;

namespace {

// Exposes the protected per-direction row distortions of the IV-PSNR processor
// (libs/pmbb-ivqm/src/xIVPSNR.cpp:100-133) and the WS-PSNR row SSDs
// (libs/pmbb-ivqm/src/xWSPSNR.cpp:101-126).
class ProbeIVPSNR : public xIVPSNRM {
 public:
  flt64 qualI(const xPicI* T, const xPicI* R, const int32V4& G, std::vector<uint64V4>& Rows) {
    flt64 Q = xCalcQualAsymmetricPic(T, R, G);
    Rows = m_RowDistsV4;
    return Q;
  }
  flt64 qualP(const xPicP* T, const xPicP* R, const int32V4& G, std::vector<uint64V4>& Rows) {
    flt64 Q = xCalcQualAsymmetricPic(T, R, G);
    Rows = m_RowDistsV4;
    return Q;
  }
  flt64 qualM(const xPicI* T, const xPicI* R, const xPicP* M, const int32V4& G, int32 NNM,
              std::vector<uint64V4>& Rows) {
    flt64 Q = xCalcQualAsymmetricPicM(T, R, M, G, NNM);
    Rows = m_RowDistsV4;
    return Q;
  }
  const std::vector<uint64>& wsRows(int C) const { return m_RowDistortions[C]; }
  const std::vector<flt64>& weights() const { return m_EquirectangularWeights; }
  flt64 correction() const { return m_DistortionCorrection; }
};

// Exposes the MS-SSIM per-scale scores (libs/pmbb-ivqm/src/xSSIM.cpp:67-87) and
// the per-row SSIM sums of the last component computed.
class ProbeSSIM : public xIVSSIM {
 public:
  void scales(const xPicP* T, const xPicP* R, flt64V4 Out[c_NumMultiScales]) {
    Out[0] = calcPicSSIM(T, R, false);
    for (int32 i = 1; i < c_NumMultiScales; i++) {
      const xPicP* ST = i == 1 ? T : m_SubPicTst[i - 1];
      const xPicP* SR = i == 1 ? R : m_SubPicRef[i - 1];
      xDownsamplePic(m_SubPicTst[i], ST);
      xDownsamplePic(m_SubPicRef[i], SR);
      Out[i] = calcPicSSIM(m_SubPicTst[i], m_SubPicRef[i], i != c_NumMultiScales - 1);
    }
  }
  const std::vector<flt64>& rowSums(int C) const { return m_RowSums[C]; }
};

struct Args {
  int W = 0, H = 0, BD = 8, CF = 420;
  std::string I0, I1, IM;  // IM: mask file (masked mode, xAppQMIV.cpp:319-337)
  int BDM = 8, CFM = 420;
  int NumFrames = 0;
  bool Erp = false;
  int Lon = 360, Lat = 180;
  int SR = 2;
  int32V4 CWS = {4, 1, 1, 0};
  int32V4 CWA = {4, 1, 1, 0};
  flt32V4 UC = {0.01f, 0.01f, 0.01f, 0.0f};
  int Nth = 0;
  bool Ilp = true;
  bool Rows = false;       // dump per-row arrays (small pictures)
  bool Pics = false;       // dump SCP / pyramid pictures (tiny pictures)
  bool Conceal = false;    // InvalidPelActn CNCL: conceal the pictures that fail the check
  std::string Dump;        // write frame 0's unpacked pictures (P0, P1, mask: 3 planes W x H uint16 each)
  bool Time = false;       // per-stage timing
  bool M[6] = {true, true, true, false, false, false};  // PSNR WSPSNR IVPSNR SSIM MSSSIM IVSSIM
};

const char* kMetricNames[6] = {"PSNR", "WSPSNR", "IVPSNR", "SSIM", "MSSSIM", "IVSSIM"};

bool parseInt4(const char* S, int32V4& V) {
  int a, b, c, d;
  if (sscanf(S, "%d:%d:%d:%d", &a, &b, &c, &d) != 4) return false;
  V = {a, b, c, d};
  return true;
}
bool parseFlt4(const char* S, flt32V4& V) {
  float a, b, c, d;
  if (sscanf(S, "%f:%f:%f:%f", &a, &b, &c, &d) != 4) return false;
  V = {a, b, c, d};
  return true;
}

void usage() {
  fprintf(stderr,
          "qmiv_ref -w W -h H -bd BD -cf 420|444 -i0 FILE -i1 FILE [-im MASK -bdm BD -cfm CF] [-n FRAMES] [-erp LON LAT]\n"
          "         [-sr B] [-cws a:b:c:d] [-cwa a:b:c:d] [-ucf a:b:c:d] [-nth T] [-ilp 0|1]\n"
          "         [-ml PSNR,WSPSNR,IVPSNR,SSIM,MSSSIM,IVSSIM|All] [-rows] [-pics] [-time]\n"
          "         [-ipa STOP|WARN|SKIP|CNCL] [-dump FILE]\n");
}

bool parse(int argc, char** argv, Args& A) {
  for (int i = 1; i < argc; i++) {
    std::string K = argv[i];
    auto need = [&](int n) { if (i + n >= argc) { usage(); exit(2); } };
    if (K == "-w") { need(1); A.W = atoi(argv[++i]); }
    else if (K == "-h") { need(1); A.H = atoi(argv[++i]); }
    else if (K == "-bd") { need(1); A.BD = atoi(argv[++i]); }
    else if (K == "-cf") { need(1); A.CF = atoi(argv[++i]); }
    else if (K == "-i0") { need(1); A.I0 = argv[++i]; }
    else if (K == "-i1") { need(1); A.I1 = argv[++i]; }
    else if (K == "-im") { need(1); A.IM = argv[++i]; }
    else if (K == "-bdm") { need(1); A.BDM = atoi(argv[++i]); }
    else if (K == "-cfm") { need(1); A.CFM = atoi(argv[++i]); }
    else if (K == "-n") { need(1); A.NumFrames = atoi(argv[++i]); }
    else if (K == "-erp") { need(2); A.Erp = true; A.Lon = atoi(argv[++i]); A.Lat = atoi(argv[++i]); }
    else if (K == "-sr") { need(1); A.SR = atoi(argv[++i]); }
    else if (K == "-cws") { need(1); if (!parseInt4(argv[++i], A.CWS)) return false; }
    else if (K == "-cwa") { need(1); if (!parseInt4(argv[++i], A.CWA)) return false; }
    else if (K == "-ucf") { need(1); if (!parseFlt4(argv[++i], A.UC)) return false; }
    else if (K == "-nth") { need(1); A.Nth = atoi(argv[++i]); }
    else if (K == "-ilp") { need(1); A.Ilp = atoi(argv[++i]) != 0; }
    else if (K == "-rows") { A.Rows = true; }
    else if (K == "-pics") { A.Pics = true; }
    else if (K == "-time") { A.Time = true; }
    else if (K == "-ipa") { need(1); A.Conceal = std::string(argv[++i]) == "CNCL"; }
    else if (K == "-dump") { need(1); A.Dump = argv[++i]; }
    else if (K == "-ml") {
      need(1);
      std::string L = argv[++i];
      for (int m = 0; m < 6; m++) A.M[m] = false;
      if (L == "All") { for (int m = 0; m < 6; m++) A.M[m] = true; }
      else {
        size_t p = 0;
        while (p <= L.size()) {
          size_t q = L.find(',', p);
          if (q == std::string::npos) q = L.size();
          std::string T = L.substr(p, q - p);
          bool Found = false;
          for (int m = 0; m < 6; m++) if (T == kMetricNames[m]) { A.M[m] = true; Found = true; }
          if (!Found) { fprintf(stderr, "unknown metric %s\n", T.c_str()); return false; }
          p = q + 1;
        }
      }
    } else { usage(); return false; }
  }
  return A.W > 0 && A.H > 0 && !A.I0.empty() && !A.I1.empty();
}

void pd(FILE* F, double V) {
  if (V != V) fputs("NaN", F);
  else if (V == __builtin_inf()) fputs("Infinity", F);
  else if (V == -__builtin_inf()) fputs("-Infinity", F);
  else fprintf(F, "%.17g", V);
}
void pd4(FILE* F, const flt64V4& V, int N = 3) {
  fputc('[', F);
  for (int c = 0; c < N; c++) { if (c) fputc(',', F); pd(F, V[c]); }
  fputc(']', F);
}

void dumpPlaneHex(FILE* F, const xPicP* P, int C) {
  // Pels as one lowercase hex string of little-endian uint16 (no margin).
  fputc('"', F);
  for (int y = 0; y < P->getHeight(); y++) {
    const uint16* Row = P->getAddr((eCmp)C) + y * P->getStride();
    for (int x = 0; x < P->getWidth(); x++) fprintf(F, "%02x%02x", Row[x] & 0xff, Row[x] >> 8);
  }
  fputc('"', F);
}

using Clock = std::chrono::steady_clock;
double ms(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

// -cstables: the reference's Q16 integer colour matrices, [space][rgb2ycbcr | ycbcr2rgb][3][3].
// -cs DIR SPACE W H BD IN OUT: DIR 0 = RGB -> YCbCr, 1 = YCbCr -> RGB on three planar uint16
// planes (xColorSpaceSTD::ConvertRGB2YCbCr_I32 / ConvertYCbCr2RGB_I32).
static int colour_modes(int argc, char** argv) {
  const std::string M = argv[1];
  if (M == "-cstables") {
    fputc('[', stdout);
    for (int s = 0; s < 4; s++) {
      fprintf(stdout, "%s[", s ? "," : "");
      for (int d = 0; d < 2; d++) {
        const auto& T = d == 0 ? xColorSpaceCoeff<int32>::c_RGB2YCbCr[s] : xColorSpaceCoeff<int32>::c_YCbCr2RGB[s];
        fprintf(stdout, "%s[", d ? "," : "");
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) fprintf(stdout, "%s%d", (r || c) ? "," : "", (int)T[r][c]);
        fputc(']', stdout);
      }
      fputc(']', stdout);
    }
    fputs("]\n", stdout);
    return 0;
  }
  if (M == "-cs" && argc == 9) {
    const int Dir = atoi(argv[2]), Space = atoi(argv[3]), W = atoi(argv[4]), H = atoi(argv[5]), BD = atoi(argv[6]);
    const size_t N = (size_t)W * H;
    std::vector<uint16> In(3 * N), Out(3 * N);
    FILE* F = fopen(argv[7], "rb");
    if (!F || fread(In.data(), 2, 3 * N, F) != 3 * N) { fprintf(stderr, "cannot read %s\n", argv[7]); return 1; }
    fclose(F);
    if (Dir == 0)
      xColorSpaceSTD::ConvertRGB2YCbCr_I32(Out.data(), Out.data() + N, Out.data() + 2 * N, In.data(), In.data() + N,
                                           In.data() + 2 * N, W, W, W, H, BD, (eClrSpcLC)Space);
    else
      xColorSpaceSTD::ConvertYCbCr2RGB_I32(Out.data(), Out.data() + N, Out.data() + 2 * N, In.data(), In.data() + N,
                                           In.data() + 2 * N, W, W, W, H, BD, (eClrSpcLC)Space);
    F = fopen(argv[8], "wb");
    if (!F || fwrite(Out.data(), 2, 3 * N, F) != 3 * N) { fprintf(stderr, "cannot write %s\n", argv[8]); return 1; }
    fclose(F);
    return 0;
  }
  return -1;
}

int main(int argc, char** argv) {
  if (argc >= 2 && (std::string(argv[1]) == "-cstables" || std::string(argv[1]) == "-cs")) {
    const int r = colour_modes(argc, argv);
    if (r >= 0) return r;
    usage();
    return 2;
  }
  if (argc == 2 && std::string(argv[1]) == "-filter") {
    // The SSIM window table as the reference stores it (xStructSimConsts.h:30-43), fp32 bit patterns.
    fputc('[', stdout);
    for (int i = 0; i < 11; i++)
      for (int j = 0; j < 11; j++) {
        uint32_t bits;
        const float f = xStructSimConsts::c_FilterF[i][j];
        memcpy(&bits, &f, 4);
        fprintf(stdout, "%s%u", (i || j) ? "," : "", bits);
      }
    fputs("]\n", stdout);
    return 0;
  }
  Args A;
  if (!parse(argc, argv, A)) { usage(); return 2; }
  const eCrF CF = A.CF == 420 ? eCrF::CF420 : A.CF == 444 ? eCrF::CF444 : A.CF == 422 ? eCrF::CF422 : eCrF::CF400;
  const int32V2 Size(A.W, A.H);
  const bool CalcIVs = A.M[2] || A.M[5];
  const bool CalcSCP = A.M[5];
  const bool CalcSSIMs = A.M[3] || A.M[4] || A.M[5];
  const bool CalcPSNRs = A.M[0] || A.M[1] || A.M[2];
  const int32 Margin = xRoundUpToNearestMultiple(A.SR, 2);  // xAppQMIV.cpp:330

  const bool UseMask = !A.IM.empty();
  if (UseMask && CalcSSIMs) { fprintf(stderr, "SSIM metrics cannot be combined with mask mode\n"); return 2; }
  const eCrF CFM = A.CFM == 420 ? eCrF::CF420 : A.CFM == 444 ? eCrF::CF444 : A.CFM == 422 ? eCrF::CF422 : eCrF::CF400;
  xSeq Seq0(Size, A.BD, CF), Seq1(Size, A.BD, CF), SeqM(Size, A.BDM, CFM);
  if (!Seq0.openFile(A.I0, xSeq::eMode::Read) || !Seq1.openFile(A.I1, xSeq::eMode::Read) ||
      (UseMask && !SeqM.openFile(A.IM, xSeq::eMode::Read))) {
    fprintf(stderr, "cannot open inputs\n");
    return 1;
  }
  int NumFrames = std::min(Seq0.getNumOfFrames(), Seq1.getNumOfFrames());
  if (UseMask) NumFrames = std::min(NumFrames, SeqM.getNumOfFrames());
  if (A.NumFrames > 0) NumFrames = std::min(NumFrames, A.NumFrames);
  if (NumFrames <= 0) { fprintf(stderr, "no frames\n"); return 1; }

  xThreadPool* Pool = nullptr;
  if (A.Nth > 0) { Pool = new xThreadPool; Pool->create(A.Nth, A.H + 1); }

  xPicP P0(Size, A.BD, Margin), P1(Size, A.BD, Margin), PM;
  if (UseMask) PM.create(Size, A.BDM, Margin);
  xPicI I0, I1;
  // Masked IV-PSNR exists only for interleaved pictures (xIVPSNR.h:47-49).
  const bool UsePicI = (A.Ilp || UseMask) && A.M[2];
  if (UsePicI) { I0.create(Size, A.BD, Margin); I1.create(Size, A.BD, Margin); }
  xPicP SCP0, SCP1;
  if (CalcIVs) { SCP0.create(Size, A.BD, Margin); SCP1.create(Size, A.BD, Margin); }

  // Processor setup as in xAppQMIV::createProcessors (xAppQMIV.cpp:593-650).
  xGlobClrDiffProc ProcGCD;
  ProcGCD.setUnntcbCoef(A.UC);
  if (Pool) ProcGCD.initThreadPool(Pool, A.H + 1);
  xShftCompPicProc ProcSCP;
  ProcSCP.setSearchRange(A.SR);
  ProcSCP.setCmpWeightsSearch(A.CWS);
  ProcSCP.setCmpWeightsAverage(A.CWA);
  if (Pool) ProcSCP.initThreadPool(Pool, A.H + 1);
  ProbeIVPSNR ProcPSNR;
  ProcPSNR.setSearchRange(A.SR);
  ProcPSNR.setCmpWeightsSearch(A.CWS);
  ProcPSNR.setCmpWeightsAverage(A.CWA);
  ProcPSNR.setUnntcbCoef(A.UC);
  if (Pool) ProcPSNR.initThreadPool(Pool, A.H + 1);
  ProcPSNR.initRowBuffers(A.H);
  if (A.Erp) ProcPSNR.initWS(true, A.W, A.H, A.BD, A.Lon, A.Lat);
  ProbeSSIM ProcSSIM;
  if (CalcSSIMs) {
    ProcSSIM.create(Size, A.BD, Margin, true);
    ProcSSIM.setSearchRange(A.SR);
    ProcSSIM.setCmpWeightsSearch(A.CWS);
    ProcSSIM.setCmpWeightsAverage(A.CWA);
    ProcSSIM.setUnntcbCoef(A.UC);
    if (Pool) ProcSSIM.initThreadPool(Pool, A.H + 1);
    ProcSSIM.initRowBuffers(A.H);
  }
  flt64 LastR2T = 0, LastT2R = 0;
  ProcSSIM.setDebugCallbackQAP([&](flt64 R2T, flt64 T2R) { LastR2T = R2T; LastT2R = T2R; });

  const int32 Area = A.W * A.H;
  const flt64 FakePSNR = xPSNR::getFakePSNR(Area, A.BD);
  const flt64 InvW = 1.0 / (flt64)A.CWA.getSum();

  std::vector<flt64> PicVals[6];
  std::vector<flt64V4> CmpVals[6];
  double T[12] = {0};  // load validate preproc gcd scp psnr wspsnr ivpsnr ssim msssim ivssim

  FILE* F = stdout;
  fprintf(F, "{\"width\":%d,\"height\":%d,\"bit_depth\":%d,\"chroma_format\":%d,\"frames\":[", A.W, A.H, A.BD, A.CF);
  for (int f = 0; f < NumFrames; f++) {
    auto t0 = Clock::now();
    if (!Seq0.readFrame(&P0) || !Seq1.readFrame(&P1) || (UseMask && !SeqM.readFrame(&PM))) {
      fprintf(stderr, "read error\n");
      return 1;
    }
    auto t1 = Clock::now();
    bool InRange0 = P0.check(A.I0), InRange1 = P1.check(A.I1), InRangeM = UseMask ? PM.check(A.IM) : true;
    // xAppQMIV::validateFrames (xAppQMIV.cpp:753-770): CNCL conceals; WARN / SKIP go on as read.
    if (A.Conceal) {
      if (!InRange0) P0.conceal();
      if (!InRange1) P1.conceal();
      if (UseMask && !InRangeM) PM.conceal();
    }
    if (f == 0 && !A.Dump.empty()) {
      FILE* D = fopen(A.Dump.c_str(), "wb");
      if (!D) { fprintf(stderr, "cannot write %s\n", A.Dump.c_str()); return 1; }
      for (xPicP* Pic : {&P0, &P1, &PM}) {
        if (Pic == &PM && !UseMask) break;
        for (int c = 0; c < 3; c++)
          for (int y = 0; y < A.H; y++) fwrite(Pic->getAddr((eCmp)c) + y * Pic->getStride(), 2, A.W, D);
      }
      fclose(D);
    }
    auto t2 = Clock::now();
    boolV4 Exact = xMakeVec4<bool>(false);
    for (int c = 0; c < 3; c++) Exact[c] = P0.equalCmp(&P1, (eCmp)c);
    P0.extend();
    P1.extend();
    if (UseMask) PM.extend();
    if (UsePicI) { I0.rearrangeFromPlanar(&P0); I1.rearrangeFromPlanar(&P1); }
    // xAppQMIV.cpp:825-829
    const int32 NNM = UseMask ? xPixelOps::CountNonZero(PM.getAddr(eCmp::LM), PM.getStride(), PM.getWidth(), PM.getHeight()) : 0;
    auto t3 = Clock::now();
    int32V4 GCD = {0, 0, 0, 0};
    if (CalcIVs) GCD = UseMask ? ProcGCD.CalcGlobalColorDiffM(&P0, &P1, &PM, NNM) : ProcGCD.CalcGlobalColorDiff(&P0, &P1);
    auto t4 = Clock::now();
    if (CalcSCP) ProcSCP.GenShftCompPics(&SCP1, &SCP0, &P1, &P0, GCD);
    auto t5 = Clock::now();
    T[0] += ms(t0, t1); T[1] += ms(t1, t2); T[2] += ms(t2, t3); T[3] += ms(t3, t4); T[4] += ms(t4, t5);

    if (f) fputc(',', F);
    fprintf(F, "{\"frame\":%d,\"in_range\":[%d,%d],\"exact\":[%d,%d,%d],\"gcd\":[%d,%d,%d,%d]", f, InRange0, InRange1,
            (int)Exact[0], (int)Exact[1], (int)Exact[2], GCD[0], GCD[1], GCD[2], GCD[3]);
    if (UseMask) fprintf(F, ",\"in_range_mask\":%d,\"num_non_masked\":%d", InRangeM, NNM);

    if (A.M[0]) {
      auto a = Clock::now();
      flt64V4 V = UseMask ? ProcPSNR.calcPicPSNRM(&P0, &P1, &PM, NNM) : ProcPSNR.calcPicPSNR(&P0, &P1);
      for (int c = 0; c < 3; c++) if (Exact[c]) V[c] = FakePSNR;
      T[5] += ms(a, Clock::now());
      CmpVals[0].push_back(V);
      PicVals[0].push_back((V[0] * A.CWA[0] + V[1] * A.CWA[1] + V[2] * A.CWA[2]) * InvW);  // xAppQMIV.h:245-251
      fprintf(F, ",\"psnr\":"); pd4(F, V); fprintf(F, ",\"psnr_pic\":"); pd(F, PicVals[0].back());
    }
    if (A.M[1]) {
      auto a = Clock::now();
      flt64V4 V = UseMask ? ProcPSNR.calcPicWSPSNRM(&P0, &P1, &PM, NNM) : ProcPSNR.calcPicWSPSNR(&P0, &P1);
      for (int c = 0; c < 3; c++) if (Exact[c]) V[c] = FakePSNR;
      T[6] += ms(a, Clock::now());
      CmpVals[1].push_back(V);
      PicVals[1].push_back((V[0] * A.CWA[0] + V[1] * A.CWA[1] + V[2] * A.CWA[2]) * InvW);
      fprintf(F, ",\"wspsnr\":"); pd4(F, V); fprintf(F, ",\"wspsnr_pic\":"); pd(F, PicVals[1].back());
      if (A.Rows && A.Erp) {
        fprintf(F, ",\"ws_rows\":[");
        for (int c = 0; c < 3; c++) {
          const auto& R = ProcPSNR.wsRows(c);
          fprintf(F, "%s[", c ? "," : "");
          for (int y = 0; y < A.H; y++) fprintf(F, "%s%llu", y ? "," : "", (unsigned long long)R[y]);
          fputc(']', F);
        }
        fprintf(F, "],\"ws_weights\":[");
        for (int y = 0; y < A.H; y++) { if (y) fputc(',', F); pd(F, ProcPSNR.weights()[y]); }
        fprintf(F, "],\"ws_correction\":"); pd(F, ProcPSNR.correction());
      }
    }
    if (A.M[2]) {
      auto a = Clock::now();
      flt64 MR2T = 0, MT2R = 0;
      if (UseMask) ProcPSNR.setDebugCallbackQAP([&](flt64 R2T, flt64 T2R) { MR2T = R2T; MT2R = T2R; });
      flt64 V = UseMask   ? ProcPSNR.calcPicIVPSNRM(&I0, &I1, &PM, NNM, GCD)
                : UsePicI ? ProcPSNR.calcPicIVPSNR(&I0, &I1, GCD)
                          : ProcPSNR.calcPicIVPSNR(&P0, &P1, GCD);
      T[7] += ms(a, Clock::now());
      PicVals[2].push_back(V);
      fprintf(F, ",\"ivpsnr\":"); pd(F, V);
      if (UseMask) {
        // The masked processor's own labels (xIVPSNR.cpp:152-165): R2T searches P1 - g in P0.
        fprintf(F, ",\"r2t\":"); pd(F, MR2T); fprintf(F, ",\"t2r\":"); pd(F, MT2R);
        if (A.Rows) {
          std::vector<uint64V4> RowsR2T, RowsT2R;
          ProcPSNR.qualM(&I1, &I0, &PM, -GCD, NNM, RowsR2T);
          ProcPSNR.qualM(&I0, &I1, &PM, GCD, NNM, RowsT2R);
          const std::vector<uint64V4>* Dirs[2] = {&RowsR2T, &RowsT2R};
          fprintf(F, ",\"cps_rows\":[");
          for (int d = 0; d < 2; d++) {
            fprintf(F, "%s[", d ? "," : "");
            for (int c = 0; c < 3; c++) {
              fprintf(F, "%s[", c ? "," : "");
              for (int y = 0; y < A.H; y++) fprintf(F, "%s%llu", y ? "," : "", (unsigned long long)(*Dirs[d])[y][c]);
              fputc(']', F);
            }
            fputc(']', F);
          }
          fputc(']', F);
        }
      } else if (A.Rows) {
        // Both directions as calcPicIVPSNR forms them (xIVPSNR.cpp:45-57).
        std::vector<uint64V4> RowsR2T, RowsT2R;
        flt64 R2T, T2R;
        if (UsePicI) { R2T = ProcPSNR.qualI(&I0, &I1, GCD, RowsR2T); T2R = ProcPSNR.qualI(&I1, &I0, -GCD, RowsT2R); }
        else { R2T = ProcPSNR.qualP(&P0, &P1, GCD, RowsR2T); T2R = ProcPSNR.qualP(&P1, &P0, -GCD, RowsT2R); }
        fprintf(F, ",\"r2t\":"); pd(F, R2T); fprintf(F, ",\"t2r\":"); pd(F, T2R);
        const std::vector<uint64V4>* Dirs[2] = {&RowsR2T, &RowsT2R};
        fprintf(F, ",\"cps_rows\":[");
        for (int d = 0; d < 2; d++) {
          fprintf(F, "%s[", d ? "," : "");
          for (int c = 0; c < 3; c++) {
            fprintf(F, "%s[", c ? "," : "");
            for (int y = 0; y < A.H; y++) fprintf(F, "%s%llu", y ? "," : "", (unsigned long long)(*Dirs[d])[y][c]);
            fputc(']', F);
          }
          fputc(']', F);
        }
        fputc(']', F);
      }
    }
    if (A.M[3]) {
      auto a = Clock::now();
      flt64V4 V = ProcSSIM.calcPicSSIM(&P0, &P1);
      T[8] += ms(a, Clock::now());
      CmpVals[3].push_back(V);
      PicVals[3].push_back((V[0] * A.CWA[0] + V[1] * A.CWA[1] + V[2] * A.CWA[2]) * InvW);
      fprintf(F, ",\"ssim\":"); pd4(F, V); fprintf(F, ",\"ssim_pic\":"); pd(F, PicVals[3].back());
      if (A.Rows) {
        // Per-row sums of the last component (Cr); rows outside the interior are zero.
        fprintf(F, ",\"ssim_rows_cr\":[");
        for (int y = 0; y < A.H; y++) { if (y) fputc(',', F); pd(F, ProcSSIM.rowSums(2)[y]); }
        fputc(']', F);
      }
    }
    if (A.M[4]) {
      auto a = Clock::now();
      flt64V4 V = ProcSSIM.calcPicMSSSIM(&P0, &P1);
      T[9] += ms(a, Clock::now());
      CmpVals[4].push_back(V);
      PicVals[4].push_back((V[0] * A.CWA[0] + V[1] * A.CWA[1] + V[2] * A.CWA[2]) * InvW);
      fprintf(F, ",\"msssim\":"); pd4(F, V); fprintf(F, ",\"msssim_pic\":"); pd(F, PicVals[4].back());
      if (A.Rows) {
        flt64V4 S[5];
        ProcSSIM.scales(&P0, &P1, S);
        fprintf(F, ",\"msssim_scales\":[");
        for (int i = 0; i < 5; i++) { if (i) fputc(',', F); pd4(F, S[i]); }
        fputc(']', F);
      }
    }
    if (A.M[5]) {
      auto a = Clock::now();
      flt64 V = ProcSSIM.calcPicIVSSIM(&P0, &P1, &SCP0, &SCP1);
      T[10] += ms(a, Clock::now());
      PicVals[5].push_back(V);
      fprintf(F, ",\"ivssim\":"); pd(F, V);
      fprintf(F, ",\"ivssim_r2t\":"); pd(F, LastR2T); fprintf(F, ",\"ivssim_t2r\":"); pd(F, LastT2R);
      if (A.Pics) {
        fprintf(F, ",\"scp0\":[");
        for (int c = 0; c < 3; c++) { if (c) fputc(',', F); dumpPlaneHex(F, &SCP0, c); }
        fprintf(F, "],\"scp1\":[");
        for (int c = 0; c < 3; c++) { if (c) fputc(',', F); dumpPlaneHex(F, &SCP1, c); }
        fputc(']', F);
      }
    }
    fputc('}', F);
  }
  fputc(']', F);

  // Frame averages as xMetricStat::calcAvgMetric (xAppQMIV.h:256-260).
  fprintf(F, ",\"avg\":{");
  bool First = true;
  for (int m = 0; m < 6; m++) {
    if (!A.M[m]) continue;
    fprintf(F, "%s\"%s\":{\"pic\":", First ? "" : ",", kMetricNames[m]);
    First = false;
    pd(F, xKBNS::Accumulate(PicVals[m]) / (flt64)NumFrames);
    if (!CmpVals[m].empty()) {
      fprintf(F, ",\"cmp\":");
      pd4(F, xKBNS::Accumulate(CmpVals[m]) / (flt64)NumFrames);
    }
    fputc('}', F);
  }
  fputc('}', F);
  if (A.Time) {
    const char* Names[11] = {"load", "validate", "preproc", "gcd", "scp", "psnr", "wspsnr", "ivpsnr", "ssim", "msssim", "ivssim"};
    fprintf(F, ",\"time_ms_per_frame\":{");
    for (int i = 0; i < 11; i++) { fprintf(F, "%s\"%s\":", i ? "," : "", Names[i]); pd(F, T[i] / NumFrames); }
    fputc('}', F);
  }
  fprintf(F, ",\"num_frames\":%d,\"threads\":%d,\"calc_psnrs\":%d}\n", NumFrames, A.Nth, (int)CalcPSNRs);

  if (CalcSSIMs) ProcSSIM.destroy();
  if (Pool) {
    ProcGCD.uninitThreadPool();
    ProcSCP.uninitThreadPool();
    ProcPSNR.uninitThreadPool();
    ProcSSIM.uninitThreadPool();
    Pool->destroy();
    delete Pool;
  }
  return 0;
}
