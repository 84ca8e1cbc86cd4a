// The reference application's per-frame processor calls (apps/QMIV/src/xAppQMIV.cpp:693-727,
// 831-935), made through the C++ processor classes of include/qmiv.hpp on the MI355X engine.
//
//   test_processors W H BD CF FILE0 FILE1 ERP OUT_SCP0 OUT_SCP1 OUT_CLR [MASK BDM CFM]
//
// FILE0 / FILE1 hold one packed planar frame each (InputFile0 / InputFile1), read through xPicP::fromPacked
// and through xSeq. Prints one JSON
// object with every metric (doubles in %.17g) and writes the two shift-compensated pictures and
// picture 0 converted RGB -> YCbCr (BT.709) as planar 4:4:4 16-bit files. With a mask file, the
// masked processor calls of the application's masked flow are printed too ("masked"). Checks the identities the processor API promises and exits 1 on a
// failed check.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/qmiv.hpp"

using namespace qmiv;

static int g_Failures = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #cond); \
      g_Failures++;                                                        \
    }                                                                      \
  } while (0)

static std::vector<uint8_t> readFile(const char* Path, size_t Bytes) {
  std::vector<uint8_t> B(Bytes);
  FILE* F = std::fopen(Path, "rb");
  if (!F || std::fread(B.data(), 1, Bytes, F) != Bytes) {
    std::fprintf(stderr, "cannot read %zu bytes from %s\n", Bytes, Path);
    std::exit(2);
  }
  std::fclose(F);
  return B;
}

static std::vector<uint16_t> planes16(const xPicP& Pic) {
  const size_t N = 3 * (size_t)Pic.getArea();
  std::vector<uint16_t> Out(N);
  if (Pic.getBitDepth() <= 8) {
    std::vector<uint8_t> B(N);
    Pic.download(B.data());
    for (size_t i = 0; i < N; i++) Out[i] = B[i];
  } else {
    Pic.download(Out.data());
  }
  return Out;
}

static void writeScp(const xPicP& Pic, const char* Path) {
  const std::vector<uint16_t> Out = planes16(Pic);
  const size_t N = Out.size();
  FILE* F = std::fopen(Path, "wb");
  if (!F || std::fwrite(Out.data(), 2, N, F) != N) std::exit(2);
  std::fclose(F);
}

static void printV4(const char* Name, const flt64V4& V) {
  std::printf("\"%s\":[%.17g,%.17g,%.17g],", Name, V[0], V[1], V[2]);
}

int main(int argc, char** argv) {
  if (argc != 11 && argc != 14) {
    std::fprintf(stderr, "usage: test_processors W H BD CF FILE0 FILE1 ERP OUT_SCP0 OUT_SCP1 OUT_CLR [MASK BDM CFM]\n");
    return 2;
  }
  const int32 W = std::atoi(argv[1]), H = std::atoi(argv[2]), BD = std::atoi(argv[3]), CF = std::atoi(argv[4]);
  const bool Erp = std::atoi(argv[7]) != 0;
  const size_t Bytes = qmiv_frame_bytes(W, H, BD, CF);
  if (Bytes == 0) return 2;
  std::vector<uint8_t> F0 = readFile(argv[5], Bytes), F1 = readFile(argv[6], Bytes);

  try {
    xContext Ctx(0);
    xPicP PicInP0(Ctx, {W, H}, BD), PicInP1(Ctx, {W, H}, BD);
    xPicP PicSCP0(Ctx, {W, H}, BD), PicSCP1(Ctx, {W, H}, BD);
    // InputFile0 through the unpack of a packed frame, InputFile1 through the sequence reader
    // (xSeq::openFile / readFrame, xAppQMIV.cpp:537-541, :674); both paths must give the same picture.
    PicInP0.fromPacked(F0.data(), CF);
    xSeq Seq0({W, H}, BD, CF), Seq1({W, H}, BD, CF);
    CHECK(Seq0.openFile(argv[5], xSeq::eMode::Read) && Seq1.openFile(argv[6], xSeq::eMode::Read));
    CHECK(Seq1.getNumOfFrames() == 1 && Seq1.seekFrame(1) == xSeq::eRetv::WrongArg);
    CHECK(Seq1.readFrame(&PicInP1));
    CHECK(Seq1.readFrame(&PicInP1) == xSeq::eRetv::EndOfFile);
    {
      xPicP Check0(Ctx, {W, H}, BD);
      CHECK(Seq0.readFrame(&Check0));
      CHECK(planes16(Check0) == planes16(PicInP0));
      xPicP Check1(Ctx, {W, H}, BD);
      Check1.fromPacked(F1.data(), CF);
      CHECK(planes16(Check1) == planes16(PicInP1));
    }

    xGlobClrDiffProc ProcGCD(Ctx);
    xShftCompPicProc ProcSCP(Ctx);
    xIVPSNR ProcPSNR(Ctx);
    xIVSSIM ProcSSIM(Ctx);
    ProcPSNR.initWS(Erp, W, H, BD, 360, 180);
    ProcSSIM.create({W, H}, BD, 4, true);

    // xAppQMIV.cpp:693-703
    const int32V4 GCD_R2T = ProcGCD.CalcGlobalColorDiff(&PicInP0, &PicInP1);
    ProcSCP.GenShftCompPics(&PicSCP1, &PicSCP0, &PicInP1, &PicInP0, GCD_R2T);

    // xAppQMIV.cpp:831-935
    const flt64V4 PSNR = ProcPSNR.calcPicPSNR(&PicInP0, &PicInP1);
    const flt64V4 WSPSNR = ProcPSNR.calcPicWSPSNR(&PicInP0, &PicInP1);
    flt64 R2T = 0, T2R = 0;
    ProcPSNR.setDebugCallbackQAP([&](flt64 A, flt64 B) { R2T = A; T2R = B; });
    const flt64 IVPSNR = ProcPSNR.calcPicIVPSNR(&PicInP0, &PicInP1, GCD_R2T);
    const flt64V4 SSIM = ProcSSIM.calcPicSSIM(&PicInP0, &PicInP1);
    const flt64V4 MSSSIM = ProcSSIM.calcPicMSSSIM(&PicInP0, &PicInP1);
    flt64 SR2T = 0, ST2R = 0;
    ProcSSIM.setDebugCallbackQAP([&](flt64 A, flt64 B) { SR2T = A; ST2R = B; });
    const flt64 IVSSIM = ProcSSIM.calcPicIVSSIM(&PicInP0, &PicInP1, &PicSCP0, &PicSCP1);

    // The application's default, interleaved pictures (xAppQMIV.cpp:314,856-857,892):
    // rearrangeFromPlanar, then calcPicIVPSNR(const xPicI*, const xPicI*, GCD) - the same value.
    xPicI PicInI0(Ctx, {W, H}, BD), PicInI1(Ctx, {W, H}, BD);
    PicInI0.rearrangeFromPlanar(&PicInP0);
    PicInI1.rearrangeFromPlanar(&PicInP1);
    const flt64 IVPSNRI = ProcPSNR.calcPicIVPSNR(&PicInI0, &PicInI1, GCD_R2T);
    CHECK(IVPSNRI == IVPSNR);
    // The reference's host layouts, margins included: xPicP planes and xPicI pels out and back in.
    {
      const int32 M = PicInP0.getMargin(), S = PicInP0.getStride();
      CHECK(M == 4 && S == W + 8);
      std::vector<uint16_t> Buf(3 * (size_t)S * (H + 2 * M), 0xFFFF);
      uint16_t* O[3];
      for (int c = 0; c < 3; c++) O[c] = Buf.data() + (size_t)c * S * (H + 2 * M) + (size_t)M * S + M;
      PicInP0.toPlanar(O, S);
      CHECK(Buf[0] == 0xFFFF);  // the margin is left alone
      xPicP Back(Ctx, {W, H}, BD);
      Back.fromPlanar(O, S);
      CHECK(Back.equalPic(&PicInP0) && !Back.equalPic(&PicInP1));
      CHECK(Back.equalCmp(&PicInP0, eCmp::CB));
      const int32 SI = PicInI1.getStride();
      std::vector<uint16_t> Pels(4 * (size_t)SI * (H + 2 * M), 0);
      PicInI1.toInterleaved(Pels.data() + 4 * ((size_t)M * SI + M), SI);
      xPicI BackI(Ctx, {W, H}, BD);
      BackI.fromInterleaved(Pels.data() + 4 * ((size_t)M * SI + M), SI);
      CHECK(ProcPSNR.calcPicIVPSNR(&PicInI0, &BackI, GCD_R2T) == IVPSNR);
      CHECK(PicInP0.check("InputFile0"));  // in range: true, nothing printed
    }

    // Identities of the processor API.
    CHECK(IVPSNR == (R2T < T2R ? R2T : T2R));
    CHECK(IVSSIM == (ST2R < SR2T ? ST2R : SR2T));
    // The two-argument IV-SSIM computes its own GCD and SCP: same value as the four-argument form.
    int32V4 GcdSeen{};
    ProcSSIM.setDebugCallbackGCS([&](const int32V4& G) { GcdSeen = G; });
    CHECK(ProcSSIM.calcPicIVSSIM(&PicInP0, &PicInP1) == IVSSIM);
    CHECK(GcdSeen == GCD_R2T);
    // Without WS the WS-PSNR processor returns plain PSNR.
    if (!Erp) CHECK(WSPSNR == PSNR);
    // SSIM without the luminance term is never below the full SSIM for the chosen inputs' luma.
    const flt64V4 CS = ProcSSIM.calcPicSSIM(&PicInP0, &PicInP1, false);
    // IV-PSNR with the colour difference computed by the processor itself.
    int32V4 GcdAuto{};
    ProcPSNR.setDebugCallbackGCS([&](const int32V4& G) { GcdAuto = G; });
    const flt64 IVPSNRAuto = ProcPSNR.calcPicIVPSNR(&PicInP0, &PicInP1);
    CHECK(GcdAuto == GCD_R2T);
    CHECK(IVPSNRAuto == IVPSNR);
    // Errors surface as exceptions: incompatible pictures.
    xPicP Small(Ctx, {W / 2 * 2 == W ? W - 2 : W - 1, H}, BD);
    bool Threw = false;
    try {
      (void)ProcPSNR.calcPicPSNR(&PicInP0, &Small);
    } catch (const xError& E) {
      Threw = E.status() == QMIV_ERR_INVALID_ARG;
    }
    CHECK(Threw);

    // xAppQMIV.cpp:778-813 with picture 0 taken as R, G, B planes: to YCbCr out of place; back to
    // RGB in place stays within the conversion's rounding; two plane exchanges restore it.
    xPicP Clr(Ctx, {W, H}, BD);
    xColorSpace::ConvertRGB2YCbCr(&Clr, &PicInP0, eClrSpcLC::BT709);
    writeScp(Clr, argv[10]);
    const std::vector<uint16_t> Before = planes16(Clr);
    CHECK(Clr.swapComponents(0, 2));
    const std::vector<uint16_t> Swapped = planes16(Clr);
    const size_t Area = (size_t)Clr.getArea();
    CHECK(std::equal(Swapped.begin(), Swapped.begin() + Area, Before.begin() + 2 * Area));
    CHECK(Clr.swapComponents(2, 0));
    CHECK(planes16(Clr) == Before);
    xColorSpace::ConvertYCbCr2RGB(&Clr, &Clr, eClrSpcLC::BT709);
    CHECK(!(planes16(Clr) == Before));

    writeScp(PicSCP0, argv[8]);
    writeScp(PicSCP1, argv[9]);

    // Masked flow (xAppQMIV.cpp:695,825-889) through the masked processor methods.
    std::string Masked;
    if (argc == 14) {
      const int32 BDM = std::atoi(argv[12]), CFM = std::atoi(argv[13]);
      std::vector<uint8_t> FM = readFile(argv[11], qmiv_frame_bytes(W, H, BDM, CFM));
      xPicP PicMsk(Ctx, {W, H}, BDM);
      PicMsk.fromPacked(FM.data(), CFM);
      const int32 NNM = countNonZero(Ctx, &PicMsk);
      const int32V4 GM = ProcGCD.CalcGlobalColorDiffM(&PicInP0, &PicInP1, &PicMsk, NNM);
      const flt64V4 PM = ProcPSNR.calcPicPSNRM(&PicInP0, &PicInP1, &PicMsk, NNM);
      const flt64V4 WM = ProcPSNR.calcPicWSPSNRM(&PicInP0, &PicInP1, &PicMsk, NNM);
      flt64 MR2T = 0, MT2R = 0;
      int32 SeenNNM = -1;
      ProcPSNR.setDebugCallbackQAP([&](flt64 A, flt64 B) { MR2T = A; MT2R = B; });
      ProcPSNR.setDebugCallbackMSK([&](int32 N) { SeenNNM = N; });
      const flt64 IM = ProcPSNR.calcPicIVPSNRM(&PicInP0, &PicInP1, &PicMsk, NNM, GM);
      // the application's masked call takes interleaved pictures (xAppQMIV.cpp:889)
      CHECK(ProcPSNR.calcPicIVPSNRM(&PicInI0, &PicInI1, &PicMsk, NNM, GM) == IM);
      // The self-contained form counts the mask and finds the same colour difference.
      int32V4 GAuto{};
      ProcPSNR.setDebugCallbackGCS([&](const int32V4& G) { GAuto = G; });
      CHECK(ProcPSNR.calcPicIVPSNRM(&PicInP0, &PicInP1, &PicMsk) == IM);
      CHECK(GAuto == GM);
      CHECK(SeenNNM == NNM);
      char Buf[1024];
      std::snprintf(Buf, sizeof(Buf),
                    "\"masked\":{\"nnm\":%d,\"gcd\":[%d,%d,%d,%d],\"psnr\":[%.17g,%.17g,%.17g],"
                    "\"wspsnr\":[%.17g,%.17g,%.17g],\"ivpsnr\":%.17g,\"r2t\":%.17g,\"t2r\":%.17g},",
                    NNM, GM[0], GM[1], GM[2], GM[3], PM[0], PM[1], PM[2], WM[0], WM[1], WM[2], IM, MR2T, MT2R);
      Masked = Buf;
    }

    std::printf("{\"gcd\":[%d,%d,%d,%d],", GCD_R2T[0], GCD_R2T[1], GCD_R2T[2], GCD_R2T[3]);
    printV4("psnr", PSNR);
    printV4("wspsnr", WSPSNR);
    std::printf("\"ivpsnr\":%.17g,\"ivpsnr_r2t\":%.17g,\"ivpsnr_t2r\":%.17g,\"ivpsnr_interleaved\":%.17g,", IVPSNR,
                R2T, T2R, IVPSNRI);
    printV4("ssim", SSIM);
    printV4("ssim_cs", CS);
    printV4("msssim", MSSSIM);
    std::printf("\"ivssim\":%.17g,\"ivssim_r2t\":%.17g,\"ivssim_t2r\":%.17g,", IVSSIM, SR2T, ST2R);
    std::printf("%s\"failures\":%d}\n", Masked.c_str(), g_Failures);
  } catch (const xError& E) {
    std::fprintf(stderr, "error: %s\n", E.what());
    return 3;
  }
  return g_Failures ? 1 : 0;
}
