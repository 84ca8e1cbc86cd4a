// Command-line driver with the reference application's parameters, frame flow and report format
// (apps/QMIV/src/main_QMIV.cpp:36-112 program flow; apps/QMIV/src/xAppQMIV.cpp:149-339 parameters,
// :341-392 configuration report, :492-560 sequences, :661-935 frame loop, :966-1030 results;
// apps/QMIV/src/xAppQMIV.h:176-306 per-metric statistics). Every metric runs on the MI355X engine
// through the C ABI (include/qmiv.h); this file holds only parsing, file I/O ordering and printing.
//
//   qmiv -i0 A.yuv -i1 B.yuv -ps 4096x2048 -bd 10 -erp -ml "PSNR, WSPSNR, IVPSNR" -r results.txt
//   qmiv -c config.cfg [-c more.cfg] [-v 2]
//
// Parameters are "-short value" on the command line or "LongName = value" lines in files given with
// -c (a later occurrence overrides an earlier one):
//   i0 InputFile0, i1 InputFile1, ff FileFormat (RAW, PNG), ps PictureSize (WxH), pw PictureWidth,
//   ph PictureHeight, pf PictureFormat (yuv420p, yuv444p10le, ...), bd BitDepth (8..14),
//   cf ChromaFormat (420, 422, 444), s0 StartFrame0, s1 StartFrame1, nf NumberOfFrames, r ResultFile,
//   ml MetricList (comma list of PSNR WSPSNR IVPSNR SSIM MSSSIM IVSSIM, or All), im InputFileM,
//   bdm BitDepthM, cfm ChromaFormatM, erp Equirectangular (flag), lor LonRangeDeg, lar LatRangeDeg,
//   csi ColorSpaceInput, csm ColorSpaceMetric (RGB BGR GBR YCbCr YCbCr_BT601 YCbCr_SMPTE170M
//   YCbCr_BT709 YCbCr_SMPTE240M YCbCr_BT2020), sr SearchRange, cws CmpWeightsSearch (a:b:c:d),
//   cwa CmpWeightsAverage, unc UnnoticeableCoef (a:b:c:d), ipa InvalidPelActn (STOP WARN SKIP CNCL), nma
//   NameMismatchActn, nth NumberOfThreads, ilp InterleavedPic, v VerboseLevel (0..4).
// Chroma formats 4:2:0 (any size), 4:2:2 and 4:4:4 (the application's parser reads "400" as 4:2:0,
// xMiscUtilsCORE.cpp:13-21); InvalidPelActn STOP, WARN, SKIP and CNCL with the reference's check
// report (FILE BROKEN / DETECTED lines). VerboseLevel 3 runs each frame stage by stage and prints the
// application's AvgTime table (LOAD, VALIDATE, PREPROC, GCD, SCP, per metric), the stages timed on
// the GPU up to their results. Differences from the reference application: with FileFormat PNG,
// BitDepth and BitDepthM must be 8 (the files are decoded to 8-bit R, G, B planes,
// xSeqPNG.cpp:14-36); NumberOfThreads and InterleavedPic are reported and otherwise ignored (the
// engine runs on one GPU and its results do not depend on either).
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/qmiv.h"

namespace {

using Clock = std::chrono::steady_clock;

enum Metric { kPSNR = 0, kWSPSNR, kIVPSNR, kSSIM, kMSSSIM, kIVSSIM, kNumMetrics };
const char* const kMetricName[kNumMetrics] = {"PSNR", "WSPSNR", "IVPSNR", "SSIM", "MSSSIM", "IVSSIM"};
const uint32_t kMetricBit[kNumMetrics] = {QMIV_PSNR, QMIV_WSPSNR, QMIV_IVPSNR, QMIV_SSIM, QMIV_MSSSIM, QMIV_IVSSIM};
// Metric properties: xMetricInfo (apps/QMIV/src/xAppQMIV.h:146-172).
const bool kPerCmp[kNumMetrics] = {true, true, false, true, true, false};
const bool kPerPic[kNumMetrics] = {false, false, true, false, false, true};
const bool kNormalized[kNumMetrics] = {false, false, false, true, true, true};
const char* const kUnit[kNumMetrics] = {"dB", "dB", "dB", "  ", "  ", "  "};
const char* const kDescription[kNumMetrics] = {
    "Peak Signal-to-Noise Ratio", "Spherical Weighted - Peak Signal-to-Noise Ratio",
    "Immersive Video - Peak Signal-to-Noise Ratio", "Structural Similarity Index Measure",
    "Multi Scale Structural Similarity Index Measure", "Immersive Video - Structural Similarity Index Measure"};

// Colour spaces of the application: eClrSpcApp and its predicates (apps/QMIV/src/xAppQMIV.h:60-110).
enum ClrSpc { kInvalidCS = -1, kRGB = 2, kBGR = 3, kGBR = 4, kYCbCr = 16, kBT601, kSMPTE170M, kBT709, kSMPTE240M, kBT2020 };
const std::map<std::string, ClrSpc>& clr_spc_names() {
  static const std::map<std::string, ClrSpc> m = {
      {"RGB", kRGB}, {"BGR", kBGR}, {"GBR", kGBR}, {"YCbCr", kYCbCr}, {"YCbCr_BT601", kBT601},
      {"YCbCr_SMPTE170M", kSMPTE170M}, {"YCbCr_BT709", kBT709}, {"YCbCr_SMPTE240M", kSMPTE240M},
      {"YCbCr_BT2020", kBT2020}};
  return m;
}
ClrSpc str2cs(const std::string& s) {
  const auto it = clr_spc_names().find(s);
  return it == clr_spc_names().end() ? kInvalidCS : it->second;
}
const char* cs2str(ClrSpc c) {
  for (const auto& kv : clr_spc_names())
    if (kv.second == c) return kv.first.c_str();
  return "INVALID";
}
bool is_rgb(ClrSpc c) { return c == kRGB || c == kBGR || c == kGBR; }
bool is_ycbcr(ClrSpc c) { return (int)c >= (int)kYCbCr; }
bool is_defined_ycbcr(ClrSpc c) { return (int)c > (int)kYCbCr; }
int32_t lc_space(ClrSpc c) {  // xClrSpcAppToClrSpc (apps/QMIV/src/xAppQMIV.h:112-124)
  switch (c) {
    case kBT601: case kSMPTE170M: return QMIV_CLR_BT601;
    case kBT709: return QMIV_CLR_BT709;
    case kSMPTE240M: return QMIV_CLR_SMPTE240M;
    case kBT2020: return QMIV_CLR_BT2020;
    default: return -1;
  }
}

std::string format(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string format(const char* f, ...) {
  char buf[4096];
  va_list ap;
  va_start(ap, f);
  std::vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

// Neumaier summation, xKBNS (libs/pmbb-core/src/xMathUtils.h:16-40, xMathUtils.cpp:12-41).
double kbns(const std::vector<double>& x) {
  double s = 0.0, c = 0.0;
  for (const double v : x) {
    const double t = s + v;
    if (std::fabs(s) >= std::fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  return s + c;
}

// Per-metric values and their report lines: xMetricStat (apps/QMIV/src/xAppQMIV.h:176-306).
struct MetricStat {
  int m = 0;
  std::string suffix_cmp, suffix_pic;
  int32_t wavg[4] = {0, 0, 0, 0};
  double inv_denom = 0.0;
  std::vector<double> cmp[3], pic;
  double avg_cmp[3] = {0, 0, 0}, avg_pic = 0;

  void init(int metric, bool mask, bool rgb, const int32_t cmp_weights_average[4]) {
    m = metric;
    suffix_cmp = mask && rgb ? "-M R:G:B  " : mask ? "-M Y:Cb:Cr" : rgb ? " R:G:B    " : " Y:Cb:Cr  ";
    if (kPerPic[m]) suffix_pic = mask && rgb ? "-M-RGB    " : mask ? "-M        " : rgb ? "-RGB      " : "          ";
    else suffix_pic = mask && rgb ? "-M-RGB    " : mask ? "-M-YCbCr  " : rgb ? "-RGB      " : "-YCbCr    ";
    std::copy(cmp_weights_average, cmp_weights_average + 4, wavg);
    inv_denom = 1.0 / (double)(wavg[0] + wavg[1] + wavg[2] + wavg[3]);
  }
  void set_cmp(const double v[3]) {
    for (int c = 0; c < 3; c++) cmp[c].push_back(v[c]);
    pic.push_back((v[0] * wavg[0] + v[1] * wavg[1] + v[2] * wavg[2]) * inv_denom);
  }
  void set_pic(double v) { pic.push_back(v); }
  void average(int32_t frames) {
    avg_pic = kbns(pic) / (double)frames;
    if (kPerCmp[m])
      for (int c = 0; c < 3; c++) avg_cmp[c] = kbns(cmp[c]) / (double)frames;
  }
  std::string frame_cmp(size_t f) const {
    const char* one = kNormalized[m] ? "%8.6f " : "%8.4f ";
    std::string r = format("%8s%s ", kMetricName[m], suffix_cmp.c_str());
    for (int c = 0; c < 3; c++) r += format(one, cmp[c][f]);
    return r;
  }
  std::string frame_pic(size_t f) const {
    const char* one = kNormalized[m] ? "%8.6f " : "%8.4f ";
    return format("%8s%s ", kMetricName[m], suffix_pic.c_str()) + format(one, pic[f]) + "                  ";
  }
  std::string avg_lines(const std::string& header) const {
    const char* name = header.empty() ? "%-8s%s " : "%8s%s ";
    const char* one = kNormalized[m] ? "%10.8f %s  " : "%10.6f %s  ";
    std::string r;
    if (kPerCmp[m]) {
      r += header + format(name, kMetricName[m], suffix_cmp.c_str());
      for (int c = 0; c < 3; c++) r += format(one, avg_cmp[c], kUnit[m]);
      r += "\n";
    }
    r += header + format(name, kMetricName[m], suffix_pic.c_str()) + format(one, avg_pic, kUnit[m]) + "                  ";
    return r;
  }
};

// ------------------------------------------------------------------------------------------------
// Parameters (xCfgINI: command line "-short value", files "Name = Value").
// ------------------------------------------------------------------------------------------------
const std::map<std::string, std::string>& short_names() {
  static const std::map<std::string, std::string> m = {
      {"i0", "InputFile0"}, {"i1", "InputFile1"}, {"ff", "FileFormat"}, {"ps", "PictureSize"},
      {"pw", "PictureWidth"}, {"ph", "PictureHeight"}, {"pf", "PictureFormat"}, {"bd", "BitDepth"},
      {"cf", "ChromaFormat"}, {"s0", "StartFrame0"}, {"s1", "StartFrame1"}, {"nf", "NumberOfFrames"},
      {"r", "ResultFile"}, {"ml", "MetricList"}, {"im", "InputFileM"}, {"bdm", "BitDepthM"},
      {"cfm", "ChromaFormatM"}, {"erp", "Equirectangular"}, {"lor", "LonRangeDeg"}, {"lar", "LatRangeDeg"},
      {"csi", "ColorSpaceInput"}, {"csm", "ColorSpaceMetric"}, {"sr", "SearchRange"},
      {"cws", "CmpWeightsSearch"}, {"cwa", "CmpWeightsAverage"}, {"unc", "UnnoticeableCoef"},
      {"ipa", "InvalidPelActn"}, {"nma", "NameMismatchActn"}, {"nth", "NumberOfThreads"},
      {"ilp", "InterleavedPic"}, {"v", "VerboseLevel"}};
  return m;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  std::string t = s.substr(a, b - a);
  if (t.size() >= 2 && t.front() == '"' && t.back() == '"') t = t.substr(1, t.size() - 2);
  return t;
}
std::string upper(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

struct Params {
  std::map<std::string, std::string> v;  // long name -> value
  std::string log;                       // parsing and validation errors, one "!  ..." line each
  bool has(const std::string& k) const { return v.count(k) != 0; }
  std::string str(const std::string& k, const std::string& d = "") const {
    const auto it = v.find(k);
    return it == v.end() ? d : it->second;
  }
  int32_t num(const std::string& k, int32_t d) {
    const auto it = v.find(k);
    if (it == v.end()) return d;
    char* end = nullptr;
    const long x = std::strtol(it->second.c_str(), &end, 10);
    if (end == it->second.c_str() || *end != '\0') {
      log += "!  Invalid " + k + " value\n";
      return d;
    }
    return (int32_t)x;
  }
};

bool load_file(const std::string& path, Params& p) {
  std::ifstream f(path);
  if (!f) {
    p.log += "!  Cannot read configuration file " + path + "\n";
    return false;
  }
  std::string line;
  while (std::getline(f, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    const std::string k = trim(line.substr(0, eq));
    if (!k.empty()) p.v[k] = trim(line.substr(eq + 1));
  }
  return true;
}

bool load_cmdline(int argc, char** argv, Params& p) {
  const auto& sn = short_names();
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (a.size() < 2 || a[0] != '-') {
      p.log += "!  Unexpected argument " + a + "\n";
      return false;
    }
    std::string key = a.substr(a[1] == '-' ? 2 : 1);
    if (key == "c") {
      if (i + 1 >= argc) { p.log += "!  -c needs a file name\n"; return false; }
      if (!load_file(argv[++i], p)) return false;
      continue;
    }
    const auto it = sn.find(key);
    if (it != sn.end()) key = it->second;
    else if (std::none_of(sn.begin(), sn.end(), [&](const auto& kv) { return kv.second == key; })) {
      p.log += "!  Unknown parameter " + a + "\n";
      return false;
    }
    if (key == "Equirectangular") {  // a flag (xCfgINI addCmdFlag): present means 1
      p.v[key] = "1";
      continue;
    }
    if (i + 1 >= argc) { p.log += "!  " + a + " needs a value\n"; return false; }
    p.v[key] = argv[++i];
  }
  return true;
}

int32_t str2cf(const std::string& s) {  // xStr2CrF (libs/pmbb-core/src/xMiscUtilsCORE.cpp:12-21)
  const std::string u = upper(s);
  if (u == "CF444" || u == "444") return 444;
  if (u == "CF422" || u == "422") return 422;
  if (u == "CF420" || u == "420" || u == "CF400" || u == "400") return 420;
  return -1;
}
// FFmpeg pixel formats "yuvNNNp[BBle]" (xFmtScn::scanPixelFormat, libs/pmbb-core/src/xFmtScn.cpp:43-96).
bool scan_pix_fmt(const std::string& s, int32_t& bd, int32_t& cf) {
  const std::string t = trim(s);
  if (t.size() < 7 || t.compare(0, 3, "yuv") != 0 || t[6] != 'p') return false;
  cf = std::atoi(t.substr(3, 3).c_str());
  if (cf != 444 && cf != 422 && cf != 420) return false;
  const std::string rest = t.substr(7);
  if (rest.empty()) { bd = 8; return true; }
  if (rest.size() < 3 || rest.compare(rest.size() - 2, 2, "le") != 0) return false;
  bd = std::atoi(rest.substr(0, rest.size() - 2).c_str());
  return bd > 0;
}
bool scan_int4(const std::string& s, int32_t v[4]) {
  return std::sscanf(s.c_str(), "%d:%d:%d:%d", &v[0], &v[1], &v[2], &v[3]) == 4;
}
bool scan_flt4(const std::string& s, float v[4]) {
  return std::sscanf(s.c_str(), "%f:%f:%f:%f", &v[0], &v[1], &v[2], &v[3]) == 4;
}

[[noreturn]] void fail(const std::string& msg) {
  std::fflush(stdout);
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(EXIT_FAILURE);
}
void check(qmiv_status st, const std::string& what) {
  if (st != QMIV_OK) fail("ERROR --> " + what + ": " + qmiv_last_error());
}

struct Owned {
  qmiv_ctx* ctx = nullptr;
  qmiv_seq* seq[3] = {nullptr, nullptr, nullptr};
  qmiv_pic* in[3] = {nullptr, nullptr, nullptr};
  qmiv_pic* up[2] = {nullptr, nullptr};
  qmiv_pic* scp[2] = {nullptr, nullptr};  // the staged route's shift-compensated pictures
  ~Owned() {
    for (qmiv_seq* s : seq) qmiv_seq_close(s);
    for (qmiv_pic* p : in) qmiv_pic_destroy(p);
    for (qmiv_pic* p : up) qmiv_pic_destroy(p);
    for (qmiv_pic* p : scp) qmiv_pic_destroy(p);
    qmiv_destroy(ctx);
  }
};

}  // namespace

int main(int argc, char** argv) {
  std::printf("QMIV: PSNR, WS-PSNR, IV-PSNR, SSIM, MS-SSIM, IV-SSIM on an AMD Instinct GPU\n\n");
  const auto app_beg = Clock::now();

  // --- configuration: xAppQMIV::readConfiguration (apps/QMIV/src/xAppQMIV.cpp:205-339) ----------
  Params P;
  if (argc < 2 || !load_cmdline(argc, argv, P)) fail("! invalid commandline\n" + P.log + "\nParameters: see qmiv_cli.cpp");
  const std::string in_file[3] = {P.str("InputFile0"), P.str("InputFile1"), P.str("InputFileM")};
  if (in_file[0].empty()) P.log += "!  InputFile0 is empty\n";
  if (in_file[1].empty()) P.log += "!  InputFile1 is empty\n";
  const std::string file_format = upper(P.str("FileFormat", "RAW"));
  const bool png = file_format == "PNG";
  if (file_format != "RAW" && !png) P.log += "!  Invalid FileFormat " + file_format + "\n";
  int32_t w = 0, h = 0;
  if (P.has("PictureSize")) {
    if (std::sscanf(P.str("PictureSize").c_str(), "%dx%d", &w, &h) != 2 || w <= 0 || h <= 0)
      P.log += "!  Invalid PictureSize value\n";
  } else {
    w = P.num("PictureWidth", -1);
    h = P.num("PictureHeight", -1);
    if (w <= 0) P.log += "!  Invalid PictureWidth value\n";
    if (h <= 0) P.log += "!  Invalid PictureHeight value\n";
  }
  int32_t bd = 8, cf = 420;
  if (P.has("PictureFormat")) {
    if (!scan_pix_fmt(P.str("PictureFormat"), bd, cf)) P.log += "!  Invalid or unsuported PictureFormat\n";
  } else {
    bd = P.num("BitDepth", 8);
    cf = str2cf(P.str("ChromaFormat", "420"));
  }
  if (bd < 8 || bd > 14) P.log += "!  Invalid or unsuported BitDepth value\n";
  if (cf != 420 && cf != 422 && cf != 444) P.log += "!  Invalid or unsuported ChromaFormat value\n";
  const int32_t start[2] = {P.num("StartFrame0", 0), P.num("StartFrame1", 0)};
  if (start[0] < 0 || start[1] < 0) P.log += "!  StartFrame value cannot be negative\n";
  const int32_t number_of_frames = P.num("NumberOfFrames", -1);
  const std::string result_file = P.str("ResultFile");
  bool calc[kNumMetrics] = {false, false, false, false, false, false};
  if (P.has("MetricList")) {
    const std::string list = P.str("MetricList");
    size_t pos = 0;
    while (pos <= list.size()) {
      const size_t comma = std::min(list.find(',', pos), list.size());
      const std::string tok = trim(list.substr(pos, comma - pos));
      pos = comma + 1;
      if (upper(tok) == "ALL") {
        std::fill(calc, calc + kNumMetrics, true);
        break;
      }
      int found = -1;
      for (int m = 0; m < kNumMetrics; m++)
        if (upper(tok) == kMetricName[m]) found = m;
      if (found >= 0) calc[found] = true;
      else P.log += "!  MetricList contains not valid entry. Token \"" + tok + "\" in not a metric. \n";
    }
  } else {
    calc[kPSNR] = calc[kWSPSNR] = calc[kIVPSNR] = calc[kIVSSIM] = true;
  }
  const int32_t bdm = P.num("BitDepthM", bd);
  const int32_t cfm = P.has("ChromaFormatM") ? str2cf(P.str("ChromaFormatM")) : cf;
  if (bdm < 8 || bdm > 14) P.log += "!  Invalid or unsuported BitDepthM value\n";
  if (cfm != 420 && cfm != 422 && cfm != 444) P.log += "!  Invalid or unsuported ChromaFormatM value\n";
  if (png && (bd != 8 || bdm != 8)) P.log += "!  PNG input is 8-bit: BitDepth and BitDepthM must be 8\n";
  qmiv_params prm;
  qmiv_params_default(&prm);
  prm.erp = P.num("Equirectangular", 0) != 0;
  prm.lon_range_deg = P.num("LonRangeDeg", 360);
  prm.lat_range_deg = P.num("LatRangeDeg", 180);
  const ClrSpc csi = str2cs(P.str("ColorSpaceInput", "YCbCr")), csm = str2cs(P.str("ColorSpaceMetric", "YCbCr"));
  if (csi == kInvalidCS || csm == kInvalidCS) P.log += "!  Invalid ColorSpaceInput or ColorSpaceMetric value\n";
  if (csi != csm) {
    if (is_ycbcr(csi) && is_ycbcr(csm)) P.log += "!  YCbCr to YCbCr conversion is not supported.\n";
    if (csi == kYCbCr || csm == kYCbCr) P.log += "!  Generic YCbCr cannot be used for colorspace conversion.\n";
  }
  const int32_t default_sr = prm.search_range;
  int32_t default_w[4];
  float default_unc[4];
  std::copy(prm.cmp_weights_search, prm.cmp_weights_search + 4, default_w);
  std::copy(prm.unnoticeable_coef, prm.unnoticeable_coef + 4, default_unc);
  prm.search_range = P.num("SearchRange", default_sr);
  if (prm.search_range < 1) P.log += "!  SearchRange value must non-zero\n";
  if (P.has("CmpWeightsSearch") && !scan_int4(P.str("CmpWeightsSearch"), prm.cmp_weights_search))
    P.log += "!  Invalid CmpWeightsSearch value\n";
  if (P.has("CmpWeightsAverage") && !scan_int4(P.str("CmpWeightsAverage"), prm.cmp_weights_average))
    P.log += "!  Invalid CmpWeightsAverage value\n";
  if (P.has("UnnoticeableCoef") && !scan_flt4(P.str("UnnoticeableCoef"), prm.unnoticeable_coef))
    P.log += "!  Invalid UnnoticeableCoef value\n";
  const std::string ipa = upper(P.str("InvalidPelActn", "STOP")), nma = upper(P.str("NameMismatchActn", "WARN"));
  for (const std::string* a : {&ipa, &nma})
    if (*a != "SKIP" && *a != "WARN" && *a != "STOP" && *a != "CNCL") P.log += "!  Invalid action " + *a + "\n";
  prm.invalid_pel_action = ipa == "WARN" ? QMIV_PEL_WARN : ipa == "SKIP" ? QMIV_PEL_SKIP
                           : ipa == "CNCL" ? QMIV_PEL_CNCL : QMIV_PEL_STOP;
  const int32_t threads = P.num("NumberOfThreads", -2);
  const int32_t interleaved = P.num("InterleavedPic", 1);
  const int verbose = P.num("VerboseLevel", 1);

  const bool use_mask = !in_file[2].empty();
  const bool calc_ssims = calc[kSSIM] || calc[kMSSSIM] || calc[kIVSSIM];
  const bool cvt_ycbcr2rgb = is_defined_ycbcr(csi) && is_rgb(csm);
  const bool cvt_rgb2ycbcr = is_rgb(csi) && is_defined_ycbcr(csm);
  const bool reorder_rgb = is_rgb(csi) && csi != kRGB && csm == kRGB;
  const bool convert = cvt_ycbcr2rgb || cvt_rgb2ycbcr || reorder_rgb;
  if (use_mask && calc_ssims) P.log += "! Structural Similarity metrics cannot be combined with Mask mode\n";
  if (!P.log.empty()) fail("! invalid configuration\n" + P.log);
  uint32_t metrics = 0;
  for (int m = 0; m < kNumMetrics; m++)
    if (calc[m]) metrics |= kMetricBit[m];

  if (verbose >= 2) {
    std::printf("WorkingDir = %s\n\nCommandline args:\n", std::filesystem::current_path().string().c_str());
    for (int i = 0; i < argc; i++) std::printf("%s ", argv[i]);
    std::printf("\n");
  }
  // xAppQMIV::formatConfiguration (apps/QMIV/src/xAppQMIV.cpp:341-392)
  if (verbose >= 1) {
    const char* irr = "  (irrelevant)";
    std::string c = "Run-time configuration:\n";
    c += format("InputFile0        = %s\n", in_file[0].c_str());
    c += format("InputFile1        = %s\n", in_file[1].c_str());
    c += format("FileFormat        = %s\n", png ? "PNG" : "RAW");
    c += format("PictureSize       = %dx%d\n", w, h);
    c += format("BitDepth          = %d\n", bd);
    c += format("ChromaFormat      = CF%d\n", cf);
    c += format("StartFrame0       = %d\n", start[0]);
    c += format("StartFrame1       = %d\n", start[1]);
    c += format("NumberOfFrames    = %d%s\n", number_of_frames, number_of_frames == -1 ? "  (all)" : "");
    c += format("ResultFile        = %s\n", result_file.empty() ? "(unused)" : result_file.c_str());
    std::string ml;
    for (int m = 0; m < kNumMetrics; m++)
      if (calc[m]) ml += std::string(kMetricName[m]) + ", ";
    if (ml.size() >= 2) ml.resize(ml.size() - 2);
    c += "MetricList        = " + ml + "\n";
    c += format("InputFileM        = %s\n", use_mask ? in_file[2].c_str() : "(unused)");
    c += format("BitDepthM         = %d%s\n", bdm, use_mask ? "" : irr);
    c += format("ChromaFormatM     = CF%d%s\n", cfm, use_mask ? "" : irr);
    c += format("Equirectangular   = %d\n", prm.erp);
    c += format("LonRangeDeg       = %d%s\n", prm.lon_range_deg, prm.erp ? "" : irr);
    c += format("LatRangeDeg       = %d%s\n", prm.lat_range_deg, prm.erp ? "" : irr);
    c += format("ColorSpaceInput   = %s%s\n", cs2str(csi), convert ? "" : irr);
    c += format("ColorSpaceMetric  = %s%s\n", cs2str(csm), convert ? "" : irr);
    const auto dflt = [](bool d) { return d ? "  (default)" : "  (custom)"; };
    const int32_t* ws = prm.cmp_weights_search;
    const int32_t* wa = prm.cmp_weights_average;
    const float* uc = prm.unnoticeable_coef;
    c += format("SearchRange       = %d%s\n", prm.search_range, dflt(prm.search_range == default_sr));
    c += format("CmpWeightsSearch  = %d:%d:%d:%d%s\n", ws[0], ws[1], ws[2], ws[3], dflt(std::equal(ws, ws + 4, default_w)));
    c += format("CmpWeightsAverage = %d:%d:%d:%d%s\n", wa[0], wa[1], wa[2], wa[3], dflt(std::equal(wa, wa + 4, default_w)));
    c += format("UnnoticeableCoef  = %.2f:%.2f:%.2f:%.2f%s\n", uc[0], uc[1], uc[2], uc[3],
                dflt(std::equal(uc, uc + 4, default_unc)));
    c += format("InvalidPelActn    = %s\n", ipa.c_str());
    c += format("NameMismatchActn  = %s\n", nma.c_str());
    c += format("NumberOfThreads   = %d%s\n", threads, threads == -1 ? "  (all)" : threads == -2 ? "  (auto)" : "");
    c += format("InterleavedPic    = %d\n", interleaved != 0);
    c += format("VerboseLevel      = %d\n\n", verbose);
    c += "Run-time derrived parameters:\n";
    c += format("WindowSize        = %dx%d\n", 2 * prm.search_range + 1, 2 * prm.search_range + 1);
    c += format("PictureMargin     = %d\n", ((prm.search_range + 3) >> 2) << 2);
    c += format("UseMask           = %d\n\n", use_mask);
    c += "Selected metrics:\n";
    for (int m = 0; m < kNumMetrics; m++)
      if (calc[m]) c += format("%-9s - %s\n", kMetricName[m], kDescription[m]);
    std::printf("%s\n\n", c.c_str());
  }
  std::printf("\n\n\n");
  if (verbose >= 2) std::printf("Initializing:\n");

  // --- sequences: xAppQMIV::setupSeqAndBuffs (apps/QMIV/src/xAppQMIV.cpp:492-560) ---------------
  const char fid[3] = {'0', '1', 'M'};
  const int32_t num_inputs = use_mask ? 3 : 2;
  // PNG frames are 4:4:4 R, G, B pictures whatever ChromaFormat says (xSeqPNG::create, xSeqPNG.cpp:14-35).
  const int32_t bds[3] = {bd, bd, bdm};
  const int32_t cfs[3] = {png ? 444 : cf, png ? 444 : cf, png ? 444 : cfm};
  // Device pictures: 4:2:0 of even size and 4:4:4 as read; 4:2:2 and odd-size 4:2:0 frames are
  // unpacked on the device into 4:4:4 pictures, as the reference's reader unpacks every frame
  // (libs/pmbb-core/src/xSeq.cpp:179-249).
  int32_t dcf[3];
  for (int i = 0; i < 3; i++) dcf[i] = cfs[i] == 420 && !(w & 1) && !(h & 1) ? 420 : 444;
  Owned o;
  for (int i = 0; i < num_inputs && !png; i++)
    if (!std::filesystem::exists(in_file[i]))
      fail(format("ERROR --> InputFile%c does not exist (%s)", fid[i], in_file[i].c_str()));
  if (verbose >= 1 && !png)
    for (int i = 0; i < num_inputs; i++)
      std::printf("SizeOfInputFile%c = %lld\n", fid[i], (long long)std::filesystem::file_size(in_file[i]));
  int32_t frames_in[3] = {0, 0, 0};
  for (int i = 0; i < num_inputs; i++) {
    if (png) {
      if (qmiv_seq_open_png(in_file[i].c_str(), w, h, &o.seq[i]) != QMIV_OK)
        fail(format("ERROR --> InputFile%c does not exist (%s) [%s]", fid[i], in_file[i].c_str(), qmiv_last_error()));
    } else {
      check(qmiv_seq_open(in_file[i].c_str(), w, h, bds[i], cfs[i], &o.seq[i]), "InputFile opening failure (" + in_file[i] + ")");
    }
    frames_in[i] = qmiv_seq_num_frames(o.seq[i]);
    if (verbose >= 1) std::printf("DetectedFrames%d  = %d\n", i, frames_in[i]);
    const int32_t first = i < 2 ? start[i] : 0;
    if (first >= frames_in[i])
      fail(format("ERROR --> StartFrame%c >= DetectedFrames%c for (%s)", fid[i], fid[i], in_file[i].c_str()));
  }
  const int32_t min_all = std::min(frames_in[0], frames_in[1]);
  const int32_t min_rem = std::min(frames_in[0] - start[0], frames_in[1] - start[1]);
  const int32_t frames = std::min(number_of_frames > 0 ? number_of_frames : min_all, min_rem);
  if (verbose >= 1) std::printf("FramesToProcess  = %d\n", frames);
  std::printf("\n");
  if (use_mask && frames > frames_in[2]) fail("ERROR --> FramesToProcess > NumOfFramesM");

  check(qmiv_create(0, &o.ctx), "device initialisation");
  for (int i = 0; i < num_inputs; i++) check(qmiv_pic_create(o.ctx, w, h, bds[i], dcf[i], &o.in[i]), "picture allocation");
  // Colour conversion, plane reordering and the stage-by-stage route (VerboseLevel 3) work on 4:4:4
  // pictures, as the reference's do (its pictures are always 4:4:4 after reading:
  // libs/pmbb-core/src/xSeq.cpp:197-213).
  const bool up444 = (convert || verbose >= 3) && dcf[0] == 420;
  if (up444)
    for (int i = 0; i < 2; i++) check(qmiv_pic_create(o.ctx, w, h, bd, 444, &o.up[i]), "picture allocation");

  MetricStat st[kNumMetrics];
  for (int m = 0; m < kNumMetrics; m++) st[m].init(m, use_mask, is_rgb(csm), prm.cmp_weights_average);
  // With a conversion, the range check and the identical-component flags are those of the input
  // pictures before conversion (xAppQMIV.cpp:680-684, :758-765), so the engine's own substitution
  // for identical components is turned off and applied here from the input pictures' flags, with
  // the PSNR of SSD=1 (xPSNR::getFakePSNR).
  qmiv_params run_prm = prm;
  if (convert) run_prm.fake_psnr_for_exact = 0;
  const uint64_t max_pel = (uint64_t)((1 << bd) - 1);
  const double fake_psnr = 10.0 * std::log10((double)((uint64_t)(int64_t)(w * h) * (max_pel * max_pel)) / 1.0);

  // --- frame loop: xAppQMIV::processAllFrames (apps/QMIV/src/xAppQMIV.cpp:661-759) ---------------
  const bool print_frame = verbose >= 2, debug = verbose >= 4;
  // VerboseLevel 3 gathers per-stage times as the reference does (xAppQMIV.cpp:333, :670-744): each
  // frame then runs stage by stage through the processor-level calls, each stage timed up to its
  // results; below that level one fused engine call computes the whole frame.
  const bool staged = verbose >= 3;
  const int32_t act = prm.invalid_pel_action;
  const bool need_gcd = calc[kIVPSNR] || calc[kIVSSIM];
  double t_stage[5] = {0, 0, 0, 0, 0};  // LOAD, VALIDATE, PREPROC, GCD, SCP (ms, summed over frames)
  double t_metric[kNumMetrics] = {0, 0, 0, 0, 0, 0};
  if (staged && calc[kIVSSIM])
    for (int i = 0; i < 2; i++) check(qmiv_pic_create(o.ctx, w, h, bd, 444, &o.scp[i]), "picture allocation");
  auto ms_since = [](Clock::time_point a) { return std::chrono::duration<double, std::milli>(Clock::now() - a).count(); };
  // xPicP::check's report (libs/pmbb-core/src/xPic.cpp:62-82, FindOutOfRange, xPixelOpsSTD.cpp:175-197)
  // of input i as read: the first component holding a sample above the maximum, then every such sample
  // of it in raster order of the 4:4:4 picture.
  auto report = [&](int i) {
    const int32_t fmt_cf = dcf[i], bdi = bds[i];
    const size_t nb = qmiv_frame_bytes(w, h, bdi, fmt_cf);
    std::vector<uint8_t> raw(nb);
    check(qmiv_pic_download(o.ctx, o.in[i], raw.data(), nb), "picture download");
    const int32_t maxv = (1 << bdi) - 1;
    const int32_t cs = fmt_cf == 420 ? 1 : 0, cw = w >> cs;
    auto sample = [&](int c, int32_t y, int32_t x) -> int32_t {
      const size_t off = c == 0 ? (size_t)y * w + x
                                : (size_t)w * h + (size_t)(c - 1) * (size_t)cw * (h >> cs) + (size_t)(y >> cs) * cw + (x >> cs);
      return bdi <= 8 ? raw[off] : (int32_t)raw[2 * off] | ((int32_t)raw[2 * off + 1] << 8);
    };
    for (int c = 0; c < 3; c++) {
      bool broken = false;
      for (int32_t y = 0; y < h && !broken; y++)
        for (int32_t x = 0; x < w && !broken; x++) broken = sample(c, y, x) > maxv;
      if (!broken) continue;
      std::printf("FILE BROKEN %s (CMP=%d)\n", in_file[i].c_str(), c);
      for (int32_t y = 0; y < h; y++)
        for (int32_t x = 0; x < w; x++)
          if (sample(c, y, x) > maxv)
            std::printf("DETECTED out-of-range value (y=%d, x=%d, VALUE=%d, Expected=[0-%d])\n", y, x, sample(c, y, x), maxv);
      return;
    }
  };
  auto invalid = [&](const int32_t* ok, int n) {
    for (int i = 0; i < n; i++)
      if (!ok[i]) fail("ERROR --> InputFile contains invalid values (" + in_file[i] + ")");
    fail(std::string("ERROR --> ") + qmiv_last_error());
  };
  const auto prc_beg = Clock::now();
  for (int32_t f = 0; f < frames; f++) {
    if (debug) std::printf("Frame %08d  ", f);
    auto ts = Clock::now();
    for (int i = 0; i < num_inputs; i++) {
      const int32_t idx = (i < 2 ? start[i] : 0) + f;
      check(qmiv_seq_read(o.ctx, o.seq[i], idx, o.in[i]), "InputFile read error (" + in_file[i] + ")");
    }
    if (staged) check(qmiv_sync(o.ctx), "InputFile read");
    t_stage[0] += ms_since(ts);
    ts = Clock::now();
    // Validation (xAppQMIV.cpp:681-683, :753-770). The fused call checks the frame itself (its
    // report is printed below, from the pictures as read); concealment clips the pictures in place,
    // so it and the staged route check them here first.
    bool checked = false;
    int32_t in_ok[3] = {1, 1, 1};
    if (act != QMIV_PEL_SKIP && (staged || act == QMIV_PEL_CNCL)) {
      checked = true;
      for (int i = 0; i < num_inputs; i++) {
        int32_t c3[3];
        check(qmiv_pic_check(o.ctx, o.in[i], c3), "input validation");
        in_ok[i] = c3[0] && c3[1] && c3[2];
        if (!in_ok[i]) report(i);
      }
      if (act == QMIV_PEL_CNCL)
        for (int i = 0; i < num_inputs; i++)
          if (!in_ok[i]) check(qmiv_pic_conceal(o.ctx, o.in[i]), "concealment");
      if (act == QMIV_PEL_STOP && !(in_ok[0] && in_ok[1] && in_ok[2])) invalid(in_ok, num_inputs);
    }
    if (staged) t_stage[1] += ms_since(ts);
    ts = Clock::now();
    // After a failed fused call's check: the report, then STOP ends the run.
    auto after_check = [&](const int32_t* ok, int n) {
      if (checked || act == QMIV_PEL_SKIP) return;
      bool all_ok = true;
      for (int i = 0; i < n; i++)
        if (!ok[i]) {
          report(i);
          all_ok = false;
        }
      if (act == QMIV_PEL_STOP && !all_ok) invalid(ok, n);
    };
    int32_t exact[3] = {0, 0, 0};
    if (staged) {
      check(qmiv_pic_equal(o.ctx, o.in[0], o.in[1], exact), "exact comparison");
    } else if (convert) {
      qmiv_frame_result pre;
      const qmiv_status rc = qmiv_frame(o.ctx, o.in[0], o.in[1], &prm, QMIV_PSNR, &pre);
      after_check(pre.in_range, 2);
      if (rc != QMIV_ERR_PEL_RANGE) check(rc, "input validation");
      std::copy(pre.exact, pre.exact + 3, exact);
      checked = true;
    }
    qmiv_pic* p[2] = {o.in[0], o.in[1]};
    if (up444)
      for (int i = 0; i < 2; i++) {
        check(qmiv_pic_upsample(o.ctx, o.in[i], o.up[i]), "4:4:4 conversion");
        p[i] = o.up[i];
      }
    for (int i = 0; i < 2 && convert; i++) {  // xAppQMIV::preprocessFrames (xAppQMIV.cpp:767-813)
      if (cvt_ycbcr2rgb) check(qmiv_pic_colorspace(o.ctx, p[i], p[i], lc_space(csi), 1), "colour conversion");
      if (cvt_rgb2ycbcr) check(qmiv_pic_colorspace(o.ctx, p[i], p[i], lc_space(csm), 0), "colour conversion");
      if (reorder_rgb && csi == kBGR) check(qmiv_pic_swap_planes(o.ctx, p[i], 0, 2), "plane reorder");
      if (reorder_rgb && csi == kGBR) {
        check(qmiv_pic_swap_planes(o.ctx, p[i], 0, 1), "plane reorder");
        check(qmiv_pic_swap_planes(o.ctx, p[i], 0, 2), "plane reorder");
      }
    }
    double psnr[3] = {0, 0, 0}, wspsnr[3] = {0, 0, 0}, ssim[3] = {0, 0, 0}, msssim[3] = {0, 0, 0};
    double ivpsnr = 0, r2t = 0, t2r = 0, ivssim = 0, s_r2t = 0, s_t2r = 0;
    int32_t gcd[4] = {0, 0, 0, 0};
    int32_t nnm = 0;
    if (staged) {
      // xAppQMIV.cpp:686-744 stage by stage; the processors validate nothing and substitute
      // nothing, so the exact-component rule below applies as in the fused route.
      qmiv_params sp = prm;
      sp.fake_psnr_for_exact = 0;
      if (use_mask) check(qmiv_count_nonzero(o.ctx, o.in[2], &nnm), "mask count");
      check(qmiv_sync(o.ctx), "preprocessing");
      t_stage[2] += ms_since(ts);
      ts = Clock::now();
      if (need_gcd) {
        if (use_mask) check(qmiv_calc_gcd_masked(o.ctx, p[0], p[1], o.in[2], &sp, nnm, gcd), "GCD");
        else check(qmiv_calc_gcd(o.ctx, p[0], p[1], &sp, gcd), "GCD");
      }
      t_stage[3] += ms_since(ts);
      ts = Clock::now();
      if (calc[kIVSSIM]) check(qmiv_gen_scp_pics(o.ctx, p[1], p[0], &sp, gcd, o.scp[1], o.scp[0]), "SCP");
      t_stage[4] += ms_since(ts);
      for (int m = 0; m < kNumMetrics; m++) {
        if (!calc[m]) continue;
        ts = Clock::now();
        switch (m) {
          case kPSNR:
            check(use_mask ? qmiv_calc_psnr_masked(o.ctx, p[0], p[1], o.in[2], &sp, nnm, psnr)
                           : qmiv_calc_psnr(o.ctx, p[0], p[1], psnr), "PSNR");
            break;
          case kWSPSNR:
            check(use_mask ? qmiv_calc_wspsnr_masked(o.ctx, p[0], p[1], o.in[2], &sp, nnm, wspsnr)
                           : qmiv_calc_wspsnr(o.ctx, p[0], p[1], &sp, wspsnr), "WS-PSNR");
            break;
          case kIVPSNR:
            check(use_mask ? qmiv_calc_ivpsnr_masked(o.ctx, p[0], p[1], o.in[2], &sp, nnm, gcd, &ivpsnr, &r2t, &t2r)
                           : qmiv_calc_ivpsnr(o.ctx, p[0], p[1], &sp, gcd, &ivpsnr, &r2t, &t2r, nullptr), "IV-PSNR");
            break;
          case kSSIM: check(qmiv_calc_ssim(o.ctx, p[0], p[1], 1, ssim), "SSIM"); break;
          case kMSSSIM: check(qmiv_calc_msssim(o.ctx, p[0], p[1], msssim, nullptr), "MS-SSIM"); break;
          default:
            check(qmiv_calc_ivssim_scp(o.ctx, p[0], p[1], o.scp[0], o.scp[1], &sp, &ivssim, &s_r2t, &s_t2r), "IV-SSIM");
        }
        t_metric[m] += ms_since(ts);
      }
      if (debug && use_mask) std::printf("NNM %d    ", nnm);
    } else if (use_mask) {
      qmiv_masked_result r;
      const qmiv_status rc = qmiv_frame_masked(o.ctx, p[0], p[1], o.in[2], &run_prm, metrics, &r);
      after_check(r.in_range, 3);
      if (rc == QMIV_ERR_PEL_RANGE) invalid(r.in_range, 3);
      check(rc, "metric computation");
      std::copy(r.psnr, r.psnr + 3, psnr);
      std::copy(r.wspsnr, r.wspsnr + 3, wspsnr);
      std::copy(r.gcd, r.gcd + 4, gcd);
      if (!convert) std::copy(r.exact, r.exact + 3, exact);
      ivpsnr = r.ivpsnr, r2t = r.ivpsnr_r2t, t2r = r.ivpsnr_t2r;
      if (debug) std::printf("NNM %d    ", r.num_non_masked);
    } else {
      qmiv_frame_result r;
      const qmiv_status rc = qmiv_frame(o.ctx, p[0], p[1], &run_prm, metrics, &r);
      after_check(r.in_range, 2);
      if (rc == QMIV_ERR_PEL_RANGE) invalid(r.in_range, 2);
      check(rc, "metric computation");
      std::copy(r.psnr, r.psnr + 3, psnr);
      std::copy(r.wspsnr, r.wspsnr + 3, wspsnr);
      std::copy(r.ssim, r.ssim + 3, ssim);
      std::copy(r.msssim, r.msssim + 3, msssim);
      std::copy(r.gcd, r.gcd + 4, gcd);
      if (!convert) std::copy(r.exact, r.exact + 3, exact);
      ivpsnr = r.ivpsnr, r2t = r.ivpsnr_r2t, t2r = r.ivpsnr_t2r;
      ivssim = r.ivssim, s_r2t = r.ivssim_r2t, s_t2r = r.ivssim_t2r;
    }
    if (debug) {
      if (metrics & (QMIV_IVPSNR | QMIV_IVSSIM)) std::printf("GCD-R2T %d %d %d %d    ", gcd[0], gcd[1], gcd[2], gcd[3]);
      std::printf("\n");
    }
    if (convert || staged)  // xAppQMIV.cpp:837-844, :864-871
      for (int c = 0; c < 3; c++)
        if (exact[c]) psnr[c] = wspsnr[c] = fake_psnr;

    const std::string fh = format("Frame %08d ", f);
    const std::string tags = std::string(exact[0] ? " ExactY" : "") + (exact[1] ? " ExactU" : "") + (exact[2] ? " ExactV" : "");
    std::string log;
    if (calc[kPSNR]) {
      st[kPSNR].set_cmp(psnr);
      if (print_frame) log += fh + st[kPSNR].frame_cmp(f) + tags + "\n" + fh + st[kPSNR].frame_pic(f) + "\n";
    }
    if (calc[kWSPSNR]) {
      st[kWSPSNR].set_cmp(wspsnr);
      if (print_frame) log += fh + st[kWSPSNR].frame_cmp(f) + tags + "\n" + fh + st[kWSPSNR].frame_pic(f) + "\n";
    }
    if (calc[kIVPSNR]) {
      st[kIVPSNR].set_pic(ivpsnr);
      if (print_frame)
        log += fh + st[kIVPSNR].frame_pic(f) + (debug ? format("    R2T %7.4f  T2R %7.4f", r2t, t2r) : "") + "\n";
    }
    if (calc[kSSIM]) {
      st[kSSIM].set_cmp(ssim);
      if (print_frame) log += fh + st[kSSIM].frame_cmp(f) + "\n" + fh + st[kSSIM].frame_pic(f) + "\n";
    }
    if (calc[kMSSSIM]) {  // printed at every verbose level, as the reference application does (:914-920)
      st[kMSSSIM].set_cmp(msssim);
      log += fh + st[kMSSSIM].frame_cmp(f) + "\n" + fh + st[kMSSSIM].frame_pic(f) + "\n";
    }
    if (calc[kIVSSIM]) {
      st[kIVSSIM].set_pic(ivssim);
      if (print_frame)
        log += fh + st[kIVSSIM].frame_pic(f) + (debug ? format("    R2T %7.4f  T2R %7.4f", s_r2t, s_t2r) : "") + "\n";
    }
    std::fputs(log.c_str(), stdout);
  }
  const auto prc_end = Clock::now();
  const double prc_s = std::chrono::duration<double>(prc_end - prc_beg).count();

  // --- results: xAppQMIV::combineFrameStats / formatResultsFile / formatResultsStdOut ------------
  // (apps/QMIV/src/xAppQMIV.cpp:954-1030)
  if (verbose >= 1) std::printf("\n");
  std::printf("\n\n");
  for (int m = 0; m < kNumMetrics; m++)
    if (calc[m]) st[m].average(frames);
  if (!result_file.empty()) {
    std::string r = "FILE0  \"" + in_file[0] + "\"\nFILE1  \"" + in_file[1] + "\"\n";
    if (use_mask) r += "FILEM  \"" + in_file[2] + "\"\n";
    const std::time_t now = std::time(nullptr);
    char ts[64];
    std::strftime(ts, sizeof(ts), "%Y-%m-%d  %H:%M:%S", std::localtime(&now));
    r += std::string("TIME   ") + ts + "\n";
    for (int m = 0; m < kNumMetrics; m++)
      if (calc[m]) r += st[m].avg_lines("") + "\n";
    std::ofstream rf(result_file, std::ios::app);
    rf << r;
    if (!rf) fail("ERROR --> cannot write ResultFile (" + result_file + ")");
  }
  std::string out;
  for (int m = 0; m < kNumMetrics; m++)
    if (calc[m]) out += st[m].avg_lines("Average      ") + "\n";
  if (staged && frames > 0) {  // xAppQMIV::formatResultsStdOut (apps/QMIV/src/xAppQMIV.cpp:995-1030)
    const double n = (double)frames;
    const double pre = t_stage[2] / n, gcd_t = t_stage[3] / n, scp_t = t_stage[4] / n;
    out += "\n";
    out += format("AvgTime          LOAD %9.2f ms\n", t_stage[0] / n);
    out += format("AvgTime      VALIDATE %9.2f ms\n", t_stage[1] / n);
    out += format("AvgTime       PREPROC %9.2f ms\n", pre);
    if (need_gcd) out += format("AvgTime           GCD %9.2f ms\n", gcd_t);
    if (calc[kIVSSIM]) out += format("AvgTime           SCP %9.2f ms\n", scp_t);
    for (int m = 0; m < kNumMetrics; m++) {
      if (!calc[m]) continue;
      const double avg = t_metric[m] / n;
      const double before = pre + (m == kIVPSNR ? gcd_t : 0.0) + (m == kIVSSIM ? gcd_t + scp_t : 0.0);
      out += format("AvgTime      %8s %9.2f ms", kMetricName[m], avg);
      if (before > 0) out += format("   Total %7.2f ms", avg + before);
      out += "\n";
    }
  }
  std::printf("%s\n", out.c_str());
  std::printf("TotalProcessingTime  = %.3f s\n", prc_s);
  std::printf("TotalApplicationTime = %.3f s\n", std::chrono::duration<double>(Clock::now() - app_beg).count());
  std::printf("END-OF-LOG\n");
  std::fflush(stdout);
  return EXIT_SUCCESS;
}
