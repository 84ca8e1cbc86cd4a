// PNG frame decoding for PNG-list sequences: the read step of xSeqPNG::xBackendRead
// (libs/pmbb-spng/src/xSeqPNG.cpp:91-148), which decodes one 8-bit PNG file with libspng into RGB8
// (no transparency, gamma or significant-bit processing: decode flags 0) and stores it as three
// planes R, G, B of 8-bit samples. Written from the PNG specification (ISO/IEC 15948, W3C PNG
// 2nd edition): chunk layout and CRC-32, zlib-compressed IDAT stream, the five scanline filters,
// Adam7 interlacing, and the colour types' expansion to RGB (grey replicated, palette looked up,
// alpha dropped). Inflation uses the system zlib.
//
// The reference accepts only files whose width and height equal the sequence's and whose bit
// depth is 8 (xSeqPNG.cpp:119-121); every colour type at bit depth 8 is decoded. As libspng's
// defaults do, a CRC mismatch in a critical chunk is an error and ancillary chunks are skipped;
// reading stops where the image data is complete.
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "png_reader.h"

namespace qmiv {
namespace png {
namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// Paeth predictor (PNG specification, 9.4).
inline uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return (uint8_t)a;
  return (uint8_t)(pb <= pc ? b : c);
}

// Reverses the filter of one scanline in place; prev is the previous reconstructed scanline of the
// same (sub-)image or null for its first row.
bool unfilter(uint8_t type, uint8_t* row, const uint8_t* prev, size_t len, size_t bpp) {
  switch (type) {
    case 0: return true;
    case 1:
      for (size_t i = bpp; i < len; i++) row[i] = (uint8_t)(row[i] + row[i - bpp]);
      return true;
    case 2:
      if (prev)
        for (size_t i = 0; i < len; i++) row[i] = (uint8_t)(row[i] + prev[i]);
      return true;
    case 3:
      for (size_t i = 0; i < len; i++) {
        const int a = i >= bpp ? row[i - bpp] : 0, b = prev ? prev[i] : 0;
        row[i] = (uint8_t)(row[i] + ((a + b) >> 1));
      }
      return true;
    case 4:
      for (size_t i = 0; i < len; i++) {
        const int a = i >= bpp ? row[i - bpp] : 0, b = prev ? prev[i] : 0;
        const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
        row[i] = (uint8_t)(row[i] + paeth(a, b, c));
      }
      return true;
    default: return false;
  }
}

struct Header {
  uint32_t w = 0, h = 0;
  uint8_t depth = 0, colour = 0, interlace = 0;
};

int channels(uint8_t colour) {
  switch (colour) {
    case 0: return 1;  // greyscale
    case 2: return 3;  // truecolour
    case 3: return 1;  // indexed
    case 4: return 2;  // greyscale + alpha
    case 6: return 4;  // truecolour + alpha
    default: return 0;
  }
}

// Adam7 passes (PNG specification, 8.2): starting column / row and spacing.
constexpr int kX0[7] = {0, 4, 0, 2, 0, 1, 0}, kY0[7] = {0, 0, 4, 0, 2, 0, 1};
constexpr int kDX[7] = {8, 8, 4, 4, 2, 2, 1}, kDY[7] = {8, 8, 8, 4, 4, 2, 2};

}  // namespace

bool decode_rgb8_planar(const uint8_t* data, size_t size, int32_t width, int32_t height, uint8_t* out,
                        std::string& err) {
  static const uint8_t kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (size < 8 || std::memcmp(data, kSig, 8) != 0) {
    err = "not a PNG file";
    return false;
  }
  auto crc_ok = [&](size_t at, uint32_t len) {  // chunk at `at`: CRC over type and data
    return (uint32_t)crc32(crc32(0L, Z_NULL, 0), data + at + 4, len + 4) == be32(data + at + 8 + len);
  };
  Header hd;
  bool have_ihdr = false;
  uint8_t plte[256][3];
  std::memset(plte, 0, sizeof(plte));  // entries past the palette's length read as black
  uint32_t plte_n = 0;

  // Chunks before the image data. As the reference's decoder, nothing after the image data is
  // read: IEND, later chunks and the CRC of the IDAT chunk in which the image data ends are not
  // checked.
  size_t pos = 8;
  for (;;) {
    if (size - pos < 8) {
      err = "truncated file";
      return false;
    }
    const uint32_t len = be32(data + pos);
    const uint8_t* type = data + pos + 4;
    if (len > 0x7fffffffu) {
      err = "invalid chunk length";
      return false;
    }
    if (!have_ihdr && std::memcmp(type, "IHDR", 4) != 0) {
      err = "IHDR is not the first chunk";
      return false;
    }
    if (std::memcmp(type, "IDAT", 4) == 0) break;
    if (size - pos - 8 < (size_t)len + 4) {
      err = "truncated chunk";
      return false;
    }
    const uint8_t* body = data + pos + 8;
    const bool critical = !(type[0] & 0x20);
    if (critical && !crc_ok(pos, len)) {
      err = "invalid chunk checksum";
      return false;
    }
    pos += 12 + (size_t)len;
    if (std::memcmp(type, "IHDR", 4) == 0) {
      if (have_ihdr || len != 13) {
        err = "invalid IHDR";
        return false;
      }
      hd.w = be32(body);
      hd.h = be32(body + 4);
      hd.depth = body[8];
      hd.colour = body[9];
      hd.interlace = body[12];
      if (hd.w == 0 || hd.h == 0 || hd.w > 0x7fffffffu || hd.h > 0x7fffffffu || channels(hd.colour) == 0 ||
          body[10] != 0 || body[11] != 0 || hd.interlace > 1) {
        err = "invalid IHDR";
        return false;
      }
      have_ihdr = true;
      if ((int32_t)hd.w != width) { err = "Width does not match"; return false; }
      if ((int32_t)hd.h != height) { err = "Height does not match"; return false; }
      if (hd.depth != 8) { err = "BitDepth does not match"; return false; }
    } else if (std::memcmp(type, "PLTE", 4) == 0) {
      if (len % 3 != 0 || len == 0 || len > 768) {
        err = "invalid PLTE";
        return false;
      }
      plte_n = len / 3;
      for (uint32_t i = 0; i < plte_n; i++)
        for (int c = 0; c < 3; c++) plte[i][c] = body[3 * i + c];
    } else if (std::memcmp(type, "IEND", 4) == 0) {
      err = "missing IDAT";
      return false;
    } else if (critical) {
      err = "unknown critical chunk";
      return false;
    }
  }
  if (hd.colour == 3 && plte_n == 0) {
    err = "missing PLTE";
    return false;
  }

  // Raw (filtered) size of all passes.
  const size_t ch = (size_t)channels(hd.colour), bpp = ch;
  struct Pass { size_t w, h; };
  std::vector<Pass> passes;
  if (hd.interlace == 0) {
    passes.push_back({hd.w, hd.h});
  } else {
    for (int p = 0; p < 7; p++) {
      const size_t pw = hd.w > (uint32_t)kX0[p] ? (hd.w - kX0[p] + kDX[p] - 1) / kDX[p] : 0;
      const size_t ph = hd.h > (uint32_t)kY0[p] ? (hd.h - kY0[p] + kDY[p] - 1) / kDY[p] : 0;
      passes.push_back({pw, ph});
    }
  }
  size_t raw_size = 0;
  for (const Pass& p : passes)
    if (p.w && p.h) raw_size += p.h * (1 + p.w * ch);
  std::vector<uint8_t> raw(raw_size);

  // Inflate the IDAT run chunk by chunk until the image data is complete. A chunk's CRC is checked
  // when its data is used up and more is needed.
  z_stream zs;
  std::memset(&zs, 0, sizeof(zs));
  if (inflateInit(&zs) != Z_OK) {
    err = "zlib initialisation failed";
    return false;
  }
  zs.next_out = raw.data();
  zs.avail_out = (uInt)raw.size();
  bool ok = false;
  err = "truncated image data";
  for (;;) {
    const uint32_t len = be32(data + pos);
    const size_t avail = std::min<size_t>(len, size - pos - 8);
    zs.next_in = const_cast<uint8_t*>(data + pos + 8);
    zs.avail_in = (uInt)avail;
    const int zr = inflate(&zs, Z_NO_FLUSH);
    if (zs.avail_out == 0) {
      ok = true;
      break;
    }
    if (zr != Z_OK && zr != Z_BUF_ERROR) {
      err = zr == Z_STREAM_END ? "image data too short" : "corrupt image data";
      break;
    }
    if (avail < len || size - pos - 8 - len < 4) break;  // truncated
    if (!crc_ok(pos, len)) {
      err = "invalid chunk checksum";
      break;
    }
    pos += 12 + (size_t)len;
    if (size - pos < 8 || std::memcmp(data + pos + 4, "IDAT", 4) != 0 || be32(data + pos) > 0x7fffffffu) break;
  }
  inflateEnd(&zs);
  if (!ok) return false;
  err.clear();

  const size_t n = (size_t)width * (size_t)height;
  uint8_t* dst[3] = {out, out + n, out + 2 * n};
  size_t off = 0;
  for (size_t p = 0; p < passes.size(); p++) {
    const size_t pw = passes[p].w, ph = passes[p].h, stride = pw * ch;
    if (!pw || !ph) continue;
    const uint8_t* prev = nullptr;
    for (size_t y = 0; y < ph; y++) {
      uint8_t* row = raw.data() + off + 1;
      if (!unfilter(raw[off], row, prev, stride, bpp)) {
        err = "invalid filter type";
        return false;
      }
      prev = row;
      off += 1 + stride;
      const size_t oy = hd.interlace ? kY0[p] + y * kDY[p] : y;
      for (size_t x = 0; x < pw; x++) {
        const size_t ox = hd.interlace ? kX0[p] + x * kDX[p] : x;
        const uint8_t* s = row + x * ch;
        uint8_t r, g, b;
        switch (hd.colour) {
          case 0: case 4: r = g = b = s[0]; break;
          case 3: r = plte[s[0]][0], g = plte[s[0]][1], b = plte[s[0]][2]; break;
          default: r = s[0], g = s[1], b = s[2]; break;
        }
        const size_t i = oy * (size_t)width + ox;
        dst[0][i] = r;
        dst[1][i] = g;
        dst[2][i] = b;
      }
    }
  }
  return true;
}

bool decode_file_rgb8_planar(const std::string& path, int32_t width, int32_t height, uint8_t* out,
                             std::string& err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::vector<uint8_t> buf;
  uint8_t chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  const bool read_error = std::ferror(f) != 0;
  std::fclose(f);
  if (read_error) {
    err = "read error in " + path;
    return false;
  }
  if (!decode_rgb8_planar(buf.data(), buf.size(), width, height, out, err)) {
    err += " (" + path + ")";
    return false;
  }
  return true;
}

// The reference names frame files with fmt::format(FileName, Index) (xSeqPNG.cpp:55-71): one
// replacement field per occurrence of {}, {:d}, {:Nd} or {:0Nd} (an optional argument id 0 before
// the colon), with {{ and }} as literal braces.
bool format_index(const std::string& pattern, int32_t index, std::string& out, std::string& err) {
  out.clear();
  for (size_t i = 0; i < pattern.size(); i++) {
    const char c = pattern[i];
    if (c == '{' && i + 1 < pattern.size() && pattern[i + 1] == '{') { out += '{'; i++; continue; }
    if (c == '}' && i + 1 < pattern.size() && pattern[i + 1] == '}') { out += '}'; i++; continue; }
    if (c == '}') { err = "unmatched } in the file name pattern"; return false; }
    if (c != '{') { out += c; continue; }
    const size_t close = pattern.find('}', i);
    if (close == std::string::npos) { err = "unmatched { in the file name pattern"; return false; }
    std::string spec = pattern.substr(i + 1, close - i - 1);
    if (!spec.empty() && spec[0] == '0' && (spec.size() == 1 || spec[1] == ':')) spec = spec.substr(1);
    bool zero = false;
    int wdt = 0;
    if (!spec.empty()) {
      if (spec[0] != ':') { err = "unsupported replacement field {" + spec + "}"; return false; }
      size_t k = 1;
      if (k < spec.size() && spec[k] == '0') { zero = true; k++; }
      while (k < spec.size() && spec[k] >= '0' && spec[k] <= '9') wdt = wdt * 10 + (spec[k++] - '0');
      if (k < spec.size() && spec[k] == 'd') k++;
      if (k != spec.size() || wdt > 64) { err = "unsupported replacement field {" + spec + "}"; return false; }
    }
    char num[80];
    std::snprintf(num, sizeof(num), zero ? "%0*d" : "%*d", wdt, (int)index);
    out += num;
    i = close;
  }
  return true;
}

}  // namespace png
}  // namespace qmiv
