// PNG decoding for PNG-list sequences (host code).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace qmiv {
namespace png {

// Decodes one PNG image of exactly width x height at bit depth 8 into three planes R, G, B of
// 8-bit samples (out holds 3 * width * height bytes). false with a message in err otherwise.
bool decode_rgb8_planar(const uint8_t* data, size_t size, int32_t width, int32_t height, uint8_t* out,
                        std::string& err);
bool decode_file_rgb8_planar(const std::string& path, int32_t width, int32_t height, uint8_t* out,
                             std::string& err);
// fmt-style frame file name: pattern with {} / {:Nd} / {:0Nd} fields replaced by index.
bool format_index(const std::string& pattern, int32_t index, std::string& out, std::string& err);

}  // namespace png
}  // namespace qmiv
