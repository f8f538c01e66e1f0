// ImageCodecs.h -- decoders for LDR texture files (PNG, baseline JPEG), the
// formats ImageHandler::LoadFromFile reads through its image library
// (src/tools/ImageHandler.cpp:20-90: any 8-bit image with >= 3 channels).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace OmochiRenderer {
namespace codec {

// 8-bit samples, interleaved, first row = top of the image.
struct Pixels8 {
  uint32_t width = 0, height = 0, channels = 0;
  std::vector<uint8_t> px;
};

// PNG (ISO/IEC 15948): bit depth 8, colour types 0, 2, 3, 4, 6, Adam7
// interlacing.  A palette image gives 3 channels, or 4 with a tRNS chunk.
bool decode_png(const uint8_t *data, size_t n, Pixels8 &out, std::string &err);

// JPEG (ITU-T T.81) baseline / extended sequential Huffman, 8-bit samples,
// 1 or 3 components with sampling factors 1..4, restart intervals.  YCbCr
// (JFIF) to RGB; chroma upsampled by the triangle filter.
bool decode_jpeg(const uint8_t *data, size_t n, Pixels8 &out, std::string &err);

// Picks the decoder from the file signature.
bool decode_file(const std::string &path, Pixels8 &out, std::string &err);

}  // namespace codec
}  // namespace OmochiRenderer
