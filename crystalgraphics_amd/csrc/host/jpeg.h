// jpeg.h — JPEG images (ITU-T T.81) to RGBA8 for glTF textures.
#ifndef CRYSTALGRAPHICS_AMD_JPEG_H_
#define CRYSTALGRAPHICS_AMD_JPEG_H_

#include <cstdint>
#include <string>
#include <vector>

#include "scene_impl.h"

namespace crystal::graphics {

// True when f starts with the SOI marker followed by another marker.
bool is_jpeg(const std::vector<uint8_t>& f);
// Sequential or progressive Huffman JPEG, 8-bit samples, one or three
// components, to RGBA8 (alpha 255); false with a message otherwise.
bool decode_jpeg(const std::vector<uint8_t>& f, HostTexture& t, std::string& err);

}  // namespace crystal::graphics

#endif  // CRYSTALGRAPHICS_AMD_JPEG_H_
