// gltf.cpp — LoadScene for glTF 2.0 files (.gltf with external or data-URI
// buffers, and .glb), restating the reference's loader (src/scene.cpp:
// 829-1003 over cgltf v1.14) on this library's own JSON reader:
//   * materials: the default material at index 0, then every glTF material
//     (src/scene.cpp:869-872, LoadMaterial :708-821); a primitive's material
//     index is its glTF material + 1, or 0 (:941-950);
//   * texture slots: index, texCoord, and scale copied into both scale and
//     strength (LoadTextureInfo :635-645; occlusion's "strength" is the
//     view's scale, as cgltf stores it);
//   * vertices: POSITION, NORMAL, TEXCOORD_0/1 read as floats (normalised
//     integer components scaled as cgltf_accessor_read_float does); indices
//     rebased onto the global vertex array (:903-925);
//   * nodes: the default scene's nodes depth first (a node's primitives, then
//     its children), each primitive object with its absolute transform
//     (matrix, or T * R * S), all under a +90 degree rotation about x that
//     turns glTF's y-up into z-up (:969-1000);
//   * textures: images decoded to RGBA8 (PNG here: zlib inflate and the five
//     PNG filters; 16-bit channels keep their high byte; JPEG: jpeg.cpp,
//     sequential and progressive), samplers' wrap modes kept for the
//     emissive buffer;
//   * environment: an EXT_lights_image_based cube turned equirectangular, or
//     an HDR file named in extras/extensions (Radiance RGBE or PNG, decoded as
//     stb_image's stbi_loadf does), else the fallback sky.
// Not supported, with an error rather than a different picture: KTX2 and WebP
// images (which stb_image, the reference's decoder, does not read either),
// interlaced PNGs, sparse accessors, non-triangle primitives.  An environment HDR file in another
// format leaves the fallback sky, as an undecodable one does in the reference.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "jpeg.h"
#include "json.h"
#include "scene_impl.h"

namespace crystal::graphics {
namespace {

using json::Value;

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  std::error_code ec;
  if (!std::filesystem::is_regular_file(path, ec)) return false;
  const uintmax_t n = std::filesystem::file_size(path, ec);
  if (ec || n > (uintmax_t{1} << 34)) return false;
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out.resize(static_cast<size_t>(n));
  if (n > 0) f.read(reinterpret_cast<char*>(out.data()), static_cast<std::streamsize>(n));
  return static_cast<bool>(f);
}

std::string dir_of(const std::string& path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

int hexv(char c) {
  if (c >= '0' && c <= '9') return c - '0';
  if (c >= 'a' && c <= 'f') return c - 'a' + 10;
  if (c >= 'A' && c <= 'F') return c - 'A' + 10;
  return -1;
}
std::string percent_decode(const std::string& s) {
  std::string out;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '%' && i + 2 < s.size() && hexv(s[i + 1]) >= 0 && hexv(s[i + 2]) >= 0) {
      out.push_back(static_cast<char>(hexv(s[i + 1]) * 16 + hexv(s[i + 2])));
      i += 2;
    } else {
      out.push_back(s[i]);
    }
  }
  return out;
}
bool base64_decode(const std::string& s, std::vector<uint8_t>& out) {
  int acc = 0, bits = -8;
  for (char c : s) {
    if (c == '=') break;
    if (c == ' ' || c == '\n' || c == '\r' || c == '\t') continue;
    int d = (c >= 'A' && c <= 'Z') ? c - 'A'
            : (c >= 'a' && c <= 'z') ? c - 'a' + 26
            : (c >= '0' && c <= '9') ? c - '0' + 52
            : c == '+' ? 62
            : c == '/' ? 63 : -1;
    if (d < 0) return false;
    acc = (acc << 6) | d;
    bits += 6;
    if (bits >= 0) {
      out.push_back(static_cast<uint8_t>((acc >> bits) & 0xFF));
      bits -= 8;
    }
  }
  return true;
}
// data: URIs (base64 or percent-encoded) and paths relative to the file.
bool load_uri(const std::string& uri, const std::string& base, std::vector<uint8_t>& out, std::string& mime,
              std::string& err) {
  if (uri.rfind("data:", 0) == 0) {
    const size_t comma = uri.find(',');
    if (comma == std::string::npos) return err = "invalid data URI", false;
    const std::string header = uri.substr(5, comma - 5);
    mime = header.substr(0, header.find(';'));
    const std::string payload = uri.substr(comma + 1);
    if (header.find(";base64") != std::string::npos) {
      if (!base64_decode(payload, out)) return err = "invalid base64 data URI", false;
    } else {
      const std::string d = percent_decode(payload);
      out.assign(d.begin(), d.end());
    }
    return true;
  }
  const std::string rel = percent_decode(uri);
  const std::string path = (!rel.empty() && rel[0] == '/') ? rel : base + rel;
  if (!read_file(path, out)) return err = "cannot read " + path, false;
  return true;
}

// ---------------------------------------------------------------------------
// PNG to RGBA8.
uint32_t be32(const uint8_t* p) {
  return (uint32_t{p[0]} << 24) | (uint32_t{p[1]} << 16) | (uint32_t{p[2]} << 8) | uint32_t{p[3]};
}
bool decode_png(const std::vector<uint8_t>& f, HostTexture& t, std::string& err) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return err = "not a PNG image", false;
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  size_t p = 8;
  while (p + 12 <= f.size()) {
    const uint32_t len = be32(&f[p]);
    if (p + 12 + static_cast<size_t>(len) > f.size()) return err = "truncated PNG chunk", false;
    const std::string type(reinterpret_cast<const char*>(&f[p + 4]), 4);
    const uint8_t* d = &f[p + 8];
    if (type == "IHDR" && len >= 13) {
      w = be32(d);
      h = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      interlace = d[12];
    } else if (type == "PLTE") {
      plte.assign(d, d + len);
    } else if (type == "tRNS") {
      trns.assign(d, d + len);
    } else if (type == "IDAT") {
      idat.insert(idat.end(), d, d + len);
    } else if (type == "IEND") {
      break;
    }
    p += 12 + static_cast<size_t>(len);
  }
  if (w == 0 || h == 0 || w > 16384 || h > 16384 || static_cast<uint64_t>(w) * h > (uint64_t{1} << 26))
    return err = "bad PNG size (at most 16384 on a side, 2^26 texels)", false;
  if (interlace != 0) return err = "interlaced PNG images are not supported", false;
  int ch;
  switch (ctype) {
    case 0: ch = 1; break;
    case 2: ch = 3; break;
    case 3: ch = 1; break;
    case 4: ch = 2; break;
    case 6: ch = 4; break;
    default: return err = "bad PNG colour type", false;
  }
  const bool ok_depth = ctype == 3 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8)
                        : ctype == 0 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)
                                     : (depth == 8 || depth == 16);
  if (!ok_depth) return err = "bad PNG bit depth", false;
  const size_t row = (static_cast<size_t>(w) * ch * depth + 7) / 8;
  const size_t bpp = std::max<size_t>(1, static_cast<size_t>(ch) * depth / 8);
  std::vector<uint8_t> raw((row + 1) * h);
  uLongf raw_len = static_cast<uLongf>(raw.size());
  if (uncompress(raw.data(), &raw_len, idat.data(), static_cast<uLong>(idat.size())) != Z_OK || raw_len != raw.size())
    return err = "corrupt PNG image data", false;
  // undo the per-row filters (PNG 1.2 section 6)
  std::vector<uint8_t> img(row * h);
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t ft = raw[y * (row + 1)];
    const uint8_t* src = &raw[y * (row + 1) + 1];
    uint8_t* cur = &img[y * row];
    const uint8_t* prev = y ? &img[(y - 1) * row] : nullptr;
    for (size_t x = 0; x < row; ++x) {
      const int a = x >= bpp ? cur[x - bpp] : 0;
      const int b = prev ? prev[x] : 0;
      const int c = (prev && x >= bpp) ? prev[x - bpp] : 0;
      int v = src[x];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) / 2; break;
        case 4: {
          const int pa = std::abs(b - c), pb = std::abs(a - c), pc = std::abs(a + b - 2 * c);
          v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
          break;
        }
        default: return err = "bad PNG filter", false;
      }
      cur[x] = static_cast<uint8_t>(v);
    }
  }
  t.width = w;
  t.height = h;
  t.rgba8.assign(static_cast<size_t>(w) * h * 4, 255);
  auto sample = [&](const uint8_t* r, uint32_t x, int c) -> uint32_t {  // channel c of pixel x, at `depth` bits
    if (depth == 8) return r[x * ch + c];
    if (depth == 16) return r[(x * ch + c) * 2];  // high byte
    const size_t bit = static_cast<size_t>(x) * depth;
    return (r[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1u);
  };
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t* r = &img[y * row];
    for (uint32_t x = 0; x < w; ++x) {
      uint8_t* o = &t.rgba8[(static_cast<size_t>(y) * w + x) * 4];
      if (ctype == 3) {
        const uint32_t i = sample(r, x, 0);
        if (3 * i + 2 >= plte.size()) return err = "PNG palette index out of range", false;
        o[0] = plte[3 * i];
        o[1] = plte[3 * i + 1];
        o[2] = plte[3 * i + 2];
        o[3] = i < trns.size() ? trns[i] : 255;
      } else if (ctype == 0 || ctype == 4) {
        uint32_t g = sample(r, x, 0);
        if (depth < 8) g = g * 255u / ((1u << depth) - 1u);
        o[0] = o[1] = o[2] = static_cast<uint8_t>(g);
        if (ctype == 4) o[3] = static_cast<uint8_t>(sample(r, x, 1));
      } else {
        for (int c = 0; c < ch; ++c) o[c] = static_cast<uint8_t>(sample(r, x, c));
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Four-channel float images, as stb_image's stbi_loadf(..., 4) returns them
// (the reference's DecodeImageRgba32f, src/image_decode.cpp:51-116):
//   * Radiance RGBE (.hdr): a "#?RADIANCE" or "#?RGBE" line, header lines up
//     to an empty one (one must read FORMAT=32-bit_rle_rgbe), "-Y <h> +X <w>",
//     then rows top first, flat RGBE quadruples when w < 8 or w >= 32768, else
//     new-style run-length rows (2, 2, w as two bytes, then each of the four
//     components as runs: a count above 128 repeats the next byte count - 128
//     times, otherwise count literal bytes follow).  A row that does not
//     start with that marker makes those four bytes pixel 0 of row 0 and the
//     rest of the data flat pixels from pixel 1 of row 0 on (stb's reading).
//     A pixel is c * 2^(e - 136) per colour channel (zero when e = 0), alpha 1;
//   * PNG: the RGBA8 image, colour channels as powf(v / 255, 2.2), alpha v / 255.
// Truncated data is an error here (stb would return unset pixels).
struct FloatImage {
  uint32_t w = 0, h = 0;
  std::vector<float> rgba;
};

void rgbe_pixel(const uint8_t* q, float* o) {
  if (q[3] != 0) {
    const float f1 = std::ldexp(1.0f, static_cast<int>(q[3]) - 136);
    o[0] = static_cast<float>(q[0]) * f1;
    o[1] = static_cast<float>(q[1]) * f1;
    o[2] = static_cast<float>(q[2]) * f1;
  } else {
    o[0] = o[1] = o[2] = 0.0f;
  }
  o[3] = 1.0f;
}

bool is_hdr(const std::vector<uint8_t>& f) {
  auto starts = [&](const char* sig) {
    const size_t n = std::strlen(sig);
    return f.size() >= n && std::memcmp(f.data(), sig, n) == 0;
  };
  return starts("#?RADIANCE\n") || starts("#?RGBE\n");
}

bool decode_hdr(const std::vector<uint8_t>& f, FloatImage& img, std::string& err) {
  size_t p = 0;
  // A header line without its newline; long lines keep their first 1023 bytes.
  auto line = [&]() {
    std::string t;
    while (p < f.size() && f[p] != '\n') {
      if (t.size() < 1023) t.push_back(static_cast<char>(f[p]));
      ++p;
    }
    if (p < f.size()) ++p;
    return t;
  };
  const std::string magic = line();
  if (magic != "#?RADIANCE" && magic != "#?RGBE") return err = "not a Radiance HDR image", false;
  bool format = false;
  for (;;) {
    if (p >= f.size()) return err = "truncated HDR header", false;
    const std::string t = line();
    if (t.empty()) break;
    if (t == "FORMAT=32-bit_rle_rgbe") format = true;
  }
  if (!format) return err = "unsupported HDR format", false;
  const std::string res = line();
  if (res.compare(0, 3, "-Y ") != 0) return err = "unsupported HDR data layout", false;
  const char* c = res.c_str() + 3;
  char* e = nullptr;
  const long h = std::strtol(c, &e, 10);
  c = e;
  while (*c == ' ') ++c;
  if (std::strncmp(c, "+X ", 3) != 0) return err = "unsupported HDR data layout", false;
  const long w = std::strtol(c + 3, nullptr, 10);
  if (w <= 0 || h <= 0 || w > (1L << 24) || h > (1L << 24) || w * h > (1L << 27))  // at most 2 GiB of floats
    return err = "bad HDR image size", false;
  img.w = static_cast<uint32_t>(w);
  img.h = static_cast<uint32_t>(h);
  img.rgba.assign(static_cast<size_t>(w) * static_cast<size_t>(h) * 4u, 0.0f);
  const size_t n = static_cast<size_t>(w) * static_cast<size_t>(h);
  auto flat_from = [&](size_t first) {
    if ((n - first) * 4u > f.size() - p) return err = "truncated HDR data", false;
    for (size_t i = first; i < n; ++i, p += 4) rgbe_pixel(&f[p], &img.rgba[i * 4u]);
    return true;
  };
  if (w < 8 || w >= 32768) return flat_from(0);
  std::vector<uint8_t> row(static_cast<size_t>(w) * 4u);
  for (long y = 0; y < h; ++y) {
    if (f.size() - p < 4) return err = "truncated HDR data", false;
    if (f[p] != 2 || f[p + 1] != 2 || (f[p + 2] & 0x80)) {
      rgbe_pixel(&f[p], &img.rgba[0]);
      p += 4;
      return flat_from(1);
    }
    if (((static_cast<long>(f[p + 2]) << 8) | f[p + 3]) != w) return err = "invalid decoded HDR scanline length", false;
    p += 4;
    for (int k = 0; k < 4; ++k) {
      long x = 0;
      while (x < w) {
        if (p >= f.size()) return err = "truncated HDR data", false;
        int count = f[p++];
        const bool run = count > 128;
        if (run) count -= 128;
        if (count == 0 || count > w - x) return err = "bad RLE data in HDR", false;
        if (run) {
          if (p >= f.size()) return err = "truncated HDR data", false;
          const uint8_t v = f[p++];
          for (int z = 0; z < count; ++z) row[static_cast<size_t>(x++) * 4u + static_cast<size_t>(k)] = v;
        } else {
          if (f.size() - p < static_cast<size_t>(count)) return err = "truncated HDR data", false;
          for (int z = 0; z < count; ++z) row[static_cast<size_t>(x++) * 4u + static_cast<size_t>(k)] = f[p++];
        }
      }
    }
    for (long x = 0; x < w; ++x)
      rgbe_pixel(&row[static_cast<size_t>(x) * 4u], &img.rgba[(static_cast<size_t>(y) * static_cast<size_t>(w) +
                                                              static_cast<size_t>(x)) * 4u]);
  }
  return true;
}

// An 8-bit image (PNG or JPEG) to RGBA8.
bool decode_image(const std::vector<uint8_t>& f, HostTexture& t, std::string& err) {
  if (is_jpeg(f)) return decode_jpeg(f, t, err);
  return decode_png(f, t, err);
}

bool decode_float_image(const std::vector<uint8_t>& f, FloatImage& img, std::string& err) {
  if (is_hdr(f)) return decode_hdr(f, img, err);
  HostTexture t;
  if (!decode_image(f, t, err)) return false;
  img.w = t.width;
  img.h = t.height;
  img.rgba.resize(t.rgba8.size());
  for (size_t i = 0; i < t.rgba8.size(); ++i)
    img.rgba[i] = (i & 3u) == 3u ? static_cast<float>(t.rgba8[i]) / 255.0f
                                 : std::pow(static_cast<float>(t.rgba8[i]) / 255.0f, 2.2f) * 1.0f;
  return true;
}

// ---------------------------------------------------------------------------
// An EXT_lights_image_based cube as an equirectangular environment texture
// (src/scene.cpp:461-583): 4 w x 2 h texels for w x h faces; texel (x, y)
// looks along the engine direction of polar angle pi (y + 1/2) / (2 h) from
// +z and azimuth 2 pi ((x + 1/2) / (4 w) - 1/2), taken to glTF axes
// (x, -z, y), turned by the inverse of the light's rotation quaternion, and
// reads the nearest texel of the face it falls on.  Single-precision
// arithmetic in the reference's order of operations.
struct V3f {
  float x, y, z;
};
V3f cross3(V3f a, V3f b) { return V3f{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
V3f rotate(const float q[4], V3f d) {
  V3f im{q[0], q[1], q[2]};
  float re = q[3];
  const float norm = ((im.x * im.x + im.y * im.y) + im.z * im.z) + re * re;
  if (norm <= 0.0f) return d;
  const float sn = std::sqrt(norm);
  im = V3f{im.x / sn, im.y / sn, im.z / sn};
  re = re / sn;
  const V3f c = cross3(im, d);
  const V3f t{2.0f * c.x, 2.0f * c.y, 2.0f * c.z};
  const V3f u = cross3(im, t);
  return V3f{(d.x + re * t.x) + u.x, (d.y + re * t.y) + u.y, (d.z + re * t.z) + u.z};
}
const float* cube_texel(const FloatImage& img, float u, float v) {
  u = std::clamp(u, 0.0f, 1.0f);
  v = std::clamp(v, 0.0f, 1.0f);
  const uint32_t x = std::min(img.w - 1u, static_cast<uint32_t>(u * static_cast<float>(img.w)));
  const uint32_t y = std::min(img.h - 1u, static_cast<uint32_t>(v * static_cast<float>(img.h)));
  return &img.rgba[(static_cast<size_t>(y) * img.w + x) * 4u];
}
const float* sample_cube(const FloatImage* faces, V3f d) {
  const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  int face;
  float u, v;
  if (ax >= ay && ax >= az) {
    face = d.x >= 0.0f ? 1 : 0;
    u = d.x >= 0.0f ? (-d.z / ax + 1.0f) * 0.5f : (d.z / ax + 1.0f) * 0.5f;
    v = (-d.y / ax + 1.0f) * 0.5f;
  } else if (ay >= az) {
    face = d.y >= 0.0f ? 2 : 3;
    u = (d.x / ay + 1.0f) * 0.5f;
    v = d.y >= 0.0f ? (d.z / ay + 1.0f) * 0.5f : (-d.z / ay + 1.0f) * 0.5f;
  } else {
    face = d.z >= 0.0f ? 4 : 5;
    u = d.z >= 0.0f ? (d.x / az + 1.0f) * 0.5f : (-d.x / az + 1.0f) * 0.5f;
    v = (1.0f - d.y / az) * 0.5f;
  }
  return cube_texel(faces[face], 1.0f - u, 1.0f - v);
}
void cube_to_env(const FloatImage* faces, const float rot[4], Scene::Impl& s) {
  const float pi = 3.14159265358979323846f;
  const float inv[4] = {-rot[0], -rot[1], -rot[2], rot[3]};
  s.env_w = faces[0].w * 4u;
  s.env_h = faces[0].h * 2u;
  s.env_rgba32f.assign(static_cast<size_t>(s.env_w) * s.env_h * 4u, 0.0f);
  for (uint32_t y = 0; y < s.env_h; ++y) {
    const float v = (static_cast<float>(y) + 0.5f) / static_cast<float>(s.env_h);
    const float theta = v * pi;
    const float sin_theta = std::sin(theta), z = std::cos(theta);
    for (uint32_t x = 0; x < s.env_w; ++x) {
      const float u = (static_cast<float>(x) + 0.5f) / static_cast<float>(s.env_w);
      const float phi = ((u - 0.5f) * 2.0f) * pi;
      const V3f eng{std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, z};
      const float* c = sample_cube(faces, rotate(inv, V3f{eng.x, -eng.z, eng.y}));
      float* o = &s.env_rgba32f[(static_cast<size_t>(y) * s.env_w + x) * 4u];
      o[0] = c[0];
      o[1] = c[1];
      o[2] = c[2];
      o[3] = 1.0f;
    }
  }
}

// ---------------------------------------------------------------------------
struct Doc {
  Value root;
  std::vector<std::vector<uint8_t>> buffers;
  std::string base;
};

// Integers from the file, clamped to [-1, 2^31 - 1] (negative means absent or invalid).
int as_int(const Value* v, int dflt) {
  if (!v || !v->is_num()) return dflt;
  const double x = v->num;
  if (!(x >= 0.0)) return -1;
  return x >= 2147483647.0 ? 2147483647 : static_cast<int>(x);
}
float as_float(const Value* v, float dflt) { return v && v->is_num() ? static_cast<float>(v->num) : dflt; }
const Value* arr_at(const Value& root, const char* key, int i) {
  const Value* a = root.get(key);
  return (a && i >= 0) ? a->at(static_cast<size_t>(i)) : nullptr;
}

// Raw bytes of a buffer view: pointer and length.
bool view_bytes(const Doc& d, int view, const uint8_t*& ptr, size_t& len, size_t& stride, std::string& err) {
  const Value* bv = arr_at(d.root, "bufferViews", view);
  if (!bv) return err = "bad bufferView index", false;
  const int b = as_int(bv->get("buffer"), -1);
  if (b < 0 || static_cast<size_t>(b) >= d.buffers.size()) return err = "bad buffer index", false;
  const int o = as_int(bv->get("byteOffset"), 0), l = as_int(bv->get("byteLength"), 0),
            st = as_int(bv->get("byteStride"), 0);
  if (o < 0 || l < 0 || st < 0 || st > 252) return err = "bad bufferView", false;
  const size_t off = static_cast<size_t>(o);
  len = static_cast<size_t>(l);
  stride = static_cast<size_t>(st);
  if (off > d.buffers[b].size() || len > d.buffers[b].size() - off) return err = "bufferView past its buffer", false;
  ptr = d.buffers[b].data() + off;
  return true;
}

struct Accessor {
  const uint8_t* data = nullptr;  // null: all zeros
  size_t count = 0, stride = 0;
  int comp = 5126, ncomp = 1;
  bool normalized = false;
};
int comp_size(int c) { return c == 5120 || c == 5121 ? 1 : c == 5122 || c == 5123 ? 2 : 4; }
bool accessor(const Doc& d, int index, Accessor& a, std::string& err) {
  const Value* v = arr_at(d.root, "accessors", index);
  if (!v) return err = "bad accessor index", false;
  if (v->get("sparse")) return err = "sparse accessors are not supported", false;
  a.comp = as_int(v->get("componentType"), 5126);
  if (a.comp != 5120 && a.comp != 5121 && a.comp != 5122 && a.comp != 5123 && a.comp != 5125 && a.comp != 5126)
    return err = "bad accessor componentType", false;
  const int cnt = as_int(v->get("count"), 0);
  if (cnt < 0 || cnt > (1 << 28)) return err = "bad accessor count", false;
  a.count = static_cast<size_t>(cnt);
  a.normalized = v->get("normalized") && v->get("normalized")->b;
  const Value* type = v->get("type");
  const std::string t = type ? type->str : "SCALAR";
  a.ncomp = t == "VEC2" ? 2 : t == "VEC3" ? 3 : t == "VEC4" ? 4 : t == "MAT4" ? 16 : t == "MAT3" ? 9 : t == "MAT2" ? 4 : 1;
  const size_t elem = static_cast<size_t>(comp_size(a.comp)) * a.ncomp;
  if (const Value* bvi = v->get("bufferView")) {
    const uint8_t* p;
    size_t len, stride;
    if (!view_bytes(d, as_int(bvi, -1), p, len, stride, err)) return false;
    const int o = as_int(v->get("byteOffset"), 0);
    if (o < 0) return err = "bad accessor byteOffset", false;
    const size_t off = static_cast<size_t>(o);
    a.stride = stride ? stride : elem;
    if (a.count && (off > len || (a.count - 1) * a.stride + elem > len - off))
      return err = "accessor past its bufferView", false;
    a.data = p + off;
  } else {
    a.stride = elem;
  }
  return true;
}
// cgltf_accessor_read_float: normalised integers map to [0, 1] / [-1, 1].
float read_comp(const Accessor& a, size_t i, int c) {
  if (!a.data) return 0.0f;
  const uint8_t* p = a.data + i * a.stride + static_cast<size_t>(c) * comp_size(a.comp);
  switch (a.comp) {
    case 5126: { float f; std::memcpy(&f, p, 4); return f; }
    case 5121: return a.normalized ? static_cast<float>(p[0]) / 255.0f : static_cast<float>(p[0]);
    case 5120: {
      const float v = static_cast<float>(static_cast<int8_t>(p[0]));
      return a.normalized ? std::max(v / 127.0f, -1.0f) : v;
    }
    case 5123: {
      uint16_t u; std::memcpy(&u, p, 2);
      return a.normalized ? static_cast<float>(u) / 65535.0f : static_cast<float>(u);
    }
    case 5122: {
      int16_t s; std::memcpy(&s, p, 2);
      return a.normalized ? std::max(static_cast<float>(s) / 32767.0f, -1.0f) : static_cast<float>(s);
    }
    case 5125: { uint32_t u; std::memcpy(&u, p, 4); return static_cast<float>(u); }
    default: return 0.0f;
  }
}
uint32_t read_index(const Accessor& a, size_t i) {
  if (!a.data) return 0u;
  const uint8_t* p = a.data + i * a.stride;
  switch (a.comp) {
    case 5121: return p[0];
    case 5123: { uint16_t u; std::memcpy(&u, p, 2); return u; }
    case 5125: { uint32_t u; std::memcpy(&u, p, 4); return u; }
    default: return 0u;
  }
}

// An image's encoded bytes: its buffer view, or its URI (src/scene.cpp:170-200).
bool image_bytes(const Doc& d, const Value& img, std::vector<uint8_t>& bytes, std::string& mime, std::string& err) {
  mime = img.get("mimeType") ? img.get("mimeType")->str : "";
  if (const Value* bv = img.get("bufferView")) {
    const uint8_t* p;
    size_t len, stride;
    if (!view_bytes(d, as_int(bv, -1), p, len, stride, err)) return false;
    bytes.assign(p, p + len);
    return true;
  }
  if (const Value* uri = img.get("uri")) return load_uri(uri->str, d.base, bytes, mime, err);
  return err = "image has no source", false;
}

cg_tex_info tex_info(const Value* view, const char* scale_key) {  // LoadTextureInfo
  cg_tex_info t{-1, 0, 1.0f, 1.0f};
  if (!view || !view->get("index")) return t;
  t.index = as_int(view->get("index"), -1);
  t.tex_coord = static_cast<uint32_t>(std::max(as_int(view->get("texCoord"), 0), 0));
  t.scale = t.strength = scale_key ? as_float(view->get(scale_key), 1.0f) : 1.0f;
  return t;
}
void floats(const Value* a, float* out, int n) {
  for (int i = 0; a && i < n; ++i) out[i] = as_float(a->at(static_cast<size_t>(i)), out[i]);
}

cg_material load_material(const Value& m) {  // src/scene.cpp:708-821
  cg_material r = DefaultMaterial();
  if (const Value* pbr = m.get("pbrMetallicRoughness")) {
    floats(pbr->get("baseColorFactor"), r.base_color_factor, 4);
    r.metallic_factor = as_float(pbr->get("metallicFactor"), 1.0f);
    r.roughness_factor = as_float(pbr->get("roughnessFactor"), 1.0f);
    r.tex[CG_TEX_BASE_COLOR] = tex_info(pbr->get("baseColorTexture"), nullptr);
    r.tex[CG_TEX_METALLIC_ROUGHNESS] = tex_info(pbr->get("metallicRoughnessTexture"), nullptr);
  }
  floats(m.get("emissiveFactor"), r.emissive_factor, 3);
  const Value* am = m.get("alphaMode");
  r.alpha_mode = !am ? 0u : am->str == "MASK" ? 1u : am->str == "BLEND" ? 2u : 0u;
  r.alpha_cutoff = as_float(m.get("alphaCutoff"), 0.5f);
  if (m.get("doubleSided") && m.get("doubleSided")->b) r.flags |= 1u;
  r.tex[CG_TEX_NORMAL] = tex_info(m.get("normalTexture"), "scale");
  r.tex[CG_TEX_OCCLUSION] = tex_info(m.get("occlusionTexture"), "strength");
  r.tex[CG_TEX_EMISSIVE] = tex_info(m.get("emissiveTexture"), nullptr);
  const Value* ext = m.get("extensions");
  if (!ext) return r;
  if (ext->get("KHR_materials_unlit")) r.flags |= 2u;
  if (const Value* e = ext->get("KHR_materials_emissive_strength"))
    r.emissive_strength = as_float(e->get("emissiveStrength"), 1.0f);
  if (const Value* e = ext->get("KHR_materials_transmission")) {
    r.transmission_factor = as_float(e->get("transmissionFactor"), 0.0f);
    r.tex[CG_TEX_TRANSMISSION] = tex_info(e->get("transmissionTexture"), nullptr);
  }
  if (const Value* e = ext->get("KHR_materials_ior")) r.ior = as_float(e->get("ior"), 1.5f);
  if (const Value* e = ext->get("KHR_materials_dispersion")) r.dispersion = as_float(e->get("dispersion"), 0.0f);
  if (const Value* e = ext->get("KHR_materials_volume")) {
    r.thickness_factor = as_float(e->get("thicknessFactor"), 0.0f);
    r.attenuation_color[0] = r.attenuation_color[1] = r.attenuation_color[2] = 1.0f;
    floats(e->get("attenuationColor"), r.attenuation_color, 3);
    r.attenuation_distance = as_float(e->get("attenuationDistance"), std::numeric_limits<float>::max());
    r.tex[16] = tex_info(e->get("thicknessTexture"), nullptr);
  }
  if (const Value* e = ext->get("KHR_materials_clearcoat")) {
    r.clearcoat_factor = as_float(e->get("clearcoatFactor"), 0.0f);
    r.clearcoat_roughness_factor = as_float(e->get("clearcoatRoughnessFactor"), 0.0f);
    r.tex[6] = tex_info(e->get("clearcoatTexture"), nullptr);
    r.tex[7] = tex_info(e->get("clearcoatRoughnessTexture"), nullptr);
    r.tex[8] = tex_info(e->get("clearcoatNormalTexture"), "scale");
  }
  if (const Value* e = ext->get("KHR_materials_iridescence")) {
    r.iridescence_factor = as_float(e->get("iridescenceFactor"), 0.0f);
    r.iridescence_ior = as_float(e->get("iridescenceIor"), 1.3f);
    r.iridescence_thickness_minimum = as_float(e->get("iridescenceThicknessMinimum"), 100.0f);
    r.iridescence_thickness_maximum = as_float(e->get("iridescenceThicknessMaximum"), 400.0f);
    r.tex[9] = tex_info(e->get("iridescenceTexture"), nullptr);
    r.tex[10] = tex_info(e->get("iridescenceThicknessTexture"), nullptr);
  }
  if (const Value* e = ext->get("KHR_materials_sheen")) {
    r.sheen_color_factor[0] = r.sheen_color_factor[1] = r.sheen_color_factor[2] = 0.0f;
    floats(e->get("sheenColorFactor"), r.sheen_color_factor, 3);
    r.sheen_roughness_factor = as_float(e->get("sheenRoughnessFactor"), 0.0f);
    r.tex[11] = tex_info(e->get("sheenColorTexture"), nullptr);
    r.tex[12] = tex_info(e->get("sheenRoughnessTexture"), nullptr);
  }
  if (const Value* e = ext->get("KHR_materials_specular")) {
    r.specular_color_factor[0] = r.specular_color_factor[1] = r.specular_color_factor[2] = 1.0f;
    floats(e->get("specularColorFactor"), r.specular_color_factor, 3);
    r.specular_factor = as_float(e->get("specularFactor"), 1.0f);
    r.tex[13] = tex_info(e->get("specularTexture"), nullptr);
    r.tex[14] = tex_info(e->get("specularColorTexture"), nullptr);
  }
  if (const Value* e = ext->get("KHR_materials_anisotropy")) {
    r.anisotropy_strength = as_float(e->get("anisotropyStrength"), 0.0f);
    r.anisotropy_rotation = as_float(e->get("anisotropyRotation"), 0.0f);
    r.tex[5] = tex_info(e->get("anisotropyTexture"), nullptr);
  }
  return r;
}

// Column-major 4x4 in double.
struct M4d {
  double m[16];
};
M4d mul(const M4d& a, const M4d& b) {
  M4d r{};
  for (int c = 0; c < 4; ++c)
    for (int rr = 0; rr < 4; ++rr) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a.m[k * 4 + rr] * b.m[c * 4 + k];
      r.m[c * 4 + rr] = s;
    }
  return r;
}
M4d identity() {
  M4d r{};
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
  return r;
}
// A node's local transform: its matrix, else T * R * S (glTF 2.0 section 3.5.3).
M4d node_local(const Value& n) {
  M4d r = identity();
  if (const Value* mat = n.get("matrix")) {
    for (int i = 0; i < 16; ++i) r.m[i] = mat->at(static_cast<size_t>(i)) ? mat->at(static_cast<size_t>(i))->num : r.m[i];
    return r;
  }
  double t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
  if (const Value* v = n.get("translation"))
    for (int i = 0; i < 3; ++i) t[i] = v->at(i) ? v->at(i)->num : t[i];
  if (const Value* v = n.get("rotation"))
    for (int i = 0; i < 4; ++i) q[i] = v->at(i) ? v->at(i)->num : q[i];
  if (const Value* v = n.get("scale"))
    for (int i = 0; i < 3; ++i) s[i] = v->at(i) ? v->at(i)->num : s[i];
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  r.m[0] = (1 - 2 * (y * y + z * z)) * s[0];
  r.m[1] = (2 * (x * y + z * w)) * s[0];
  r.m[2] = (2 * (x * z - y * w)) * s[0];
  r.m[4] = (2 * (x * y - z * w)) * s[1];
  r.m[5] = (1 - 2 * (x * x + z * z)) * s[1];
  r.m[6] = (2 * (y * z + x * w)) * s[1];
  r.m[8] = (2 * (x * z + y * w)) * s[2];
  r.m[9] = (2 * (y * z - x * w)) * s[2];
  r.m[10] = (1 - 2 * (x * x + y * y)) * s[2];
  r.m[12] = t[0];
  r.m[13] = t[1];
  r.m[14] = t[2];
  return r;
}
// Inverse of an affine transform (upper 3x3 by cofactors, then translation).
M4d affine_inverse(const M4d& a) {
  const double* m = a.m;
  const double c00 = m[5] * m[10] - m[9] * m[6], c01 = m[9] * m[2] - m[1] * m[10], c02 = m[1] * m[6] - m[5] * m[2];
  const double c10 = m[8] * m[6] - m[4] * m[10], c11 = m[0] * m[10] - m[8] * m[2], c12 = m[4] * m[2] - m[0] * m[6];
  const double c20 = m[4] * m[9] - m[8] * m[5], c21 = m[8] * m[1] - m[0] * m[9], c22 = m[0] * m[5] - m[4] * m[1];
  const double det = m[0] * c00 + m[4] * c01 + m[8] * c02;
  const double id = det != 0.0 ? 1.0 / det : 0.0;
  M4d r = identity();
  r.m[0] = c00 * id; r.m[1] = c01 * id; r.m[2] = c02 * id;
  r.m[4] = c10 * id; r.m[5] = c11 * id; r.m[6] = c12 * id;
  r.m[8] = c20 * id; r.m[9] = c21 * id; r.m[10] = c22 * id;
  for (int i = 0; i < 3; ++i) r.m[12 + i] = -(r.m[i] * m[12] + r.m[4 + i] * m[13] + r.m[8 + i] * m[14]);
  return r;
}

// The first member named `key` in document order at any depth whose value is
// a string (want_string) or an array.
const Value* find_member(const Value& v, const char* key, bool want_string) {
  if (v.kind == Value::kObject) {
    for (const auto& m : v.obj) {
      if (m.first == key && m.second.kind == (want_string ? Value::kString : Value::kArray)) return &m.second;
      if (const Value* r = find_member(m.second, key, want_string)) return r;
    }
  } else if (v.kind == Value::kArray) {
    for (const Value& e : v.arr)
      if (const Value* r = find_member(e, key, want_string)) return r;
  }
  return nullptr;
}

// The environment HDR file (src/scene.cpp:307-359): the sources in order are
// the file's extras, each of its extensions, then the scene's extras and each
// of its extensions; within a source the keys in the order below, each
// taking its first string value; an empty value passes to the next key.  The
// path is percent-decoded and taken relative to the glTF file's folder.
std::string environment_path(const Value& root, const Value& scene, const std::string& base) {
  static const char* keys[] = {"environment_hdr", "environmentHdr", "environmentHDR", "environmentMap", "hdr", "hdrFile"};
  std::vector<const Value*> sources;
  // cgltf keeps the root extensions it parses itself out of the list.
  auto add = [&](const Value& owner, bool is_root) {
    if (const Value* x = owner.get("extras")) sources.push_back(x);
    if (const Value* e = owner.get("extensions"))
      if (e->kind == Value::kObject)
        for (const auto& m : e->obj)
          if (!is_root || (m.first != "KHR_lights_punctual" && m.first != "KHR_materials_variants"))
            sources.push_back(&m.second);
  };
  add(root, true);
  add(scene, false);
  for (const Value* src : sources)
    for (const char* k : keys) {
      const Value* v = find_member(*src, k, true);
      if (!v || v->str.empty()) continue;
      const std::string path = percent_decode(v->str);
      return (!path.empty() && path[0] == '/') ? path : base + path;
    }
  return "";
}

}  // namespace

namespace {
std::string load_gltf(const std::string& path, Scene::Impl& s);
}  // namespace

std::string LoadGltf(const std::string& path, Scene::Impl& s) {
  try {
    return load_gltf(path, s);
  } catch (const std::exception& e) {
    return std::string("cannot load: ") + e.what();
  }
}

namespace {
std::string load_gltf(const std::string& path, Scene::Impl& s) {
  std::vector<uint8_t> file;
  if (!read_file(path, file)) return "cannot read " + path;
  Doc d;
  d.base = dir_of(path);
  std::string json_text;
  std::vector<uint8_t> glb_bin;
  bool glb = file.size() >= 12 && std::memcmp(file.data(), "glTF", 4) == 0;
  if (glb) {  // GLB container: header, JSON chunk, optional BIN chunk
    size_t p = 12;
    while (p + 8 <= file.size()) {
      uint32_t len, type;
      std::memcpy(&len, &file[p], 4);
      std::memcpy(&type, &file[p + 4], 4);
      if (p + 8 + static_cast<size_t>(len) > file.size()) return "truncated GLB chunk";
      if (type == 0x4E4F534Au) json_text.assign(reinterpret_cast<const char*>(&file[p + 8]), len);
      if (type == 0x004E4942u) glb_bin.assign(file.begin() + static_cast<long>(p + 8), file.begin() + static_cast<long>(p + 8 + len));
      p += 8 + static_cast<size_t>(len);
    }
  } else {
    json_text.assign(file.begin(), file.end());
  }
  json::Parser parser(json_text);
  if (!parser.parse(d.root)) return parser.error();
  const Value* scenes = d.root.get("scenes");
  const Value* scene = scenes ? scenes->at(static_cast<size_t>(as_int(d.root.get("scene"), 0))) : nullptr;
  if (!scene) return "glTF file has no scene";
  std::string err;

  // buffers
  if (const Value* bufs = d.root.get("buffers")) {
    for (size_t i = 0; i < bufs->size(); ++i) {
      const Value* b = bufs->at(i);
      const Value* uri = b->get("uri");
      std::vector<uint8_t> data;
      if (!uri) {
        if (!glb || i != 0) return "buffer " + std::to_string(i) + " has no data";
        data = glb_bin;
      } else {
        std::string mime;
        if (!load_uri(uri->str, d.base, data, mime, err)) return err;
      }
      if (data.size() < static_cast<size_t>(as_int(b->get("byteLength"), 0))) return "buffer shorter than byteLength";
      d.buffers.push_back(std::move(data));
    }
  }

  s.vertices.clear();
  s.indices.clear();
  s.prims.clear();
  s.textures.clear();
  s.env_rgba32f.clear();
  s.env_w = s.env_h = 0;

  // environment (src/scene.cpp:586-624, 853-859; scene_data.h:52-69): an
  // EXT_lights_image_based cube (its first specular image set and rotation)
  // wins; else the HDR file named in extras or extensions, decoded now; an HDR
  // file that cannot be read or decoded leaves the fallback sky, as the
  // reference's scene data does.
  const std::string hdr_path = environment_path(d.root, *scene, d.base);
  const Value* root_ext = d.root.get("extensions");
  if (const Value* ibl = root_ext ? root_ext->get("EXT_lights_image_based") : nullptr) {
    const Value* sets = find_member(*ibl, "specularImages", false);
    const Value* first = sets ? sets->at(0) : nullptr;
    int idx[6];
    bool ok = first && first->size() >= 6;
    for (int i = 0; ok && i < 6; ++i) {
      const Value* v = first->at(static_cast<size_t>(i));
      ok = v->is_num() && v->num >= 0.0 && v->num < 2147483647.0 && v->num == std::floor(v->num) &&
           v->str.find_first_not_of("0123456789") == std::string::npos;
      if (ok) idx[i] = static_cast<int>(v->num);
    }
    if (ok) {
      float rot[4] = {0.0f, 0.0f, 0.0f, 1.0f};
      const Value* r = find_member(*ibl, "rotation", false);
      if (r && r->size() >= 4 && r->at(0)->is_num() && r->at(1)->is_num() && r->at(2)->is_num() && r->at(3)->is_num())
        for (int i = 0; i < 4; ++i) rot[i] = std::stof(r->at(static_cast<size_t>(i))->str);
      FloatImage faces[6];
      for (int i = 0; i < 6; ++i) {
        const Value* img = arr_at(d.root, "images", idx[i]);
        if (!img) return "image based light references an invalid image";
        std::vector<uint8_t> bytes;
        std::string mime;
        if (!image_bytes(d, *img, bytes, mime, err)) return "image based light: " + err;
        if (!decode_float_image(bytes, faces[i], err)) return "failed to decode image based light face as float: " + err;
        if (faces[i].w == 0 || faces[i].h == 0) return "image based light has an empty cube face";
        if (i > 0 && (faces[i].w != faces[0].w || faces[i].h != faces[0].h))
          return "image based light cube faces do not share one size";
      }
      if (uint64_t{faces[0].w} * faces[0].h * 8u > (uint64_t{1} << 27))  // the 4w x 2h texture: at most 2 GiB
        return "image based light cube faces are too large";
      cube_to_env(faces, rot, s);
    }
  }
  if (s.env_rgba32f.empty() && !hdr_path.empty()) {
    std::vector<uint8_t> bytes;
    FloatImage env;
    std::string ignored;
    if (read_file(hdr_path, bytes) && decode_float_image(bytes, env, ignored) && env.w && env.h) {
      s.env_w = env.w;
      s.env_h = env.h;
      s.env_rgba32f = std::move(env.rgba);
    }
  }

  // textures (src/scene.cpp:196-230)
  if (const Value* texs = d.root.get("textures")) {
    for (size_t i = 0; i < texs->size(); ++i) {
      const Value* t = texs->at(i);
      HostTexture ht;
      if (const Value* smp = arr_at(d.root, "samplers", as_int(t->get("sampler"), -1))) {
        ht.wrap_s = as_int(smp->get("wrapS"), CG_WRAP_REPEAT);
        ht.wrap_t = as_int(smp->get("wrapT"), CG_WRAP_REPEAT);
      }
      const Value* img = arr_at(d.root, "images", as_int(t->get("source"), -1));
      if (!img) return "texture " + std::to_string(i) + " has no image source";
      std::vector<uint8_t> bytes;
      std::string mime;
      if (!image_bytes(d, *img, bytes, mime, err)) return "texture " + std::to_string(i) + ": " + err;
      if (!decode_image(bytes, ht, err))
        return "texture " + std::to_string(i) + ": " + err + (mime.empty() ? "" : " (" + mime + ")");
      s.textures.push_back(std::move(ht));
    }
  }

  // materials: the default first, then every glTF material
  s.materials.assign(1, DefaultMaterial());
  if (const Value* mats = d.root.get("materials"))
    for (size_t i = 0; i < mats->size(); ++i) s.materials.push_back(load_material(*mats->at(i)));

  // meshes -> primitive records over the global vertex and index arrays
  std::vector<std::vector<cg_primitive>> mesh_prims;
  if (const Value* meshes = d.root.get("meshes")) {
    for (size_t mi = 0; mi < meshes->size(); ++mi) {
      mesh_prims.emplace_back();
      const Value* ps = meshes->at(mi)->get("primitives");
      for (size_t pi = 0; ps && pi < ps->size(); ++pi) {
        const Value& pr = *ps->at(pi);
        if (as_int(pr.get("mode"), 4) != 4) return "only triangle primitives are supported";
        const Value* attrs = pr.get("attributes");
        if (!attrs || !attrs->get("POSITION")) return "primitive without POSITION";
        Accessor pos;
        if (!accessor(d, as_int(attrs->get("POSITION"), -1), pos, err)) return err;
        const uint32_t vo = static_cast<uint32_t>(s.vertices.size());
        s.vertices.resize(vo + pos.count, cg_vertex{});
        auto fill = [&](const char* name, int n, size_t field_off) -> bool {
          const Value* a = attrs->get(name);
          if (!a) return true;
          Accessor acc;
          if (!accessor(d, as_int(a, -1), acc, err)) return false;
          for (size_t v = 0; v < acc.count && v < pos.count; ++v) {
            float* dst = reinterpret_cast<float*>(reinterpret_cast<char*>(&s.vertices[vo + v]) + field_off);
            for (int c = 0; c < n; ++c) dst[c] = c < acc.ncomp ? read_comp(acc, v, c) : 0.0f;
          }
          return true;
        };
        if (!fill("POSITION", 3, offsetof(cg_vertex, pos)) || !fill("NORMAL", 3, offsetof(cg_vertex, nrm)) ||
            !fill("TEXCOORD_0", 2, offsetof(cg_vertex, uv0)) || !fill("TEXCOORD_1", 2, offsetof(cg_vertex, uv1)))
          return err;
        cg_primitive p{};
        p.vertex_offset = vo;
        p.vertex_count = static_cast<uint32_t>(pos.count);
        p.index_offset = static_cast<uint32_t>(s.indices.size());
        if (const Value* ia = pr.get("indices")) {
          Accessor idx;
          if (!accessor(d, as_int(ia, -1), idx, err)) return err;
          for (size_t k = 0; k < idx.count; ++k) {
            const uint32_t li = read_index(idx, k);
            if (li >= pos.count) return "index past the primitive's vertices";
            s.indices.push_back(li + vo);
          }
          p.index_count = static_cast<uint32_t>(idx.count);
        }
        const Value* mat = pr.get("material");
        if (mat && as_int(mat, -1) < 0) return "bad primitive material";
        p.material_idx = mat ? static_cast<uint32_t>(as_int(mat, -1) + 1) : 0u;
        if (p.material_idx >= s.materials.size()) return "primitive material index out of range";
        mesh_prims.back().push_back(p);
      }
    }
  }

  // nodes of the default scene, depth first, under the y-up -> z-up rotation
  M4d root = identity();  // +90 degrees about x: (x, y, z) -> (x, -z, y)
  root.m[5] = 0.0;
  root.m[6] = 1.0;
  root.m[9] = -1.0;
  root.m[10] = 0.0;
  const Value* nodes = d.root.get("nodes");
  int visited = 0;
  std::function<bool(int, const M4d&, int)> walk = [&](int ni, const M4d& parent, int depth) -> bool {
    const Value* n = (nodes && ni >= 0) ? nodes->at(static_cast<size_t>(ni)) : nullptr;
    if (!n || ++visited > 1000000 || depth > 256) return err = "bad node index, node cycle or nesting too deep", false;
    const M4d world = mul(parent, node_local(*n));
    if (const Value* mesh = n->get("mesh")) {
      const int mi = as_int(mesh, -1);
      if (mi < 0 || static_cast<size_t>(mi) >= mesh_prims.size()) return err = "bad mesh index", false;
      const M4d inv = affine_inverse(world);
      for (cg_primitive p : mesh_prims[static_cast<size_t>(mi)]) {
        for (int i = 0; i < 16; ++i) {
          p.world[i] = static_cast<float>(world.m[i]);
          p.inv_world[i] = static_cast<float>(inv.m[i]);
        }
        s.prims.push_back(p);
      }
    }
    if (const Value* ch = n->get("children"))
      for (size_t i = 0; i < ch->size(); ++i)
        if (!walk(as_int(ch->at(i), -1), world, depth + 1)) return false;
    return true;
  };
  if (const Value* sn = scene->get("nodes"))
    for (size_t i = 0; i < sn->size(); ++i)
      if (!walk(as_int(sn->at(i), -1), root, 0)) return err;

  // a suggested view: the scene's bounds from -y, slightly above
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  for (const auto& p : s.prims)
    for (uint32_t v = p.vertex_offset; v < p.vertex_offset + p.vertex_count; ++v)
      for (int a = 0; a < 3; ++a) {
        const float* m = p.world;
        const float* q = s.vertices[v].pos;
        const float w = m[a] * q[0] + m[4 + a] * q[1] + m[8 + a] * q[2] + m[12 + a];
        lo[a] = std::min(lo[a], w);
        hi[a] = std::max(hi[a], w);
      }
  if (lo[0] > hi[0]) lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0.0f;
  const float c[3] = {0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])};
  const float ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-3f});
  const float eye[3] = {c[0], c[1] - 1.8f * ext, c[2] + 0.6f * ext};
  float dir[3] = {c[0] - eye[0], c[1] - eye[1], c[2] - eye[2]};
  const float dl = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  s.camera = cg_camera{};
  for (int a = 0; a < 3; ++a) {
    s.camera.position[a] = eye[a];
    s.camera.direction[a] = dir[a] / dl;
  }
  s.camera.viewport[0] = 0.5f * 16.0f / 9.0f;
  s.camera.viewport[1] = 0.5f;
  s.conf = cg_conf{1920, 1080, 640, 360, 4, 4, 3, 1, 1, 1};
  // the tag names the content (the facade caches a scene's prepared data per tag)
  uint64_t h = 0xCBF29CE484222325ull;
  auto mix = [&](const uint8_t* p, size_t n) {
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 0x100000001B3ull;
  };
  mix(file.data(), file.size());
  for (const auto& b : d.buffers) mix(b.data(), b.size());
  for (const auto& t : s.textures) mix(t.rgba8.data(), t.rgba8.size());
  mix(reinterpret_cast<const uint8_t*>(s.env_rgba32f.data()), s.env_rgba32f.size() * sizeof(float));
  s.tag = h ^ std::hash<std::string>{}(path);
  return "";
}
}  // namespace

}  // namespace crystal::graphics
