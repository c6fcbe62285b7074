// jpeg.cpp — JPEG (ITU-T T.81) to RGBA8 for glTF textures, written from the
// specification: the reference decodes images with stb_image
// (src/image_decode.cpp:23-29), which is fetched at build time and absent
// here.  Covered: sequential (SOF0, SOF1) and progressive (SOF2) Huffman
// frames with 8-bit samples; one (grey) or three (YCbCr, or RGB when an Adobe
// APP14 marker says so) components with any sampling factors 1-4; restart
// intervals; 8- and 16-bit quantisation tables.  Refused with an error:
// lossless and hierarchical frames, arithmetic coding, 12-bit samples, four
// components (CMYK).
//
// The coefficients of every scan are collected per component, then each 8x8
// block goes through the exact inverse DCT (T.81 A.3.3) in double precision,
// the level shift and rounding; chroma planes subsampled by 2 in a direction
// are upsampled with the triangle filter of the IJG library (weights 3/4 and
// 1/4 of the nearest and next samples), other factors by replication; YCbCr
// becomes RGB by the JFIF equations.  T.81 leaves upsampling and colour
// conversion to the application and bounds the IDCT only by its accuracy
// test, so decoders differ in the last units: texels can differ from
// stb_image's by a few levels.
#include "jpeg.h"

#include <cmath>
#include <cstring>

namespace crystal::graphics {
namespace {

// Zig-zag order: zz[k] is the natural (row-major) index of the k-th coefficient.
constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
  bool defined = false;
  // canonical code: for each length l (1-16), the first code, the number of
  // codes and the index of its first symbol (T.81 C, F.2.2.3)
  int32_t mincode[17] = {}, maxcode[18] = {};
  int32_t valptr[17] = {};
  uint8_t vals[256] = {};
};

struct Component {
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;          // Huffman tables of the current scan
  int bw = 0, bh = 0;          // blocks per row / column covering the component
  int bw_pad = 0, bh_pad = 0;  // blocks in the MCU grid (whole MCUs)
  std::vector<int16_t> coef;   // bh_pad * bw_pad blocks of 64 coefficients, natural order
  int pred = 0;                // DC predictor
};

class Decoder {
 public:
  Decoder(const uint8_t* d, size_t n) : d_(d), n_(n) {}
  bool run(HostTexture& t, std::string& err);

 private:
  const uint8_t* d_;
  size_t n_;
  size_t p_ = 0;
  std::string err_;
  uint16_t qt_[4][64] = {};
  bool qt_defined_[4] = {};
  Huffman dc_[4], ac_[4];
  std::vector<Component> comp_;
  int width_ = 0, height_ = 0, hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0;
  bool progressive_ = false, frame_ = false;
  int restart_ = 0;
  int adobe_transform_ = -1;
  bool jfif_ = false;
  // entropy-coded segment reader
  uint32_t bits_ = 0;
  int nbits_ = 0;
  bool marker_hit_ = false;
  int eobrun_ = 0;

  bool fail(const std::string& e) {
    if (err_.empty()) err_ = e;
    return false;
  }
  bool need(size_t k) { return p_ + k <= n_ ? true : fail("truncated JPEG data"); }
  int u8() { return d_[p_++]; }
  int u16() {
    const int v = (d_[p_] << 8) | d_[p_ + 1];
    p_ += 2;
    return v;
  }

  bool read_dqt(size_t end);
  bool read_dht(size_t end);
  bool read_sof(size_t end, int type);
  bool read_sos(size_t end);
  bool decode_scan(const std::vector<int>& sc, int ss, int se, int ah, int al);
  bool decode_block(Component& c, int16_t* blk, int ss, int se, int ah, int al);

  // bits: the next bit of the entropy-coded data; 0xFF 0x00 is a stuffed
  // 0xFF, any other marker ends the data (zeros are fed from then on)
  void fill() {
    while (nbits_ <= 24) {
      int byte = 0;
      if (!marker_hit_ && p_ < n_) {
        byte = d_[p_];
        if (byte == 0xFF) {
          const int next = p_ + 1 < n_ ? d_[p_ + 1] : 0;
          if (next == 0x00) {
            p_ += 2;
          } else {
            marker_hit_ = true;
            byte = 0;
          }
        } else {
          ++p_;
        }
      }
      bits_ |= static_cast<uint32_t>(byte) << (24 - nbits_);
      nbits_ += 8;
    }
  }
  int bit() {
    if (nbits_ == 0) fill();
    const int b = static_cast<int>(bits_ >> 31);
    bits_ <<= 1;
    --nbits_;
    return b;
  }
  int receive(int s) {  // T.81 F.2.2.4
    int v = 0;
    for (int i = 0; i < s; ++i) v = (v << 1) | bit();
    return v;
  }
  static int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }  // F.2.2.1
  int decode(const Huffman& h) {  // T.81 F.2.2.3
    int code = bit();
    for (int l = 1; l <= 16; ++l) {
      if (h.maxcode[l] >= 0 && code <= h.maxcode[l]) return h.vals[h.valptr[l] + code - h.mincode[l]];
      code = (code << 1) | bit();
    }
    return -1;
  }
  void reset_bits() {
    bits_ = 0;
    nbits_ = 0;
    marker_hit_ = false;
  }
  // at a restart interval's end: the RSTn marker, then fresh predictors
  bool restart() {
    reset_bits();
    while (p_ + 1 < n_ && !(d_[p_] == 0xFF && d_[p_ + 1] >= 0xD0 && d_[p_ + 1] <= 0xD7)) ++p_;
    if (p_ + 1 >= n_) return fail("missing restart marker");
    p_ += 2;
    for (Component& c : comp_) c.pred = 0;
    eobrun_ = 0;
    return true;
  }
  void finish(HostTexture& t);
};

bool Decoder::read_dqt(size_t end) {
  while (p_ < end) {
    const int pq = d_[p_] >> 4, tq = d_[p_] & 15;
    ++p_;
    if (tq > 3 || pq > 1) return fail("bad quantisation table");
    if (!need(pq ? 128 : 64)) return false;
    for (int k = 0; k < 64; ++k) qt_[tq][kZigzag[k]] = static_cast<uint16_t>(pq ? u16() : u8());
    qt_defined_[tq] = true;
  }
  return true;
}

bool Decoder::read_dht(size_t end) {
  while (p_ < end) {
    if (!need(17)) return false;
    const int tc = d_[p_] >> 4, th = d_[p_] & 15;
    ++p_;
    if (tc > 1 || th > 3) return fail("bad Huffman table");
    Huffman& h = tc ? ac_[th] : dc_[th];
    int counts[17] = {};
    int total = 0;
    for (int l = 1; l <= 16; ++l) total += counts[l] = u8();
    if (total > 256 || !need(static_cast<size_t>(total))) return fail("bad Huffman table");
    for (int i = 0; i < total; ++i) h.vals[i] = static_cast<uint8_t>(u8());
    int code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {  // T.81 C.2: canonical codes, length by length
      h.valptr[l] = k;
      h.mincode[l] = code;
      code += counts[l];
      k += counts[l];
      h.maxcode[l] = counts[l] ? code - 1 : -1;
      if (code > (1 << l)) return fail("bad Huffman table");
      code <<= 1;
    }
    h.maxcode[17] = -1;
    h.defined = true;
  }
  return true;
}

bool Decoder::read_sof(size_t end, int type) {
  if (frame_) return fail("more than one frame");
  if (!need(6)) return false;
  const int precision = u8();
  height_ = u16();
  width_ = u16();
  const int nc = u8();
  if (precision != 8) return fail("only 8-bit JPEG samples are supported");
  if (width_ == 0 || height_ == 0) return fail("JPEG without a size in its frame header is not supported");
  if (nc != 1 && nc != 3) return fail("JPEG with " + std::to_string(nc) + " components is not supported");
  if (!need(3 * static_cast<size_t>(nc)) || p_ + 3 * static_cast<size_t>(nc) > end) return fail("bad frame header");
  progressive_ = type == 2;
  comp_.resize(nc);
  for (Component& c : comp_) {
    c.id = u8();
    c.h = d_[p_] >> 4;
    c.v = d_[p_] & 15;
    ++p_;
    c.tq = u8();
    if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return fail("bad frame component");
    hmax_ = std::max(hmax_, c.h);
    vmax_ = std::max(vmax_, c.v);
  }
  if (uint64_t{static_cast<uint32_t>(width_)} * static_cast<uint32_t>(height_) > (uint64_t{1} << 28))
    return fail("JPEG image too large");
  mcux_ = (width_ + 8 * hmax_ - 1) / (8 * hmax_);
  mcuy_ = (height_ + 8 * vmax_ - 1) / (8 * vmax_);
  for (Component& c : comp_) {  // T.81 A.1.1: the component's size, rounded up
    const int cw = (width_ * c.h + hmax_ - 1) / hmax_, ch = (height_ * c.v + vmax_ - 1) / vmax_;
    c.bw = (cw + 7) / 8;
    c.bh = (ch + 7) / 8;
    c.bw_pad = mcux_ * c.h;
    c.bh_pad = mcuy_ * c.v;
    c.coef.assign(static_cast<size_t>(c.bw_pad) * c.bh_pad * 64, 0);
  }
  frame_ = true;
  return true;
}

bool Decoder::decode_block(Component& c, int16_t* blk, int ss, int se, int ah, int al) {
  if (!progressive_) {  // T.81 F.2.2: DC difference, then run-length coded AC
    const int t = decode(dc_[c.td]);
    if (t < 0 || t > 16) return fail("bad Huffman code");
    c.pred += extend(receive(t), t);
    blk[0] = static_cast<int16_t>(c.pred * 1);
    for (int k = 1; k < 64;) {
      const int rs = decode(ac_[c.ta]);
      if (rs < 0) return fail("bad Huffman code");
      const int r = rs >> 4, s = rs & 15;
      if (s == 0) {
        if (r != 15) break;  // end of block
        k += 16;
        continue;
      }
      k += r;
      if (k > 63) return fail("bad AC run");
      blk[kZigzag[k]] = static_cast<int16_t>(extend(receive(s), s));
      ++k;
    }
    return true;
  }
  if (ss == 0) {  // progressive DC (G.1.2.1)
    if (ah == 0) {
      const int t = decode(dc_[c.td]);
      if (t < 0 || t > 16) return fail("bad Huffman code");
      c.pred += extend(receive(t), t);
      blk[0] = static_cast<int16_t>(c.pred * (1 << al));
    } else if (bit()) {
      blk[0] = static_cast<int16_t>(blk[0] | (1 << al));
    }
    return true;
  }
  if (ah == 0) {  // progressive AC, first pass (G.1.2.2)
    if (eobrun_ > 0) {
      --eobrun_;
      return true;
    }
    for (int k = ss; k <= se;) {
      const int rs = decode(ac_[c.ta]);
      if (rs < 0) return fail("bad Huffman code");
      const int r = rs >> 4, s = rs & 15;
      if (s == 0) {
        if (r < 15) {
          eobrun_ = (1 << r) - 1;
          if (r) eobrun_ += receive(r);
          break;
        }
        k += 16;
        continue;
      }
      k += r;
      if (k > se) return fail("bad AC run");
      blk[kZigzag[k]] = static_cast<int16_t>(extend(receive(s), s) * (1 << al));
      ++k;
    }
    return true;
  }
  // progressive AC, refinement (G.1.2.3): nonzero coefficients get a
  // correction bit as they are passed; new ones (magnitude 1 << al) land
  // after the run of zero-history coefficients
  const int p1 = 1 << al, m1 = -(1 << al);
  int k = ss;
  if (eobrun_ == 0) {
    for (; k <= se;) {
      const int rs = decode(ac_[c.ta]);
      if (rs < 0) return fail("bad Huffman code");
      int r = rs >> 4;
      const int s = rs & 15;
      int value = 0;
      if (s == 0) {
        if (r < 15) {
          eobrun_ = 1 << r;
          if (r) eobrun_ += receive(r);
          break;
        }
      } else {
        if (s != 1) return fail("bad refinement value");
        value = bit() ? p1 : m1;
      }
      while (k <= se) {
        int16_t& z = blk[kZigzag[k]];
        if (z != 0) {
          if (bit() && (z & p1) == 0) z = static_cast<int16_t>(z >= 0 ? z + p1 : z + m1);
        } else {
          if (r == 0) {
            if (value) z = static_cast<int16_t>(value);
            ++k;
            break;
          }
          --r;
        }
        ++k;
      }
    }
  }
  if (eobrun_ > 0) {
    for (; k <= se; ++k) {
      int16_t& z = blk[kZigzag[k]];
      if (z != 0 && bit() && (z & p1) == 0) z = static_cast<int16_t>(z >= 0 ? z + p1 : z + m1);
    }
    --eobrun_;
  }
  return true;
}

bool Decoder::decode_scan(const std::vector<int>& sc, int ss, int se, int ah, int al) {
  reset_bits();
  eobrun_ = 0;
  for (Component& c : comp_) c.pred = 0;
  int todo = restart_;
  auto next = [&]() {
    if (restart_ == 0) return true;
    if (todo == 0) {
      if (!restart()) return false;
      todo = restart_;
    }
    --todo;
    return true;
  };
  if (sc.size() == 1) {  // non-interleaved: the component's own blocks in raster order (A.2.2)
    Component& c = comp_[static_cast<size_t>(sc[0])];
    for (int by = 0; by < c.bh; ++by)
      for (int bx = 0; bx < c.bw; ++bx) {
        if (!next()) return false;
        int16_t* blk = &c.coef[(static_cast<size_t>(by) * c.bw_pad + bx) * 64];
        if (!decode_block(c, blk, ss, se, ah, al)) return false;
      }
    return true;
  }
  for (int my = 0; my < mcuy_; ++my)  // interleaved: MCUs of h x v blocks per component (A.2.3)
    for (int mx = 0; mx < mcux_; ++mx) {
      if (!next()) return false;
      for (int ci : sc) {
        Component& c = comp_[static_cast<size_t>(ci)];
        for (int v = 0; v < c.v; ++v)
          for (int h = 0; h < c.h; ++h) {
            const int by = my * c.v + v, bx = mx * c.h + h;
            int16_t* blk = &c.coef[(static_cast<size_t>(by) * c.bw_pad + bx) * 64];
            if (!decode_block(c, blk, ss, se, ah, al)) return false;
          }
      }
    }
  return true;
}

bool Decoder::read_sos(size_t end) {
  if (!frame_) return fail("scan before the frame header");
  if (!need(1)) return false;
  const int ns = u8();
  if (ns < 1 || ns > 4 || !need(2 * static_cast<size_t>(ns) + 3)) return fail("bad scan header");
  std::vector<int> sc;
  for (int i = 0; i < ns; ++i) {
    const int id = u8();
    const int tables = u8();
    int ci = -1;
    for (size_t k = 0; k < comp_.size(); ++k)
      if (comp_[k].id == id) ci = static_cast<int>(k);
    if (ci < 0) return fail("scan names an unknown component");
    comp_[static_cast<size_t>(ci)].td = tables >> 4;
    comp_[static_cast<size_t>(ci)].ta = tables & 15;
    if ((tables >> 4) > 3 || (tables & 15) > 3) return fail("bad scan tables");
    sc.push_back(ci);
  }
  const int ss = u8(), se = u8(), a = u8();
  const int ah = a >> 4, al = a & 15;
  if (p_ != end) return fail("bad scan header length");
  if (!progressive_ && (ss != 0 || se != 63 || a != 0)) return fail("bad sequential scan");
  if (progressive_ && (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13))
    return fail("bad progressive scan");
  for (int ci : sc) {
    const Component& c = comp_[static_cast<size_t>(ci)];
    if (ss == 0 && ah == 0 && !dc_[c.td].defined) return fail("scan uses an undefined DC table");
    if (se > 0 && !ac_[c.ta].defined) return fail("scan uses an undefined AC table");
  }
  if (!decode_scan(sc, ss, se, ah, al)) return false;
  // skip to the marker that ends the entropy-coded data
  while (p_ + 1 < n_ && !(d_[p_] == 0xFF && d_[p_ + 1] != 0x00 && !(d_[p_ + 1] >= 0xD0 && d_[p_ + 1] <= 0xD7))) ++p_;
  return true;
}

// Exact 8-point inverse DCT basis (T.81 A.3.3): c(u) cos((2x + 1) u pi / 16) / 2.
struct IdctTable {
  double m[8][8];
  IdctTable() {
    for (int x = 0; x < 8; ++x)
      for (int u = 0; u < 8; ++u)
        m[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16.0) * 0.5;
  }
};

void Decoder::finish(HostTexture& t) {
  static const IdctTable T;
  // every component's samples, at its own resolution (whole blocks)
  std::vector<std::vector<uint8_t>> planes(comp_.size());
  for (size_t ci = 0; ci < comp_.size(); ++ci) {
    Component& c = comp_[ci];
    const int pw = c.bw_pad * 8;
    std::vector<uint8_t>& pl = planes[ci];
    pl.assign(static_cast<size_t>(pw) * c.bh_pad * 8, 0);
    const uint16_t* q = qt_[c.tq];
    for (int by = 0; by < c.bh_pad; ++by)
      for (int bx = 0; bx < c.bw_pad; ++bx) {
        const int16_t* blk = &c.coef[(static_cast<size_t>(by) * c.bw_pad + bx) * 64];
        double f[64], tmp[64];
        for (int i = 0; i < 64; ++i) f[i] = static_cast<double>(blk[i]) * q[i];
        for (int y = 0; y < 8; ++y)  // rows: tmp[v][x] = sum_u m[x][u] F[v][u]
          for (int x = 0; x < 8; ++x) {
            double s = 0.0;
            for (int u = 0; u < 8; ++u) s += T.m[x][u] * f[y * 8 + u];
            tmp[y * 8 + x] = s;
          }
        for (int y = 0; y < 8; ++y)  // columns: s(y, x) = sum_v m[y][v] tmp[v][x]
          for (int x = 0; x < 8; ++x) {
            double s = 0.0;
            for (int v = 0; v < 8; ++v) s += T.m[y][v] * tmp[v * 8 + x];
            const double val = std::floor(s + 128.0 + 0.5);
            pl[(static_cast<size_t>(by) * 8 + y) * pw + bx * 8 + x] =
                static_cast<uint8_t>(val < 0.0 ? 0.0 : (val > 255.0 ? 255.0 : val));
          }
      }
  }
  // sample (x, y) of the image in component ci's plane
  auto sample = [&](size_t ci, int x, int y) -> double {
    const Component& c = comp_[ci];
    const int pw = c.bw_pad * 8;
    const int cw = (width_ * c.h + hmax_ - 1) / hmax_, ch = (height_ * c.v + vmax_ - 1) / vmax_;
    const uint8_t* pl = planes[ci].data();
    auto at = [&](int sx, int sy) {
      sx = std::min(std::max(sx, 0), cw - 1);
      sy = std::min(std::max(sy, 0), ch - 1);
      return static_cast<double>(pl[static_cast<size_t>(sy) * pw + sx]);
    };
    const int fx = hmax_ / c.h, fy = vmax_ / c.v;
    const bool tri_x = fx == 2 && hmax_ % c.h == 0, tri_y = fy == 2 && vmax_ % c.v == 0;
    if (!tri_x && !tri_y) return at(x * c.h / hmax_, y * c.v / vmax_);
    // triangle filter: the nearest sample 3/4, the next one on the output
    // sample's side 1/4, in each direction subsampled by 2
    const int sx = tri_x ? x / 2 : x * c.h / hmax_, sy = tri_y ? y / 2 : y * c.v / vmax_;
    const int nx = tri_x ? (x % 2 ? sx + 1 : sx - 1) : sx, ny = tri_y ? (y % 2 ? sy + 1 : sy - 1) : sy;
    const double wx = tri_x ? 0.75 : 1.0, wy = tri_y ? 0.75 : 1.0;
    return wy * (wx * at(sx, sy) + (1.0 - wx) * at(nx, sy)) + (1.0 - wy) * (wx * at(sx, ny) + (1.0 - wx) * at(nx, ny));
  };
  t.width = static_cast<uint32_t>(width_);
  t.height = static_cast<uint32_t>(height_);
  t.rgba8.assign(static_cast<size_t>(width_) * height_ * 4, 255);
  auto clamp8 = [](double v) {
    const double r = std::floor(v + 0.5);
    return static_cast<uint8_t>(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
  };
  const bool rgb = comp_.size() == 3 && (adobe_transform_ == 0 ||
                                          (adobe_transform_ < 0 && !jfif_ && comp_[0].id == 'R' &&
                                           comp_[1].id == 'G' && comp_[2].id == 'B'));
  for (int y = 0; y < height_; ++y)
    for (int x = 0; x < width_; ++x) {
      uint8_t* o = &t.rgba8[(static_cast<size_t>(y) * width_ + x) * 4];
      if (comp_.size() == 1) {
        o[0] = o[1] = o[2] = clamp8(sample(0, x, y));
        continue;
      }
      const double Y = sample(0, x, y), cb = sample(1, x, y) - 128.0, cr = sample(2, x, y) - 128.0;
      if (rgb) {
        o[0] = clamp8(Y);
        o[1] = clamp8(cb + 128.0);
        o[2] = clamp8(cr + 128.0);
      } else {  // JFIF (ITU-R BT.601 full range)
        o[0] = clamp8(Y + 1.402 * cr);
        o[1] = clamp8(Y - 0.344136 * cb - 0.714136 * cr);
        o[2] = clamp8(Y + 1.772 * cb);
      }
    }
}

bool Decoder::run(HostTexture& t, std::string& err) {
  if (n_ < 4 || d_[0] != 0xFF || d_[1] != 0xD8) {
    err = "not a JPEG image";
    return false;
  }
  p_ = 2;
  bool scanned = false;
  for (;;) {
    while (p_ < n_ && d_[p_] != 0xFF) ++p_;  // tolerate fill before a marker
    while (p_ < n_ && d_[p_] == 0xFF) ++p_;
    if (p_ >= n_) {
      if (!scanned) fail("JPEG without image data");
      break;
    }
    const int m = d_[p_++];
    if (m == 0xD9) break;  // EOI
    if (m >= 0xD0 && m <= 0xD7) continue;
    if (!need(2)) break;
    const int len = u16();
    if (len < 2 || !need(static_cast<size_t>(len - 2))) {
      fail("bad JPEG segment length");
      break;
    }
    const size_t end = p_ + static_cast<size_t>(len - 2);
    bool ok = true;
    switch (m) {
      case 0xDB: ok = read_dqt(end); break;
      case 0xC4: ok = read_dht(end); break;
      case 0xC0: case 0xC1: ok = read_sof(end, 0); break;
      case 0xC2: ok = read_sof(end, 2); break;
      case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD: case 0xCE:
      case 0xCF:
        ok = fail("lossless, hierarchical and arithmetic-coded JPEG are not supported");
        break;
      case 0xDD:
        ok = len == 4 ? (restart_ = u16(), true) : fail("bad restart interval");
        break;
      case 0xDA:
        ok = read_sos(end);
        scanned = ok;
        break;
      case 0xE0:
        jfif_ = jfif_ || (len >= 7 && std::memcmp(d_ + p_, "JFIF", 4) == 0);
        break;
      case 0xEE:
        if (len >= 14 && std::memcmp(d_ + p_, "Adobe", 5) == 0) adobe_transform_ = d_[p_ + 11];
        break;
      default: break;  // APPn, COM and others carry nothing the pixels need
    }
    if (!ok) break;
    if (m != 0xDA) p_ = end;
  }
  if (err_.empty() && !frame_) err_ = "JPEG without a frame header";
  if (err_.empty())
    for (const Component& c : comp_)
      if (!qt_defined_[c.tq]) err_ = "JPEG component uses an undefined quantisation table";
  if (!err_.empty()) {
    err = err_;
    return false;
  }
  finish(t);
  return true;
}

}  // namespace

bool is_jpeg(const std::vector<uint8_t>& f) { return f.size() >= 3 && f[0] == 0xFF && f[1] == 0xD8 && f[2] == 0xFF; }

bool decode_jpeg(const std::vector<uint8_t>& f, HostTexture& t, std::string& err) {
  Decoder d(f.data(), f.size());
  return d.run(t, err);
}

}  // namespace crystal::graphics

// C export for the bindings and tests: decode a JPEG in memory to RGBA8.
// With out == nullptr only the size is reported; `cap` is out's length in
// bytes.  Returns 0, or 1 with the message in *err (static storage until
// the next call on this thread).
extern "C" int cgh_decode_jpeg(const uint8_t* bytes, size_t n, uint32_t* w, uint32_t* h, uint8_t* out, size_t cap,
                               const char** err) {
  static thread_local std::string msg;
  crystal::graphics::HostTexture t;
  std::vector<uint8_t> f(bytes, bytes + n);
  if (!crystal::graphics::decode_jpeg(f, t, msg)) {
    if (err) *err = msg.c_str();
    return 1;
  }
  *w = t.width;
  *h = t.height;
  if (out) {
    if (cap < t.rgba8.size()) {
      msg = "output buffer too small";
      if (err) *err = msg.c_str();
      return 1;
    }
    std::memcpy(out, t.rgba8.data(), t.rgba8.size());
  }
  return 0;
}
