// cg_oracle.cpp — ORACLE: test infrastructure only (see cg_oracle.h).
//
// A scalar restatement of the reference's spectral path tracer, written to be
// read side by side with the WGSL it follows.  Each function cites the
// reference file:line it restates (paths relative to the reference root;
// WGSL paths relative to src/wgpu/shader/).  It is compiled with
// -ffp-contract=off so every float operation is the single IEEE operation the
// WGSL writes.  The product must never link or call this file.
#include "cg_oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Vector helpers with declared evaluation order (left to right).
struct v3 {
  float x, y, z;
};
struct f2 {
  float x, y;
};
inline v3 mk(float x, float y, float z) { return v3{x, y, z}; }
inline v3 mk(const float* p) { return v3{p[0], p[1], p[2]}; }
inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
inline v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline f2 operator+(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
inline f2 operator*(float s, f2 a) { return {s * a.x, s * a.y}; }
inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline v3 cross(v3 a, v3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(v3 a) { return std::sqrt(dot(a, a)); }
inline v3 normalize(v3 a) {
  float l = length(a);
  return {a.x / l, a.y / l, a.z / l};
}
inline float fmn(float a, float b) { return std::fmin(a, b); }
inline float fmx(float a, float b) { return std::fmax(a, b); }
inline float clampf(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }
inline v3 vmin(v3 a, v3 b) { return {fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
inline v3 vmax(v3 a, v3 b) { return {fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }
inline float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(x);
}
// WGSL reflect/refract/mix (declared op order).
inline v3 reflect(v3 e1, v3 e2) {
  float s = 2.0f * dot(e2, e1);
  return e1 - e2 * s;
}
inline v3 refract(v3 e1, v3 e2, float e3) {
  float d = dot(e2, e1);
  float k = 1.0f - e3 * e3 * (1.0f - d * d);
  if (k < 0.0f) return mk(0.0f, 0.0f, 0.0f);
  return e1 * e3 - e2 * (e3 * d + std::sqrt(k));
}
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// Column-major 4x4 (glm / WGSL mat4x4f): col j = m[4j .. 4j+3].
inline v3 mat_point(const float* m, v3 p) {  // (M * vec4(p, 1)).xyz
  v3 c0 = mk(m + 0), c1 = mk(m + 4), c2 = mk(m + 8), c3 = mk(m + 12);
  return ((c0 * p.x + c1 * p.y) + c2 * p.z) + c3 * 1.0f;
}
inline v3 mat_dir(const float* m, v3 d) {  // (M * vec4(d, 0)).xyz
  v3 c0 = mk(m + 0), c1 = mk(m + 4), c2 = mk(m + 8), c3 = mk(m + 12);
  return ((c0 * d.x + c1 * d.y) + c2 * d.z) + c3 * 0.0f;
}
// glm's mat4 * vec4(p, 1), pairwise: (m0*x + m1*y) + (m2*z + m3*1).  Used by
// the host-side emissive buffer (emissive_buffer.h:124-127); glm itself is
// not available, so this operation order is declared, not pinned.
inline v3 glm_point(const float* m, v3 p) {
  v3 c0 = mk(m + 0), c1 = mk(m + 4), c2 = mk(m + 8), c3 = mk(m + 12);
  return (c0 * p.x + c1 * p.y) + (c2 * p.z + c3 * 1.0f);
}
inline v3 mat_t_dir(const float* m, v3 n) {  // (transpose(M) * vec4(n, 0)).xyz
  float r[3];
  for (int i = 0; i < 3; ++i) {
    const float* c = m + 4 * i;
    r[i] = ((c[0] * n.x + c[1] * n.y) + c[2] * n.z) + c[3] * 0.0f;
  }
  return mk(r[0], r[1], r[2]);
}

// ---------------------------------------------------------------------------
// math/rng.wesl:1-21 — iterated PCG-RXS-M-XS hash on a 32-bit state.
inline void step_rng(uint32_t& s) {
  s = s * 747796405u + 2891336453u;
  uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  s = (word >> 22u) ^ word;
}
inline float rand_float(uint32_t& s) {
  step_rng(s);
  return static_cast<float>(s) / 4294967296.0f;
}
inline uint32_t rand_uint(uint32_t& s) {
  step_rng(s);
  return s;
}

constexpr float PI = 3.1415926535f;  // math/constant.wesl:1
constexpr float MAXD = 1e6f;         // math/constant.wesl:2

// ---------------------------------------------------------------------------
// physics/cie.wesl:1-140 — Wyman-Sloan-Shirley multi-lobe fit and its erf
// integral.
constexpr float SQRT_HALF_PI = 1.2533141373155001f;
constexpr float INV_SQRT_TWO = 0.7071067811865475f;

float erf_approx(float x) {  // cie.wesl:132-140 (Abramowitz-Stegun 7.1.26)
  float sign = x >= 0.0f ? 1.0f : -1.0f;
  float ax = std::fabs(x);
  float t = 1.0f / (1.0f + 0.3275911f * ax);
  float poly = (((((1.061405429f * t - 1.453152027f) * t) + 1.421413741f) * t - 0.284496736f) * t +
                0.254829592f) *
               t;
  return sign * (1.0f - poly * std::exp(-ax * ax));
}
float split_gauss(float l, float c, float lw, float rw) {  // cie.wesl:95-104
  float w = l < c ? lw : rw;
  float s = (l - c) / w;
  return std::exp(-0.5f * s * s);
}
float split_gauss_anti(float l, float c, float lw, float rw) {  // cie.wesl:117-130
  if (l < c) return SQRT_HALF_PI * lw * (1.0f + erf_approx((l - c) * INV_SQRT_TWO / lw));
  return SQRT_HALF_PI * lw + SQRT_HALF_PI * rw * erf_approx((l - c) * INV_SQRT_TWO / rw);
}
float split_gauss_int(float lo, float hi, float c, float lw, float rw) {  // cie.wesl:106-115
  return split_gauss_anti(hi, c, lw, rw) - split_gauss_anti(lo, c, lw, rw);
}
// Lobe constants (cie.wesl:16-30, 45-92).  Widths are the f32 quotients 1/w.
struct Lobe {
  float a, c, lw, rw;
};
const Lobe XL[3] = {{0.362f, 442.0f, 1.0f / 0.0624f, 1.0f / 0.0374f},
                    {1.056f, 599.8f, 1.0f / 0.0264f, 1.0f / 0.0323f},
                    {0.065f, 501.1f, 1.0f / 0.0490f, 1.0f / 0.0382f}};
const Lobe YL[2] = {{0.821f, 568.8f, 1.0f / 0.0213f, 1.0f / 0.0247f},
                    {0.286f, 530.9f, 1.0f / 0.0613f, 1.0f / 0.0322f}};
const Lobe ZL[2] = {{1.217f, 437.0f, 1.0f / 0.0845f, 1.0f / 0.0278f},
                    {0.681f, 459.0f, 1.0f / 0.0385f, 1.0f / 0.0725f}};

v3 xyz_match(float l) {  // cie.wesl:4-30
  float x = XL[0].a * split_gauss(l, XL[0].c, XL[0].lw, XL[0].rw) +
            XL[1].a * split_gauss(l, XL[1].c, XL[1].lw, XL[1].rw) -
            XL[2].a * split_gauss(l, XL[2].c, XL[2].lw, XL[2].rw);
  float y = YL[0].a * split_gauss(l, YL[0].c, YL[0].lw, YL[0].rw) +
            YL[1].a * split_gauss(l, YL[1].c, YL[1].lw, YL[1].rw);
  float z = ZL[0].a * split_gauss(l, ZL[0].c, ZL[0].lw, ZL[0].rw) +
            ZL[1].a * split_gauss(l, ZL[1].c, ZL[1].lw, ZL[1].rw);
  return mk(x, y, z);
}
float xyz_match_sum(float l) {  // cie.wesl:32-35
  v3 m = xyz_match(l);
  return m.x + m.y + m.z;
}
float xyz_match_sum_integral(float lo, float hi) {  // cie.wesl:37-92
  float x = XL[0].a * split_gauss_int(lo, hi, XL[0].c, XL[0].lw, XL[0].rw) +
            XL[1].a * split_gauss_int(lo, hi, XL[1].c, XL[1].lw, XL[1].rw) -
            XL[2].a * split_gauss_int(lo, hi, XL[2].c, XL[2].lw, XL[2].rw);
  float y = YL[0].a * split_gauss_int(lo, hi, YL[0].c, YL[0].lw, YL[0].rw) +
            YL[1].a * split_gauss_int(lo, hi, YL[1].c, YL[1].lw, YL[1].rw);
  float z = ZL[0].a * split_gauss_int(lo, hi, ZL[0].c, ZL[0].lw, ZL[0].rw) +
            ZL[1].a * split_gauss_int(lo, hi, ZL[1].c, ZL[1].lw, ZL[1].rw);
  return x + y + z;
}

// physics/wavelength.wesl:20-45 — 16-step bisection of the CMF-sum CDF.
void sample_wavelength_cdf(float u, float& lambda, float& prob) {
  float total = xyz_match_sum_integral(380.0f, 720.0f);
  float target = clampf(u, 0.0f, 0.999999f);
  float low = 380.0f, high = 720.0f;
  for (int i = 0; i < 16; ++i) {
    float mid = 0.5f * (low + high);
    float cdf = xyz_match_sum_integral(380.0f, mid) / total;
    if (cdf < target)
      low = mid;
    else
      high = mid;
  }
  lambda = 0.5f * (low + high);
  prob = fmx(xyz_match_sum(lambda) / total, 1e-6f);
}
// physics/luminance.wesl:13-16
v3 illuminate(float lambda, float radiance) {
  float avg = xyz_match_sum_integral(380.0f, 720.0f) / 3.0f;
  v3 m = xyz_match(lambda);
  return mk(m.x * radiance / avg, m.y * radiance / avg, m.z * radiance / avg);
}

// ---------------------------------------------------------------------------
// shading/smits.wesl:3-124 — Smits (1999) RGB->spectrum basis, with the
// reference's 10-bin index and robustness-clamped upper neighbour.
const float BINS[10] = {380.0f,    417.7778f, 455.5556f, 493.3333f, 531.1111f,
                        568.8889f, 606.6667f, 644.4444f, 682.2222f, 720.0f};
const float SMITS[7][10] = {
    {1.0f, 1.0f, 0.9999f, 0.9993f, 0.9992f, 0.9998f, 1.0f, 1.0f, 1.0f, 1.0f},              // white
    {0.1012f, 0.0515f, 0.0f, 0.0f, 0.0f, 0.0f, 0.8325f, 1.0149f, 1.0149f, 1.0149f},        // red
    {0.0f, 0.0f, 0.0273f, 0.7937f, 1.0f, 0.9418f, 0.1719f, 0.0f, 0.0f, 0.0025f},           // green
    {1.0f, 1.0f, 0.8916f, 0.3323f, 0.0f, 0.0f, 0.0003f, 0.0369f, 0.0483f, 0.0496f},        // blue
    {1.0f, 1.0f, 0.9685f, 0.2229f, 0.0f, 0.0458f, 0.8369f, 1.0f, 1.0f, 0.9959f},           // magenta
    {0.971f, 0.9426f, 1.0007f, 1.0007f, 1.0007f, 1.0007f, 0.1564f, 0.0f, 0.0f, 0.0f},      // cyan
    {0.0001f, 0.0f, 0.1088f, 0.6651f, 1.0f, 1.0f, 0.9996f, 0.9586f, 0.9685f, 0.984f}};     // yellow
enum { WHITE = 0, RED, GREEN, BLUE, MAGENTA, CYAN, YELLOW };

float smits(int curve, float l) {  // smits.wesl:14-27 (same body per curve)
  const float* T = SMITS[curve];
  if (l <= 380.0f) return T[0];
  if (l >= 720.0f) return T[9];
  uint32_t bin = f2u(10.0f * (l - 380.0f) / 340.0f);
  uint32_t b0 = std::min(bin, 9u), b1 = std::min(bin + 1u, 9u);
  float mix = (l - BINS[b0]) / (BINS[b1] - BINS[b0]);
  return (1.0f - mix) * T[b0] + mix * T[b1];
}
// shading/material.wesl:264-297 (reflectance) and environment.wesl:27-61 /
// pdf/emission.wesl:84-117 (radiance): same branch structure, including the
// first branch's unreachable inner else.
float rgb_lift(v3 rgb, float l) {
  float s = 0.0f;
  float r = rgb.x, g = rgb.y, b = rgb.z;
  if (r <= g && g <= b) {
    s += r * smits(WHITE, l);
    if (g <= b) {
      s += (g - r) * smits(CYAN, l);
      s += (b - g) * smits(BLUE, l);
    } else {
      s += (b - r) * smits(CYAN, l);
      s += (g - b) * smits(GREEN, l);
    }
  } else if (g <= r && g <= b) {
    s += g * smits(WHITE, l);
    if (r <= b) {
      s += (r - g) * smits(MAGENTA, l);
      s += (b - r) * smits(BLUE, l);
    } else {
      s += (b - g) * smits(MAGENTA, l);
      s += (r - b) * smits(RED, l);
    }
  } else {
    s += b * smits(WHITE, l);
    if (r <= g) {
      s += (r - b) * smits(YELLOW, l);
      s += (g - r) * smits(GREEN, l);
    } else {
      s += (g - b) * smits(YELLOW, l);
      s += (r - g) * smits(RED, l);
    }
  }
  return s;
}
float rgb_to_reflectance(v3 rgb, float l) { return clampf(0.94f * rgb_lift(rgb, l), 0.0f, 1.0f); }
float rgb_to_radiance(v3 rgb, float l) { return fmx(rgb_lift(rgb, l), 0.0f); }

// ---------------------------------------------------------------------------
// shading/material.wesl:11-110 — material decode at one wavelength.
struct IOR {
  float re, im;
};
struct Mat {
  float diffuse, alpha;
  IOR ior;
  float transmission;
  bool volume;
  float emission;
};
float dispersed_ior(float nd, float disp, float l) {  // material.wesl:80-97
  if (disp <= 0.0f) return nd;
  float inv_abbe = disp / 20.0f;
  return fmx(nd + (nd - 1.0f) * inv_abbe * (523655.0f / (l * l) - 1.5168f), 1.0f);
}
float imag_ior(float n, float c) {  // material.wesl:99-110
  float np1 = n + 1.0f;
  return std::sqrt(clampf((4.0f * n) / (1.0f - c) - np1 * np1, 0.0f, 10.0f));
}
// The material texture array as WriteMaterialTextures uploads it
// (src/wgpu/resource.cpp:471-546); layers == 0 is the 1x1 white default.
struct TexArr {
  const uint8_t* rgba8 = nullptr;
  uint32_t w = 0, h = 0, layers = 0;
};
// textureSampleLevel(material_textures, sampler, uv, layer, 0) with the
// linear / repeat sampler of resource.cpp:589-606 (material.wesl:181-262), in
// the declared fp32 semantics:
//   s' = s - floor(s), NaN -> 0            repeat, in normalised coordinates
//   u = s' * W - 0.5, i0 = floor(u), alpha = u - i0   (texel centres at i + 1/2)
//   i0 -> i0 mod W, i1 = (i0 + 1) mod W   (same in t with H, beta)
//   texel channel = k / 255               (rgba8unorm)
//   value = lerp(lerp(T00, T10, alpha), lerp(T01, T11, alpha), beta),
//   lerp(p, q, w) = p + w * (q - p)
// A layer index past the array reads the last layer (WGSL clamps it).
struct Rgba {
  float c[4];
};
Rgba sample_texture(const TexArr& T, int32_t index, f2 uv) {
  Rgba out{{1.0f, 1.0f, 1.0f, 1.0f}};
  if (T.layers == 0) return out;  // the white 1x1 default layer
  const uint32_t layer = std::min(static_cast<uint32_t>(index), T.layers - 1u);
  const uint32_t dims[2] = {T.w, T.h};
  const float st[2] = {uv.x, uv.y};
  uint32_t lo[2], hi[2];
  float frac[2];
  for (int a = 0; a < 2; ++a) {
    float r = st[a] - std::floor(st[a]);
    if (std::isnan(r)) r = 0.0f;
    const float u = r * static_cast<float>(dims[a]) - 0.5f;
    const float i0 = std::floor(u);
    frac[a] = u - i0;
    const int64_t n = static_cast<int64_t>(dims[a]);
    const int64_t k = ((static_cast<int64_t>(i0) % n) + n) % n;
    lo[a] = static_cast<uint32_t>(k);
    hi[a] = static_cast<uint32_t>((k + 1) % n);
  }
  auto texel = [&](uint32_t x, uint32_t y, int ch) {
    const size_t off = ((static_cast<size_t>(layer) * T.h + y) * T.w + x) * 4 + static_cast<size_t>(ch);
    return static_cast<float>(T.rgba8[off]) / 255.0f;
  };
  auto lerp = [](float p, float q, float w) { return p + w * (q - p); };
  for (int ch = 0; ch < 4; ++ch) {
    const float top = lerp(texel(lo[0], lo[1], ch), texel(hi[0], lo[1], ch), frac[0]);
    const float bottom = lerp(texel(lo[0], hi[1], ch), texel(hi[0], hi[1], ch), frac[0]);
    out.c[ch] = lerp(top, bottom, frac[1]);
  }
  return out;
}
f2 texture_uv(const f2 uvs[2], uint32_t tex_coord) { return tex_coord == 1u ? uvs[1] : uvs[0]; }  // :173-178

Mat decode_material(const cg_material& pm, float l, const TexArr& T, const f2 uvs[2]) {  // material.wesl:19-53
  // baseColor (:180-191)
  float color[4] = {pm.base_color_factor[0], pm.base_color_factor[1], pm.base_color_factor[2],
                    pm.base_color_factor[3]};
  const cg_tex_info& tb = pm.tex[CG_TEX_BASE_COLOR];
  if (tb.index >= 0) {
    Rgba t = sample_texture(T, tb.index, texture_uv(uvs, tb.tex_coord));
    for (int k = 0; k < 4; ++k) color[k] *= t.c[k];
  }
  float refl = rgb_to_reflectance(mk(color[0], color[1], color[2]), l);
  // roughness (:192-203) and metallic (:205-216): channels 1 and 2
  const cg_tex_info& tmr = pm.tex[CG_TEX_METALLIC_ROUGHNESS];
  float rough = pm.roughness_factor;
  if (tmr.index >= 0) rough *= sample_texture(T, tmr.index, texture_uv(uvs, tmr.tex_coord)).c[1];
  rough = clampf(rough, 0.0f, 1.0f);
  float metal = pm.metallic_factor;
  if (tmr.index >= 0) metal *= sample_texture(T, tmr.index, texture_uv(uvs, tmr.tex_coord)).c[2];
  metal = clampf(metal, 0.0f, 1.0f);
  // transmission (:218-229): channel 0
  const cg_tex_info& tt = pm.tex[CG_TEX_TRANSMISSION];
  float trans = pm.transmission_factor;
  if (tt.index >= 0) trans *= sample_texture(T, tt.index, texture_uv(uvs, tt.tex_coord)).c[0];
  trans = clampf(trans, 0.0f, 1.0f);
  float raw_ior = pm.ior;
  float real_ior = (metal == 0.0f && trans > 0.0f) ? dispersed_ior(raw_ior, pm.dispersion, l) : raw_ior;
  IOR ior{real_ior, metal * imag_ior(real_ior, refl)};
  // emission (:250-262)
  float strength[3] = {pm.emissive_factor[0], pm.emissive_factor[1], pm.emissive_factor[2]};
  const cg_tex_info& te = pm.tex[CG_TEX_EMISSIVE];
  if (te.index >= 0) {
    Rgba t = sample_texture(T, te.index, texture_uv(uvs, te.tex_coord));
    for (int k = 0; k < 3; ++k) strength[k] *= t.c[k];
  }
  v3 em = mk(fmx(strength[0] * pm.emissive_strength, 0.0f), fmx(strength[1] * pm.emissive_strength, 0.0f),
             fmx(strength[2] * pm.emissive_strength, 0.0f));
  Mat m;
  m.diffuse = (1.0f - metal) * (1.0f - trans) * refl;
  m.alpha = rough * rough;
  m.ior = ior;
  m.transmission = trans;
  m.volume = pm.thickness_factor > 0.0f;
  m.emission = rgb_to_reflectance(em, l);
  return m;
}

// ---------------------------------------------------------------------------
// shading/util.wesl:4-81
inline v3 oriented(v3 d, v3 n) { return dot(d, n) > 0.0f ? -n : n; }
inline IOR inv_ior(IOR i) {
  float nsq = i.re * i.re + i.im * i.im;
  return IOR{i.re / nsq, -i.im / nsq};
}
inline IOR rel_ior(v3 sn, v3 d, const Mat& m) { return dot(sn, d) < 0.0f ? inv_ior(m.ior) : m.ior; }
float complex_fresnel(float c, IOR ior) {  // util.wesl:62-81
  float re = ior.re, im = ior.im;
  float rsq = re * re, isq = im * im, csq = c * c;
  float nsq = rsq + isq;
  float trc = 2.0f * re * c;
  float perp = (nsq - trc + csq) / fmx(nsq + trc + csq, 1e-6f);
  float par = (nsq * csq - trc + 1.0f) / fmx(nsq * csq + trc + 1.0f, 1e-6f);
  return clampf(0.5f * (perp + par), 0.0f, 1.0f);
}
float fresnel(v3 d, v3 n, IOR rel) {  // util.wesl:26-60
  float ci = clampf(dot(-d, n), 0.0f, 1.0f);
  if (std::fabs(rel.im) > 1e-6f) return complex_fresnel(ci, rel);
  float r = rel.re;
  float ts2 = r * r * fmx(0.0f, 1.0f - ci * ci);
  if (ts2 >= 1.0f) return 1.0f;
  float tc = std::sqrt(fmx(0.0f, 1.0f - ts2));
  float perp = (r * ci - tc) / (r * ci + tc);
  float par = (ci - r * tc) / (ci + r * tc);
  return clampf(0.5f * (perp * perp + par * par), 0.0f, 1.0f);
}

// math/geometric.wesl:28-33
struct Frame {
  v3 t, b, n;
};
inline Frame tangent_space(v3 n) {
  v3 helper = std::fabs(n.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
  v3 t = normalize(cross(helper, n));
  v3 b = cross(n, t);
  return Frame{t, b, n};
}
inline v3 to_world(const Frame& f, v3 v) { return (f.t * v.x + f.b * v.y) + f.n * v.z; }
inline v3 to_local(const Frame& f, v3 v) { return mk(dot(f.t, v), dot(f.b, v), dot(f.n, v)); }

// structural/types.wesl:3-19
struct Surface {
  v3 pos, sn, gn;
  uint32_t material;
  f2 uv0, uv1;
};

// sin and cos of a sampling angle in [0, 2 pi], as declared for this path:
// Cody-Waite reduction by pi/2 (three-part constant) and Taylor polynomials of
// degree 9 (sin) and 10 (cos) on [-pi/4, pi/4], evaluated in this order.
// Accuracy is about one ulp, well inside WGSL's 2^-11 bound for sin/cos.
void sincos_decl(float x, float& s, float& c) {
  const float fk = std::floor(x * 0x1.45f306p-1f + 0.5f);
  const float r = ((x - fk * 0x1.92p+0f) - fk * 0x1.fb54p-12f) - fk * 0x1.10b46p-30f;
  const float r2 = r * r;
  float ps = 0x1.71de3ap-19f;
  ps = -0x1.a01a02p-13f + r2 * ps;
  ps = 0x1.111112p-7f + r2 * ps;
  ps = -0x1.555556p-3f + r2 * ps;
  const float sr = r + (r * r2) * ps;
  float pc = -0x1.27e4fcp-22f;
  pc = 0x1.a01a02p-16f + r2 * pc;
  pc = -0x1.6c16c2p-10f + r2 * pc;
  pc = 0x1.555556p-5f + r2 * pc;
  pc = -0.5f + r2 * pc;
  const float cr = 1.0f + r2 * pc;
  switch (static_cast<int>(fk) & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
  }
}

// ---------------------------------------------------------------------------
// Lobes.  pdf/*.wesl and dist/*.wesl.
struct TSample {
  v3 dir;
  float prob;
};

// pdf/diffuse.wesl:7-36, dist/diffuse.wesl:6-15
TSample diffuse_sample(v3 d, const Surface& s, uint32_t& rng) {
  v3 macro = oriented(d, s.sn);
  float r = std::sqrt(rand_float(rng));
  float angle = 2.0f * PI * rand_float(rng);
  float sa, ca;
  sincos_decl(angle, sa, ca);
  v3 local = mk(r * ca, r * sa, std::sqrt(fmx(0.0f, 1.0f - r * r)));
  v3 next = to_world(tangent_space(macro), local);
  return TSample{next, fmx(dot(next, macro), 0.0f) / PI};
}
float diffuse_pdf(v3 d, v3 next, const Surface& s) {
  v3 macro = oriented(d, s.sn);
  return fmx(dot(next, macro), 0.0f) / PI;
}
float diffuse_dist(v3 d, v3 next, const Surface& s, const Mat& m) {
  v3 macro = oriented(d, s.sn);
  float np = fmx(dot(next, macro), 0.0f);
  return m.diffuse * np / PI;
}

// pdf/specular_reflection.wesl:7-39, dist/specular_reflection.wesl:8-25
inline v3 specular_dir(v3 d, const Surface& s) { return reflect(d, oriented(d, s.sn)); }
float spec_reflect_pdf(v3 d, v3 next, const Surface& s) {
  v3 e = specular_dir(d, s);
  v3 df = next - e;
  return dot(df, df) <= 1e-6f ? 1.0f : 0.0f;
}
float spec_reflect_dist(v3 d, v3 next, const Surface& s, const Mat& m) {
  v3 macro = oriented(d, s.sn);
  v3 e = reflect(d, macro);
  v3 df = next - e;
  float distribution = dot(df, df) <= 1e-6f ? 1.0f : 0.0f;
  IOR rel = rel_ior(s.sn, d, m);
  return fresnel(d, macro, rel) * distribution;
}

// pdf/specular_transmission.wesl:7-69, dist/specular_transmission.wesl:8-52
inline v3 spec_trans_dir(v3 d, const Surface& s, const Mat& m) {
  if (m.volume) {
    IOR rel = rel_ior(s.sn, d, m);
    return refract(d, oriented(d, s.sn), rel.re);
  }
  return d;
}
float spec_trans_pdf(v3 d, v3 next, const Surface& s, const Mat& m) {
  v3 e = spec_trans_dir(d, s, m);
  v3 df = next - e;
  return (dot(e, e) > 1e-6f && dot(df, df) <= 1e-6f) ? 1.0f : 0.0f;
}
float spec_trans_dist(v3 d, v3 next, const Surface& s, const Mat& m) {
  v3 macro = oriented(d, s.sn);
  v3 e = spec_trans_dir(d, s, m);
  v3 df = next - e;
  float distribution = dot(df, df) <= 1e-6f ? 1.0f : 0.0f;
  IOR rel = rel_ior(s.sn, d, m);
  return m.transmission * (1.0f - fresnel(d, macro, rel)) * distribution;
}

// GGX pieces shared by rough reflection and transmission, with the
// reference's use of alpha where standard GGX has alpha^2
// (pdf/rough_reflection.wesl:120-139, dist/rough_reflection.wesl:51-72).
float ggx_d(v3 micro, v3 macro, float alpha) {
  float c = clampf(dot(micro, macro), 0.0f, 1.0f);
  float c2 = c * c;
  float t2 = fmx(0.0f, 1.0f - c2) / fmx(c2, 1e-6f);
  float at = alpha + t2;
  float den = PI * (c2 * c2) * (at * at);
  return alpha / fmx(den, 1e-6f);
}
// The reference writes tan(acos(c)) * tan(acos(c)) for a projection
// c in [1e-6, 1] (rough_reflection.wesl:114, 122-123, 185-195).  That
// composition is ill-conditioned at grazing c (one ulp of acos near pi/2 moves
// tan by ~1e-7 / c), and WGSL leaves acos's accuracy to the implementation, so
// its value is declared here as the exact identity max(0, 1 - c^2) / c^2.
inline float tan2_acos(float c) { return fmx(0.0f, 1.0f - c * c) / (c * c); }
float ggx_g1(float c, float alpha) {
  float t2 = tan2_acos(c);
  return 1.0f / (1.0f + (std::sqrt(1.0f + alpha * t2) - 1.0f) / 2.0f);
}
float ggx_g2(float c1, float c2, float alpha) {
  float t1 = tan2_acos(c1), t2 = tan2_acos(c2);
  return 1.0f / (1.0f + (std::sqrt(1.0f + alpha * t1) - 1.0f) / 2.0f +
                 (std::sqrt(1.0f + alpha * t2) - 1.0f) / 2.0f);
}
// pdf/rough_reflection.wesl:52-103 (identical in rough_transmission.wesl:68-119)
v3 sample_micro_normal(v3 d, const Surface& s, float alpha, uint32_t& rng) {
  v3 macro = oriented(d, s.sn);
  Frame f = tangent_space(macro);
  v3 view = -d;
  v3 local = normalize(to_local(f, view));
  v3 st = normalize(mk(alpha * local.x, alpha * local.y, fmx(local.z, 1e-6f)));
  float tl2 = st.x * st.x + st.y * st.y;
  v3 tx = tl2 > 1e-6f ? mk(-st.y, st.x, 0.0f) / std::sqrt(fmx(tl2, 1e-6f)) : mk(1.0f, 0.0f, 0.0f);
  v3 ty = cross(st, tx);
  float radius = std::sqrt(rand_float(rng));
  float angle = 2.0f * PI * rand_float(rng);
  float sa, ca;
  sincos_decl(angle, sa, ca);
  float px = radius * ca;
  float py = radius * sa;
  float vis = 0.5f * (1.0f + st.z);
  py = mixf(std::sqrt(fmx(0.0f, 1.0f - px * px)), py, vis);
  v3 lsm = px * tx + py * ty + std::sqrt(fmx(0.0f, 1.0f - px * px - py * py)) * st;
  v3 lm = normalize(mk(alpha * lsm.x, alpha * lsm.y, fmx(lsm.z, 1e-6f)));
  return normalize(to_world(f, lm));
}
float reflect_pdf(v3 d, v3 micro, v3 macro, float alpha) {  // rough_reflection.wesl:105-118
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float cmi = fmx(dot(-d, micro), 1e-6f);
  float vnp = ggx_d(micro, macro, alpha) * ggx_g1(cmp, alpha) * cmi / cmp;
  return vnp / (4.0f * cmi);
}
TSample rough_reflect_sample(v3 d, const Surface& s, const Mat& m, uint32_t& rng) {
  v3 micro = oriented(d, sample_micro_normal(d, s, m.alpha, rng));
  v3 next = reflect(d, micro);
  v3 macro = oriented(d, s.sn);
  if (dot(next, macro) <= 0.0f) return TSample{next, 0.0f};
  return TSample{next, reflect_pdf(d, micro, macro, m.alpha)};
}
float rough_reflect_pdf(v3 d, v3 next, const Surface& s, const Mat& m) {
  v3 macro = oriented(d, s.sn);
  if (dot(next, macro) <= 0.0f) return 0.0f;
  v3 mu = next - d;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  v3 micro = oriented(d, normalize(mu));
  return reflect_pdf(d, micro, macro, m.alpha);
}
float rough_reflect_dist(v3 d, v3 next, const Surface& s, const Mat& m) {  // dist/rough_reflection.wesl:8-49
  v3 macro = oriented(d, s.sn);
  float np = fmx(dot(next, macro), 0.0f);
  if (np <= 0.0f) return 0.0f;
  v3 mu = next - d;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  v3 raw = normalize(mu);
  v3 micro = oriented(d, raw);
  IOR rel = rel_ior(raw, d, m);
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float nmp = fmx(dot(next, macro), 1e-6f);
  float rd = ggx_d(micro, macro, m.alpha) * ggx_g2(cmp, nmp, m.alpha) /
             (4.0f * cmp * nmp);
  return fresnel(d, micro, rel) * rd * np;
}

// pdf/rough_transmission.wesl:9-181, dist/rough_transmission.wesl:8-115
inline float rel_value(IOR i) { return fmx(i.re, 1e-6f); }
inline float trans_ior(IOR i) { return 1.0f / rel_value(i); }
float transmit_pdf(v3 d, v3 micro, v3 macro, v3 next, float tior, float alpha) {
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float cmi = fmx(dot(-d, micro), 1e-6f);
  float vnp = ggx_d(micro, macro, alpha) * ggx_g1(cmp, alpha) * cmi / cmp;
  float nmi = dot(next, micro);
  float den = nmi + cmi / tior;
  return vnp * cmi / fmx(den * den, 1e-6f);
}
float rough_trans_pdf(v3 d, v3 next, const Surface& s, const Mat& m) {
  v3 macro = oriented(d, s.sn);
  if (dot(next, macro) >= 0.0f) return 0.0f;
  IOR mrel = rel_ior(macro, d, m);
  float mrv = rel_value(mrel);
  v3 mu = mrv * d - next;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  v3 micro = oriented(d, normalize(mu));
  IOR rel = rel_ior(micro, d, m);
  return transmit_pdf(d, micro, macro, next, trans_ior(rel), m.alpha);
}
TSample rough_trans_sample(v3 d, const Surface& s, const Mat& m, uint32_t& rng) {
  v3 macro = oriented(d, s.sn);
  v3 micro = oriented(d, sample_micro_normal(d, s, m.alpha, rng));
  IOR rel = rel_ior(micro, d, m);
  v3 next = refract(d, micro, rel_value(rel));
  if (dot(next, next) <= 1e-6f) return TSample{next, 0.0f};
  if (dot(next, macro) >= 0.0f) return TSample{next, 0.0f};
  return TSample{next, rough_trans_pdf(d, next, s, m)};
}
float rough_trans_dist(v3 d, v3 next, const Surface& s, const Mat& m) {
  v3 macro = oriented(d, s.sn);
  if (dot(next, macro) >= 0.0f) return 0.0f;
  float np = fmx(dot(-next, macro), 0.0f);
  IOR mrel = rel_ior(macro, d, m);
  float mrv = rel_value(mrel);
  v3 mu = mrv * d - next;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  v3 raw = normalize(mu);
  v3 micro = oriented(d, raw);
  IOR rel = rel_ior(raw, d, m);
  float tior = trans_ior(rel);
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float nmp = fmx(dot(-next, macro), 1e-6f);
  float cmi = fmx(dot(-d, micro), 1e-6f);
  float nmi = dot(next, micro);
  float den = nmi + cmi / tior;
  float td = m.transmission * ggx_d(micro, macro, m.alpha) *
             ggx_g2(cmp, nmp, m.alpha) * std::fabs(cmi * nmi) /
             fmx(cmp * nmp * den * den, 1e-6f);
  return (1.0f - fresnel(d, micro, rel)) * td * np;
}

// ---------------------------------------------------------------------------
// Scene held by the oracle (reference-format SceneData).
}  // namespace

struct orc_scene {
  std::vector<cg_tlas_node> tlas;
  std::vector<cg_instance> inst;
  std::vector<cg_blas_node> blas;
  std::vector<cg_index> idx;
  std::vector<uint32_t> tri_ids;
  std::vector<cg_vertex> vert;
  std::vector<cg_material> mat;
  std::vector<cg_emissive_prim> em;
  std::vector<cg_alias_entry> alias;
  std::vector<float> env;
  uint32_t env_w = 0, env_h = 0;
  std::vector<uint8_t> tex;  // material texture array
  uint32_t tex_w = 0, tex_h = 0, tex_layers = 0;
  TexArr textures() const { return TexArr{tex.data(), tex_w, tex_h, tex_layers}; }
};

namespace {

// math/geometric.wesl:5-26
inline v3 dir_inv(v3 d) {
  v3 s = mk(d.x == 0.0f ? 1e-7f : d.x, d.y == 0.0f ? 1e-7f : d.y, d.z == 0.0f ? 1e-7f : d.z);
  return mk(1.0f / s.x, 1.0f / s.y, 1.0f / s.z);
}
// Returns t_near (or +inf when the box is missed).
inline bool ray_aabb(v3 pos, v3 inv, const float* lb, const float* ub, float* t_near_out) {
  v3 t0 = (mk(lb) - pos) * inv;
  v3 t1 = (mk(ub) - pos) * inv;
  v3 tmin = vmin(t0, t1), tmax = vmax(t0, t1);
  float t_near = fmx(fmx(tmin.x, tmin.y), tmin.z);
  float t_far = fmn(fmn(tmax.x, tmax.y), tmax.z);
  *t_near_out = t_near;
  return t_far >= t_near && t_far > 0.0f;
}
// structural/bvh.wesl:188-221 (Moller-Trumbore)
struct TriHit {
  float t, u, v;
};
inline TriHit ray_triangle(v3 pos, v3 dir, v3 v0, v3 v1, v3 v2) {
  v3 e1 = v1 - v0, e2 = v2 - v0;
  v3 h = cross(dir, e2);
  float a = dot(e1, h);
  if (a > -1e-6f && a < 1e-6f) return TriHit{-1.0f, 0.0f, 0.0f};
  float f = 1.0f / a;
  v3 s = pos - v0;
  float u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return TriHit{-1.0f, 0.0f, 0.0f};
  v3 q = cross(s, e1);
  float v = f * dot(dir, q);
  if (v < 0.0f || u + v > 1.0f) return TriHit{-1.0f, 0.0f, 0.0f};
  float t = f * dot(e2, q);
  if (t > 1e-6f) return TriHit{t, u, v};
  return TriHit{-1.0f, 0.0f, 0.0f};
}

struct Counts {
  uint32_t tlas_nodes = 0, blas_nodes = 0, tris = 0, insts = 0;
};
// The two smallest hit distances over all triangles the ray hits, so callers
// can exclude ties from id comparisons.
struct Ties {
  float a = MAXD, b = MAXD;
  void note(float t) {
    if (t < a) {
      b = a;
      a = t;
    } else if (t < b) {
      b = t;
    }
  }
};
struct BlasBest {
  uint32_t slot = CG_MISS_ID;
  float u = 0, v = 0;
  v3 ln{0, 0, 0}, lgn{0, 0, 0};
  f2 uv0{0, 0}, uv1{0, 0};
};

// structural/bvh.wesl:131-186 — LIFO walk; right child pops first.  With
// cull, boxes entered beyond `bound` (the best hit so far) are skipped.
float traverse_blas(const orc_scene& sc, v3 pos, v3 dir, uint32_t root, float bound, BlasBest& out,
                    Ties& ties, Counts& cnt, bool cull) {
  uint32_t stack[64];
  int top = 0;
  stack[0] = root;
  v3 inv = dir_inv(dir);
  float closest_t = MAXD;
  while (top >= 0) {
    uint32_t ni = stack[top--];
    const cg_blas_node& n = sc.blas[std::min<size_t>(ni, sc.blas.size() - 1)];
    cnt.blas_nodes++;
    float tn;
    if (!ray_aabb(pos, inv, n.lb, n.ub, &tn)) continue;
    if (cull && tn > fmn(bound, closest_t)) continue;
    if (n.triangle_count == 0u) {
      stack[++top] = n.child;
      stack[++top] = n.child + 1u;
    } else {
      for (uint32_t i = 0; i < n.triangle_count; ++i) {
        uint32_t slot = n.child + i;
        const cg_index& ix = sc.idx[slot];
        v3 v0 = mk(sc.vert[ix.i0].pos), v1 = mk(sc.vert[ix.i1].pos), v2 = mk(sc.vert[ix.i2].pos);
        cnt.tris++;
        TriHit h = ray_triangle(pos, dir, v0, v1, v2);
        if (h.t > 0.0f) ties.note(h.t);
        if (h.t > 0.0f && h.t < closest_t) {
          closest_t = h.t;
          v3 n0 = mk(sc.vert[ix.i0].nrm), n1 = mk(sc.vert[ix.i1].nrm), n2 = mk(sc.vert[ix.i2].nrm);
          float w0 = 1.0f - h.u - h.v;
          out.slot = slot;
          out.u = h.u;
          out.v = h.v;
          out.ln = normalize(w0 * n0 + h.u * n1 + h.v * n2);
          out.lgn = normalize(cross(v1 - v0, v2 - v0));
          f2 a0{sc.vert[ix.i0].uv0[0], sc.vert[ix.i0].uv0[1]}, a1{sc.vert[ix.i1].uv0[0], sc.vert[ix.i1].uv0[1]},
              a2{sc.vert[ix.i2].uv0[0], sc.vert[ix.i2].uv0[1]};
          f2 b0{sc.vert[ix.i0].uv1[0], sc.vert[ix.i0].uv1[1]}, b1{sc.vert[ix.i1].uv1[0], sc.vert[ix.i1].uv1[1]},
              b2{sc.vert[ix.i2].uv1[0], sc.vert[ix.i2].uv1[1]};
          out.uv0 = (w0 * a0 + h.u * a1) + h.v * a2;
          out.uv1 = (w0 * b0 + h.u * b1) + h.v * b2;
        }
      }
    }
  }
  return closest_t;
}

struct Hit {
  bool hit;
  float dist;
  Surface s;
  uint32_t inst, tri;
};

// structural/bvh.wesl:56-130
Hit ray_hit(const orc_scene& sc, v3 pos, v3 dir, Counts& cnt, bool cull, float* second_t, float* uv_out) {
  Hit res{};
  res.hit = false;
  res.dist = MAXD;
  res.inst = res.tri = CG_MISS_ID;
  Ties ties;
  if (sc.tlas.empty()) {
    if (second_t) *second_t = MAXD;
    return res;
  }
  uint32_t stack[64];
  int top = 0;
  stack[0] = 0u;
  float dis = MAXD;
  BlasBest best;
  const float* inv_trans = nullptr;
  uint32_t material = 0, best_inst = CG_MISS_ID;
  v3 inv = dir_inv(dir);
  while (top >= 0) {
    uint32_t ni = stack[top--];
    const cg_tlas_node& n = sc.tlas[std::min<size_t>(ni, sc.tlas.size() - 1)];
    cnt.tlas_nodes++;
    float tn;
    if (!ray_aabb(pos, inv, n.lb, n.ub, &tn)) continue;
    if (cull && tn > dis) continue;
    if (n.child != 0u) {
      stack[++top] = n.child;
      stack[++top] = n.child + 1u;
    } else {
      const cg_instance& in = sc.inst[std::min<size_t>(n.instance_idx, sc.inst.size() - 1)];
      cnt.insts++;
      v3 lpos = mat_point(in.inv_trans, pos);
      v3 ldir = mat_dir(in.inv_trans, dir);
      BlasBest local;
      float gd = traverse_blas(sc, lpos, ldir, in.blas_root_idx, dis, local, ties, cnt, cull);
      if (gd < dis) {
        dis = gd;
        best = local;
        inv_trans = in.inv_trans;
        material = in.material_idx;
        best_inst = n.instance_idx;
      }
    }
  }
  if (second_t) *second_t = ties.b;
  if (dis >= MAXD) return res;
  v3 wgn = normalize(mat_t_dir(inv_trans, best.lgn));
  v3 wsn = normalize(mat_t_dir(inv_trans, best.ln));
  res.hit = true;
  res.dist = dis;
  res.s.pos = pos + dis * dir;
  res.s.sn = dot(wsn, wgn) >= 0.0f ? wsn : -wsn;
  res.s.gn = wgn;
  res.s.material = material;
  res.s.uv0 = best.uv0;
  res.s.uv1 = best.uv1;
  res.inst = best_inst;
  res.tri = sc.tri_ids.empty() ? best.slot : sc.tri_ids[best.slot];
  if (uv_out) {
    uv_out[0] = best.u;
    uv_out[1] = best.v;
  }
  return res;
}

// ---------------------------------------------------------------------------
// Emission: pdf/emission.wesl:28-191.
struct ESample {
  v3 dir;
  float prob, radiance, expected_dist;
  uint32_t prim;  // the emitter the alias table picked
};
inline bool can_sample(const orc_scene& sc) { return !sc.alias.empty(); }
uint32_t sample_emissive_prim(const orc_scene& sc, uint32_t& rng) {  // emission.wesl:181-191
  uint32_t n = static_cast<uint32_t>(sc.alias.size());
  uint32_t bin = f2u(rand_float(rng) * static_cast<float>(n));
  float randf = rand_float(rng);
  uint32_t b = std::min(bin, n - 1u);
  uint32_t pick = randf < sc.alias[b].prob ? bin : sc.alias[b].alias;
  return std::min<uint32_t>(pick, static_cast<uint32_t>(sc.em.size()) - 1u);
}
ESample emission_sample(const orc_scene& sc, float l, const Surface& s, uint32_t& rng) {  // :46-96
  if (!can_sample(sc)) return ESample{mk(0, 0, 0), 0.0f, 0.0f, 0.0f, 0u};
  const uint32_t prim = sample_emissive_prim(sc, rng);
  const cg_emissive_prim& p = sc.em[prim];
  float tx = rand_float(rng);
  float ty = rand_float(rng);
  float sq = std::sqrt(tx);
  v3 bary = mk(1.0f - sq, sq * (1.0f - ty), sq * ty);
  v3 p0 = mk(p.v[0]), p1 = mk(p.v[1]), p2 = mk(p.v[2]);
  v3 sp = (bary.x * p0 + bary.y * p1) + bary.z * p2;
  v3 off = sp - s.pos;
  float dsq = fmx(dot(off, off), 1e-3f);
  v3 next = off / std::sqrt(dsq);
  v3 ln = normalize(cross(p1 - p0, p2 - p0));
  float lp = std::fabs(dot(ln, -next));
  float apdf = fmx(p.weight / p.area, 1e-3f);
  float sapdf = apdf * dsq / fmx(lp, 1e-3f);
  return ESample{next, sapdf, rgb_to_radiance(mk(p.emission), l), std::sqrt(dsq), prim};
}
float triangle_hit_dist(v3 pos, v3 dir, const cg_emissive_prim& p) {  // :151-179
  v3 p0 = mk(p.v[0]), p1 = mk(p.v[1]), p2 = mk(p.v[2]);
  v3 e1 = p1 - p0, e2 = p2 - p0;
  v3 pv = cross(dir, e2);
  float det = dot(e1, pv);
  if (std::fabs(det) <= 1e-3f) return -1.0f;
  float inv = 1.0f / det;
  v3 tv = pos - p0;
  float u = dot(tv, pv) * inv;
  if (u < 0.0f || u > 1.0f) return -1.0f;
  v3 qv = cross(tv, e1);
  float v = dot(dir, qv) * inv;
  if (v < 0.0f || u + v > 1.0f) return -1.0f;
  float hd = dot(e2, qv) * inv;
  return hd > 1e-3f ? hd : -1.0f;
}
float light_pdf(const orc_scene& sc, v3 next, const Surface& s) {  // :120-149
  if (!can_sample(sc)) return 0.0f;
  float pdf = 0.0f;
  v3 rd = normalize(next);
  for (const cg_emissive_prim& p : sc.em) {
    float hd = triangle_hit_dist(s.pos, rd, p);
    if (hd <= 0.0f) continue;
    v3 p0 = mk(p.v[0]), p1 = mk(p.v[1]), p2 = mk(p.v[2]);
    v3 ln = normalize(cross(p1 - p0, p2 - p0));
    float lp = std::fabs(dot(ln, -rd));
    float apdf = fmx(p.weight / p.area, 1e-3f);
    pdf += apdf * hd * hd / fmx(lp, 1e-3f);
  }
  return pdf;
}

// ---------------------------------------------------------------------------
// shading/sample.wesl:45-501 — lobe counts, sampling and balance-heuristic MIS.
enum { M_NONE = 0, M_DIFFUSE, M_SPEC_R, M_ROUGH_R, M_SPEC_T, M_ROUGH_T };
struct Counts6 {
  uint32_t d = 0, sr = 0, rr = 0, st = 0, rt = 0, l = 0;
};
struct Params {
  int D;  // MAX_DEPTH == TRACE_DEPTH (clamped to [1, 64])
  uint32_t T, E, DIF, ROUGH;
};
float normal_reflectance(const Mat& m) {  // sample.wesl:147-160
  float lo = m.ior.re - 1.0f, hi = m.ior.re + 1.0f, isq = m.ior.im * m.ior.im;
  return clampf((lo * lo + isq) / fmx(hi * hi + isq, 1e-6f), 0.0f, 1.0f);
}
uint32_t scaled_count(float h, float scale, uint32_t mx) {  // :214-219
  if (h <= 1e-6f) return 0u;
  return std::min(std::max(1u, f2u(h * scale)), mx);
}
Counts6 sample_counts(bool can, const Mat& m, int depth, const Params& P);
Counts6 sample_counts(const orc_scene& sc, const Mat& m, int depth, const Params& P) {  // :110-246
  return sample_counts(can_sample(sc), m, depth, P);
}
Counts6 sample_counts(bool can, const Mat& m, int depth, const Params& P) {
  float hd = m.diffuse;
  float refl = normal_reflectance(m);
  float trans = m.transmission;
  float hsr = m.alpha == 0.0f ? refl : 0.0f;
  float hrr = m.alpha > 0.0f ? refl : 0.0f;
  float hst = (trans > 1e-6f && m.alpha == 0.0f) ? trans : 0.0f;
  float hrt = (trans > 1e-6f && m.alpha > 0.0f) ? trans : 0.0f;
  float hl = can ? 1.0f : 0.0f;
  float heur = static_cast<float>(P.D - depth) / static_cast<float>(P.D);  // lod.wesl:16-18
  uint32_t t_avail = f2u(heur * static_cast<float>(std::min(P.T, 4u)));
  uint32_t e_avail = f2u(heur * static_cast<float>(std::min(P.E, 1u)));
  Counts6 c;
  float sum = hd + hsr + hrr + hst + hrt;
  if (!(t_avail == 0u || sum <= 1e-6f)) {
    float scale = static_cast<float>(t_avail) / sum;
    Counts6 r;
    r.d = scaled_count(hd, scale, std::min(P.DIF, 1u));
    r.sr = scaled_count(hsr, scale, 1u);
    r.rr = scaled_count(hrr, scale, std::min(P.ROUGH, 1u));
    r.st = scaled_count(hst, scale, 1u);
    r.rt = scaled_count(hrt, scale, std::min(P.ROUGH, 1u));
    uint32_t rem = t_avail;
    c.d = std::min(r.d, rem);
    rem -= c.d;
    c.sr = std::min(r.sr, rem);
    rem -= c.sr;
    c.rr = std::min(r.rr, rem);
    rem -= c.rr;
    c.st = std::min(r.st, rem);
    rem -= c.st;
    c.rt = std::min(r.rt, rem);
  }
  c.l = (hl > 1e-6f && e_avail > 0u) ? std::max(1u, e_avail) : 0u;
  return c;
}
float balance_denominator(const orc_scene& sc, v3 d, v3 next, const Surface& s, const Mat& m,
                          const Counts6& c, int model, float model_pdf, float sampled_light_pdf) {  // :452-501
  float pd = model == M_DIFFUSE ? model_pdf : diffuse_pdf(d, next, s);
  float psr = model == M_SPEC_R ? model_pdf : spec_reflect_pdf(d, next, s);
  float prr = model == M_ROUGH_R ? model_pdf : rough_reflect_pdf(d, next, s, m);
  float pst = model == M_SPEC_T ? model_pdf : spec_trans_pdf(d, next, s, m);
  float prt = model == M_ROUGH_T ? model_pdf : rough_trans_pdf(d, next, s, m);
  float pl = sampled_light_pdf >= 0.0f ? sampled_light_pdf : light_pdf(sc, next, s);
  return fmx(static_cast<float>(c.d) * pd + static_cast<float>(c.sr) * psr + static_cast<float>(c.rr) * prr +
                 static_cast<float>(c.st) * pst + static_cast<float>(c.rt) * prt +
                 static_cast<float>(c.l) * pl,
             1e-6f);
}

struct Ctx {  // shading.wesl:24-28 + sample.wesl:73-78
  uint32_t tcount = 0, ecount = 0;
  v3 tdir[4];
  float tcoeff[4];
  v3 edir;
  float erad = 0, edist = 0;
  uint32_t eprim = 0;
  float emission = 0;
};
Ctx forward_pass(const orc_scene& sc, v3 d, float l, const Surface& s, int depth, const Params& P,
                 uint32_t& rng) {  // shading.wesl:30-46
  const cg_material& pm = sc.mat[std::min<size_t>(s.material, sc.mat.size() - 1)];
  const f2 uvs[2] = {s.uv0, s.uv1};
  Mat m = decode_material(pm, l, sc.textures(), uvs);
  Counts6 c = sample_counts(sc, m, depth, P);
  Ctx x;
  auto add = [&](v3 dir, float coeff) {
    x.tdir[x.tcount] = dir;
    x.tcoeff[x.tcount] = coeff;
    x.tcount++;
  };
  for (uint32_t i = 0; i < c.d; ++i) {
    TSample t = diffuse_sample(d, s, rng);
    add(t.dir, diffuse_dist(d, t.dir, s, m) / balance_denominator(sc, d, t.dir, s, m, c, M_DIFFUSE, t.prob, -1.0f));
  }
  for (uint32_t i = 0; i < c.sr; ++i) {
    v3 nd = specular_dir(d, s);
    float prob = spec_reflect_pdf(d, nd, s);
    add(nd, spec_reflect_dist(d, nd, s, m) / balance_denominator(sc, d, nd, s, m, c, M_SPEC_R, prob, -1.0f));
  }
  for (uint32_t i = 0; i < c.rr; ++i) {
    TSample t = rough_reflect_sample(d, s, m, rng);
    add(t.dir, rough_reflect_dist(d, t.dir, s, m) / balance_denominator(sc, d, t.dir, s, m, c, M_ROUGH_R, t.prob, -1.0f));
  }
  for (uint32_t i = 0; i < c.st; ++i) {
    v3 nd = m.volume ? spec_trans_dir(d, s, m) : d;
    float prob = spec_trans_pdf(d, nd, s, m);
    add(nd, spec_trans_dist(d, nd, s, m) / balance_denominator(sc, d, nd, s, m, c, M_SPEC_T, prob, -1.0f));
  }
  for (uint32_t i = 0; i < c.rt; ++i) {
    TSample t = rough_trans_sample(d, s, m, rng);
    add(t.dir, rough_trans_dist(d, t.dir, s, m) / balance_denominator(sc, d, t.dir, s, m, c, M_ROUGH_T, t.prob, -1.0f));
  }
  for (uint32_t i = 0; i < c.l; ++i) {
    ESample e = emission_sample(sc, l, s, rng);
    if (e.prob <= 1e-6f) continue;  // addEmissionSample, sample.wesl:392-394
    float proj = std::fabs(dot(s.sn, e.dir));
    float coeff = (diffuse_dist(d, e.dir, s, m) + rough_reflect_dist(d, e.dir, s, m)) + rough_trans_dist(d, e.dir, s, m);
    x.edir = e.dir;
    x.erad = e.radiance * coeff * proj / balance_denominator(sc, d, e.dir, s, m, c, M_NONE, -1.0f, e.prob);
    x.edist = e.expected_dist;
    x.eprim = e.prim;
    x.ecount++;
  }
  x.emission = m.emission;
  return x;
}

// shading/spawn_surface_ray.wesl:6-15
inline void spawn(const Surface& s, v3 next, v3& pos, v3& dir) {
  v3 off = dot(next, s.gn) > 0.0f ? s.gn : -s.gn;
  pos = s.pos + off * 1e-3f;
  dir = next;
}

// environment.wesl:7-25
float env_radiance(const orc_scene& sc, v3 dir, float l, uint32_t* texel) {
  v3 d = normalize(dir);
  float u = std::atan2(d.y, d.x) / (2.0f * PI) + 0.5f;
  float v = std::acos(clampf(d.z, -1.0f, 1.0f)) / PI;
  float wu = u - std::floor(u);
  float cv = clampf(1.0f - v, 0.0f, 0.999999f);
  uint32_t x = std::min(sc.env_w - 1u, f2u(wu * static_cast<float>(sc.env_w)));
  uint32_t y = std::min(sc.env_h - 1u, f2u(cv * static_cast<float>(sc.env_h)));
  const float* t = &sc.env[(static_cast<size_t>(y) * sc.env_w + x) * 4];
  *texel = y * sc.env_w + x;
  return rgb_to_radiance(mk(t[0], t[1], t[2]), l);
}

inline uint32_t hmix(uint32_t h, uint32_t w) { return (h ^ w) * 16777619u; }

struct Tracer {
  const orc_scene& sc;
  Params P;
  float lambda;
  uint32_t& rng;
  uint32_t& hash;
  uint32_t& rays;
  bool cull;
  Counts cnt;

  // trace.wesl:22-89, as recursion: frame depth == stack top == lod depth.
  float trace(v3 pos, v3 dir, int depth) {
    Hit h = ray_hit(sc, pos, dir, cnt, cull, nullptr, nullptr);
    rays++;
    hash = hmix(hmix(hash, h.inst), h.tri);
    if (!h.hit) {
      uint32_t texel = 0;
      float r = env_radiance(sc, dir, lambda, &texel);
      hash = hmix(hash, 0x5E000000u ^ texel);  // the sky texel is a discrete decision too
      return r;
    }
    int trace_depth = std::max(1, std::min(P.D, 64));
    if (depth >= P.D || depth >= trace_depth - 1) return 0.0f;
    Ctx x = forward_pass(sc, dir, lambda, h.s, depth, P, rng);
    // trace.wesl:91-113 — shadow test of each emission sample
    for (uint32_t i = 0; i < x.ecount; ++i) {
      v3 lp = h.s.pos + x.edir * x.edist;
      v3 sp, sd;
      spawn(h.s, x.edir, sp, sd);
      float expected = length(sp - lp);
      Hit sh = ray_hit(sc, sp, sd, cnt, cull, nullptr, nullptr);
      rays++;
      bool occluded = sh.hit && sh.dist < (expected - 1e-3f);
      // the shadow outcome and the emitter the alias table picked are discrete decisions
      hash = hmix(hmix(hash, 0xA5000000u | (occluded ? 1u : 0u)), x.eprim);
      if (occluded) x.erad = 0.0f;
    }
    float tr[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < x.tcount; ++i) {
      v3 cp, cd;
      spawn(h.s, x.tdir[i], cp, cd);
      tr[i] = trace(cp, cd, depth + 1);
    }
    // shading/collect.wesl:4-17 and shading.wesl:48-53
    float rad = 0.0f;
    for (uint32_t i = 0; i < x.tcount; ++i) rad += x.tcoeff[i] * tr[i];
    for (uint32_t i = 0; i < x.ecount; ++i) rad += x.erad;
    return x.emission + rad;
  }
};

// compute.wesl:64-79
void camera_ray(const cg_camera& cam, uint32_t W, uint32_t H, uint32_t px, uint32_t py, v3& pos, v3& dir) {
  float cx = (static_cast<float>(px) / static_cast<float>(W)) * 2.0f - 1.0f;
  float cy = (static_cast<float>(py) / static_cast<float>(H)) * 2.0f - 1.0f;
  v3 cd = mk(cam.direction);
  v3 dx = normalize(cross(cd, mk(0.0f, 0.0f, 1.0f))) * cam.viewport[0];
  v3 dy = normalize(cross(cd, dx)) * cam.viewport[1];
  pos = mk(cam.position);
  dir = normalize(cd + cx * dx + cy * dy);
}

// physics/srgb.wesl:5-27
inline float encode_channel(float lin) {
  float v = fmx(lin, 0.0f);
  // declared: pow(v, f32(1/2.4)) rounded once to binary32 from double precision
  float e = v <= 0.0031308f ? 12.92f * v
                            : 1.055f * static_cast<float>(std::pow(static_cast<double>(v),
                                                                   static_cast<double>(1.0f / 2.4f))) - 0.055f;
  return clampf(e, 0.0f, 1.0f);
}
v3 xyz_to_srgb(v3 xyz) {
  v3 c0 = mk(2.36461385f, -0.51516621f, 0.0052037f);
  v3 c1 = mk(-0.89654057f, 1.4264081f, -0.01440816f);
  v3 c2 = mk(-0.46807328f, 0.0887581f, 1.00920446f);
  v3 lin = (c0 * xyz.x + c1 * xyz.y) + c2 * xyz.z;
  return mk(encode_channel(lin.x), encode_channel(lin.y), encode_channel(lin.z));
}
inline uint8_t unorm8(float v) { return static_cast<uint8_t>(std::floor(v * 255.0f + 0.5f)); }

template <class F>
void parallel_for(uint32_t n, int threads, F&& f) {
  if (threads <= 1 || n < 2) {
    for (uint32_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (;;) {
        uint32_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i);
      }
    });
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Host scene preparation (src/pathtracing/*.h).
struct TriInfo {  // blas.h:120-123
  v3 center;
  uint32_t index;
};

// blas.h:171-252 — median split by longest centroid axis, leaves <= 2.
void blas_build_tree(orc_scene& sc, uint32_t node_idx, uint32_t l, uint32_t r, std::vector<TriInfo>& ti,
                     const cg_primitive& prim, const std::vector<cg_vertex>& verts,
                     const std::vector<uint32_t>& indices, uint32_t tri_offset) {
  v3 lb = mk(FLT_MAX, FLT_MAX, FLT_MAX), ub = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
  for (uint32_t i = l; i < r; ++i) {
    uint32_t o = prim.index_offset + ti[i].index * 3u;
    v3 v0 = mk(verts[indices[o]].pos), v1 = mk(verts[indices[o + 1]].pos), v2 = mk(verts[indices[o + 2]].pos);
    lb = vmin(lb, vmin(vmin(v0, v1), v2));
    ub = vmax(ub, vmax(vmax(v0, v1), v2));
  }
  v3 pad = mk(0.001f, 0.001f, 0.001f);
  lb = lb - pad;
  ub = ub + pad;
  cg_blas_node& node = sc.blas[node_idx];
  node.lb[0] = lb.x; node.lb[1] = lb.y; node.lb[2] = lb.z;
  node.ub[0] = ub.x; node.ub[1] = ub.y; node.ub[2] = ub.z;
  uint32_t count = r - l;
  if (count <= 2) {
    node.child = tri_offset + l;
    node.triangle_count = count;
    return;
  }
  v3 clb = mk(FLT_MAX, FLT_MAX, FLT_MAX), cub = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
  for (uint32_t i = l; i < r; ++i) {
    clb = vmin(clb, ti[i].center);
    cub = vmax(cub, ti[i].center);
  }
  v3 range = cub - clb;
  int axis = 0;
  if (range.y > range.x) axis = 1;
  if (range.z > comp(range, axis)) axis = 2;
  uint32_t mid = l + (r - l) / 2;
  std::nth_element(ti.begin() + l, ti.begin() + mid, ti.begin() + r,
                   [=](const TriInfo& a, const TriInfo& b) { return comp(a.center, axis) < comp(b.center, axis); });
  uint32_t left = static_cast<uint32_t>(sc.blas.size());
  sc.blas[node_idx].child = left;
  sc.blas[node_idx].triangle_count = 0;
  sc.blas.emplace_back();
  sc.blas.emplace_back();
  blas_build_tree(sc, left, l, mid, ti, prim, verts, indices, tri_offset);
  blas_build_tree(sc, left + 1, mid, r, ti, prim, verts, indices, tri_offset);
}

// Emissive buffer texture factor (emissive_buffer.h:135-212): the decoded
// image (not the resized layer) at the triangle's centroid uv, nearest texel,
// under the texture's own wrap modes.  Declared: a NaN coordinate reads as 0.
float wrap_coord(float coord, int32_t wrap) {  // :191-204
  float r;
  if (wrap == CG_WRAP_CLAMP_TO_EDGE) {
    r = coord;
  } else if (wrap == CG_WRAP_MIRRORED_REPEAT) {
    r = coord - std::floor(coord);
    if (std::fmod(std::floor(coord), 2.0f) != 0.0f) r = 1.0f - r;  // odd tile
  } else {
    r = coord - std::floor(coord);
  }
  if (std::isnan(r)) r = 0.0f;
  return std::min(std::max(r, 0.0f), 0.999999f);
}
void emissive_texel(const cg_texture& t, float u, float v, float* rgb) {  // SampleTexture, :170-189
  uint32_t x = std::min(t.width - 1, static_cast<uint32_t>(wrap_coord(u, t.wrap_s) * static_cast<float>(t.width)));
  uint32_t y = std::min(t.height - 1, static_cast<uint32_t>(wrap_coord(v, t.wrap_t) * static_cast<float>(t.height)));
  const uint8_t* px = t.rgba8 + (static_cast<size_t>(y) * t.width + x) * 4;
  for (int c = 0; c < 3; ++c) rgb[c] = static_cast<float>(px[c]) / 255.0f;
}
void emissive_texture_factor(const cg_material& m, const cg_vertex& v0, const cg_vertex& v1, const cg_vertex& v2,
                             const cg_texture* textures, uint32_t texture_count, float* rgb) {  // :135-168
  rgb[0] = rgb[1] = rgb[2] = 1.0f;
  const cg_tex_info& ti = m.tex[CG_TEX_EMISSIVE];
  if (ti.index < 0 || static_cast<uint32_t>(ti.index) >= texture_count) return;
  const bool second = ti.tex_coord == 1u;
  const float* a = second ? v0.uv1 : v0.uv0;
  const float* b = second ? v1.uv1 : v1.uv0;
  const float* c = second ? v2.uv1 : v2.uv0;
  emissive_texel(textures[ti.index], (a[0] + b[0] + c[0]) / 3.0f, (a[1] + b[1] + c[1]) / 3.0f, rgb);
}

}  // namespace

extern "C" {

void orc_sincos(float x, float* s, float* c) { sincos_decl(x, *s, *c); }

void orc_texture_sample(const uint8_t* rgba8, uint32_t w, uint32_t h, uint32_t layers, int32_t layer, float u,
                        float v, float* out) {
  Rgba r = sample_texture(TexArr{rgba8, w, h, layers}, layer, f2{u, v});
  for (int k = 0; k < 4; ++k) out[k] = r.c[k];
}

void orc_emissive_texel(const cg_texture* t, float u, float v, float* rgb) { emissive_texel(*t, u, v, rgb); }

orc_scene* orc_scene_build(const cg_vertex* vertices, uint32_t vertex_count, const uint32_t* indices,
                           uint32_t index_count, const cg_material* materials, uint32_t material_count,
                           const cg_primitive* prims, uint32_t prim_count, const cg_texture* textures,
                           uint32_t texture_count, const float* env_rgba32f, uint32_t env_w, uint32_t env_h) {
  orc_scene* sc = new orc_scene();
  std::vector<cg_vertex> verts(vertices, vertices + vertex_count);
  std::vector<uint32_t> idx(indices, indices + index_count);
  sc->vert = verts;
  sc->mat.assign(materials, materials + material_count);

  // BLAS (blas.h:94-169): one tree per primitive object, in object order.
  std::vector<uint32_t> roots;
  sc->blas.reserve(static_cast<size_t>(index_count) / 3 * 2 + 16);
  for (uint32_t p = 0; p < prim_count; ++p) {
    const cg_primitive& pr = prims[p];
    uint32_t tcount = pr.index_count / 3;
    std::vector<TriInfo> ti(tcount);
    for (uint32_t t = 0; t < tcount; ++t) {
      uint32_t o = pr.index_offset + t * 3;
      v3 v0 = mk(verts[idx[o]].pos), v1 = mk(verts[idx[o + 1]].pos), v2 = mk(verts[idx[o + 2]].pos);
      ti[t] = TriInfo{(v0 + v1 + v2) / 3.0f, t};
    }
    roots.push_back(static_cast<uint32_t>(sc->blas.size()));
    sc->blas.emplace_back();
    uint32_t offset = static_cast<uint32_t>(sc->idx.size());
    blas_build_tree(*sc, static_cast<uint32_t>(sc->blas.size() - 1), 0, tcount, ti, pr, verts, idx, offset);
    for (const TriInfo& t : ti) {
      uint32_t o = pr.index_offset + t.index * 3;
      sc->idx.push_back(cg_index{idx[o], idx[o + 1], idx[o + 2], 0});
      sc->tri_ids.push_back(t.index);
    }
  }

  // TLAS (tlas.h:34-140).
  struct Bounded {
    v3 lb, ub, center;
    uint32_t prim_idx;
  };
  std::vector<Bounded> bp;
  for (uint32_t p = 0; p < prim_count; ++p) {
    const cg_primitive& pr = prims[p];
    cg_instance in{};
    std::memcpy(in.inv_trans, pr.inv_world, sizeof(in.inv_trans));
    in.blas_root_idx = roots[p];
    in.material_idx = pr.material_idx;
    sc->inst.push_back(in);
    // AABB{prim, vertices} (aabb.h:28-33) with the default ub = FLT_MIN.
    v3 alb = mk(FLT_MAX, FLT_MAX, FLT_MAX), aub = mk(FLT_MIN, FLT_MIN, FLT_MIN);
    for (uint32_t i = pr.vertex_offset; i < pr.vertex_offset + pr.vertex_count; ++i) {
      alb = vmin(alb, mk(verts[i].pos));
      aub = vmax(aub, mk(verts[i].pos));
    }
    // trans_aabb (tlas.h:38-55)
    const float* m = pr.world;
    float lb[3] = {m[12], m[13], m[14]}, ub[3] = {m[12], m[13], m[14]};
    float alb_[3] = {alb.x, alb.y, alb.z}, aub_[3] = {aub.x, aub.y, aub.z};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        float a = m[4 * j + i] * alb_[j];
        float b = m[4 * j + i] * aub_[j];
        if (a < b) {
          lb[i] += a;
          ub[i] += b;
        } else {
          lb[i] += b;
          ub[i] += a;
        }
      }
    Bounded b;
    b.lb = mk(lb);
    b.ub = mk(ub);
    b.center = (b.lb + b.ub) * 0.5f;
    b.prim_idx = p;
    bp.push_back(b);
  }
  if (prim_count > 0) {
    sc->tlas.reserve(static_cast<size_t>(prim_count) * 2);
    sc->tlas.emplace_back();
    // Build (tlas.h:98-140), recursive.
    struct Rec {
      static void build(orc_scene& s, uint32_t ni, uint32_t l, uint32_t r, std::vector<Bounded>& b) {
        v3 lb = mk(FLT_MAX, FLT_MAX, FLT_MAX), ub = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        for (uint32_t i = l; i < r; ++i) {
          lb = vmin(lb, b[i].lb);
          ub = vmax(ub, b[i].ub);
        }
        v3 pad = mk(0.001f, 0.001f, 0.001f);
        lb = lb - pad;
        ub = ub + pad;
        cg_tlas_node& n = s.tlas[ni];
        n.lb[0] = lb.x; n.lb[1] = lb.y; n.lb[2] = lb.z;
        n.ub[0] = ub.x; n.ub[1] = ub.y; n.ub[2] = ub.z;
        uint32_t count = r - l;
        if (count <= 1) {
          n.instance_idx = b[l].prim_idx;
          n.child = 0;
          return;
        }
        // AABB center_bounds{} — default ub = FLT_MIN (aabb.h:17-22)
        v3 clb = mk(FLT_MAX, FLT_MAX, FLT_MAX), cub = mk(FLT_MIN, FLT_MIN, FLT_MIN);
        for (uint32_t i = l; i < r; ++i) {
          clb = vmin(clb, b[i].center);
          cub = vmax(cub, b[i].center);
        }
        v3 range = cub - clb;
        int axis = 0;
        if (range.y > range.x) axis = 1;
        if (range.z > comp(range, axis)) axis = 2;
        uint32_t mid = l + (r - l) / 2;
        std::nth_element(b.begin() + l, b.begin() + mid, b.begin() + r, [=](const Bounded& x, const Bounded& y) {
          return comp(x.center, axis) < comp(y.center, axis);
        });
        uint32_t lc = static_cast<uint32_t>(s.tlas.size());
        s.tlas[ni].child = lc;
        s.tlas[ni].instance_idx = 0;
        s.tlas.emplace_back();
        s.tlas.emplace_back();
        build(s, lc, l, mid, b);
        build(s, lc + 1, mid, r, b);
      }
    };
    Rec::build(*sc, 0, 0, prim_count, bp);
  }

  // Emissive buffer (emissive_buffer.h:44-100).
  float total = 0.0f;
  for (uint32_t p = 0; p < prim_count; ++p) {
    const cg_primitive& pr = prims[p];
    if (pr.material_idx >= material_count) continue;
    const cg_material& m = materials[pr.material_idx];
    v3 base = mk(fmx(m.emissive_factor[0] * m.emissive_strength, 0.0f),
                 fmx(m.emissive_factor[1] * m.emissive_strength, 0.0f),
                 fmx(m.emissive_factor[2] * m.emissive_strength, 0.0f));
    auto lum = [](v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; };
    if (lum(base) <= 0.0f) continue;
    for (uint32_t i = 0; i + 2 < pr.index_count; i += 3) {
      uint32_t o = pr.index_offset + i;
      v3 p0 = glm_point(pr.world, mk(verts[idx[o]].pos));
      v3 p1 = glm_point(pr.world, mk(verts[idx[o + 1]].pos));
      v3 p2 = glm_point(pr.world, mk(verts[idx[o + 2]].pos));
      float area = length(cross(p1 - p0, p2 - p0)) * 0.5f;
      if (area <= 0.0f) continue;
      float f[3];
      emissive_texture_factor(m, verts[idx[o]], verts[idx[o + 1]], verts[idx[o + 2]], textures, texture_count, f);
      v3 em = base * mk(f[0], f[1], f[2]);
      float w = lum(em);
      if (w <= 0.0f) continue;
      cg_emissive_prim e{};
      e.v[0][0] = p0.x; e.v[0][1] = p0.y; e.v[0][2] = p0.z;
      e.v[1][0] = p1.x; e.v[1][1] = p1.y; e.v[1][2] = p1.z;
      e.v[2][0] = p2.x; e.v[2][1] = p2.y; e.v[2][2] = p2.z;
      e.emission[0] = em.x; e.emission[1] = em.y; e.emission[2] = em.z;
      e.area = area;
      e.weight = area * w;
      sc->em.push_back(e);
      total += e.weight;
    }
  }
  if (total > 0.0f)
    for (auto& e : sc->em) e.weight /= total;

  // Alias table (alias_table.h:24-70).
  std::vector<double> w;
  for (auto& e : sc->em) w.push_back(static_cast<double>(e.weight));
  sc->alias.resize(w.size());
  orc_alias_build(w.data(), static_cast<uint32_t>(w.size()), sc->alias.data());

  // Material texture array (resource.cpp:471-546): every image resized
  // nearest to the largest width and height (ResizeRgba8Nearest, :51-75).
  if (texture_count > 0) {
    uint32_t W = 1, H = 1;
    for (uint32_t t = 0; t < texture_count; ++t) {
      W = std::max(W, textures[t].width);
      H = std::max(H, textures[t].height);
    }
    sc->tex_w = W;
    sc->tex_h = H;
    sc->tex_layers = texture_count;
    sc->tex.resize(static_cast<size_t>(W) * H * 4 * texture_count);
    for (uint32_t t = 0; t < texture_count; ++t) {
      const cg_texture& img = textures[t];
      uint8_t* dst = &sc->tex[static_cast<size_t>(t) * W * H * 4];
      for (uint32_t y = 0; y < H; ++y) {
        uint32_t sy = img.height == H ? y : std::min(img.height - 1, static_cast<uint32_t>(uint64_t{y} * img.height / H));
        for (uint32_t x = 0; x < W; ++x) {
          uint32_t sx = img.width == W ? x : std::min(img.width - 1, static_cast<uint32_t>(uint64_t{x} * img.width / W));
          for (int c = 0; c < 4; ++c)
            dst[(static_cast<size_t>(y) * W + x) * 4 + c] = img.rgba8[(static_cast<size_t>(sy) * img.width + sx) * 4 + c];
        }
      }
    }
  }

  // Environment (scene_data.h:57-98): given texture, else the fallback sky.
  if (env_rgba32f && env_w && env_h) {
    sc->env.assign(env_rgba32f, env_rgba32f + static_cast<size_t>(env_w) * env_h * 4);
    sc->env_w = env_w;
    sc->env_h = env_h;
  } else {
    sc->env_w = 512;
    sc->env_h = 256;
    sc->env.resize(512 * 256 * 4);
    for (uint32_t y = 0; y < 256; ++y) {
      float t = static_cast<float>(y) / static_cast<float>(256 - 1);
      v3 a = mk(0.7f, 0.8f, 1.0f), b = mk(0.05f, 0.05f, 0.06f);
      v3 c = a * (1.0f - t) + b * t;
      for (uint32_t x = 0; x < 512; ++x) {
        float* px = &sc->env[(static_cast<size_t>(y) * 512 + x) * 4];
        px[0] = c.x;
        px[1] = c.y;
        px[2] = c.z;
        px[3] = 1.0f;
      }
    }
  }
  return sc;
}

orc_scene* orc_scene_wrap(const cg_scene_data* d) {
  orc_scene* sc = new orc_scene();
  sc->tlas.assign(d->tlas_nodes, d->tlas_nodes + d->tlas_node_count);
  sc->inst.assign(d->instances, d->instances + d->instance_count);
  sc->blas.assign(d->blas_nodes, d->blas_nodes + d->blas_node_count);
  sc->idx.assign(d->indices, d->indices + d->index_count);
  if (d->tri_ids) sc->tri_ids.assign(d->tri_ids, d->tri_ids + d->index_count);
  sc->vert.assign(d->vertices, d->vertices + d->vertex_count);
  sc->mat.assign(d->materials, d->materials + d->material_count);
  if (d->emissive_count) sc->em.assign(d->emissive, d->emissive + d->emissive_count);
  if (d->alias_count) sc->alias.assign(d->alias, d->alias + d->alias_count);
  sc->env.assign(d->env_rgba32f, d->env_rgba32f + static_cast<size_t>(d->env_width) * d->env_height * 4);
  sc->env_w = d->env_width;
  sc->env_h = d->env_height;
  if (d->tex_layers && d->tex_rgba8) {
    sc->tex.assign(d->tex_rgba8, d->tex_rgba8 + static_cast<size_t>(d->tex_width) * d->tex_height * 4 * d->tex_layers);
    sc->tex_w = d->tex_width;
    sc->tex_h = d->tex_height;
    sc->tex_layers = d->tex_layers;
  }
  return sc;
}

void orc_scene_free(orc_scene* s) { delete s; }

void orc_scene_get(const orc_scene* s, cg_scene_data* o) {
  o->tlas_nodes = s->tlas.data();
  o->tlas_node_count = static_cast<uint32_t>(s->tlas.size());
  o->instances = s->inst.data();
  o->instance_count = static_cast<uint32_t>(s->inst.size());
  o->blas_nodes = s->blas.data();
  o->blas_node_count = static_cast<uint32_t>(s->blas.size());
  o->indices = s->idx.data();
  o->index_count = static_cast<uint32_t>(s->idx.size());
  o->tri_ids = s->tri_ids.data();
  o->vertices = s->vert.data();
  o->vertex_count = static_cast<uint32_t>(s->vert.size());
  o->materials = s->mat.data();
  o->material_count = static_cast<uint32_t>(s->mat.size());
  o->emissive = s->em.data();
  o->emissive_count = static_cast<uint32_t>(s->em.size());
  o->alias = s->alias.data();
  o->alias_count = static_cast<uint32_t>(s->alias.size());
  o->env_rgba32f = s->env.data();
  o->env_width = s->env_w;
  o->env_height = s->env_h;
  o->tex_rgba8 = s->tex.empty() ? nullptr : s->tex.data();
  o->tex_width = s->tex_w;
  o->tex_height = s->tex_h;
  o->tex_layers = s->tex_layers;
}

void orc_trace_closest(const orc_scene* s, const cg_ray* rays, uint32_t n, cg_hit* hits, float* uv,
                       float* second_t, uint32_t* counts, int cull, int threads) {
  parallel_for((n + 1023) / 1024, threads, [&](uint32_t chunk) {
    uint32_t b = chunk * 1024, e = std::min(n, b + 1024);
    for (uint32_t i = b; i < e; ++i) {
      Counts c;
      float st = MAXD;
      float uvv[2] = {0, 0};
      Hit h = ray_hit(*s, mk(rays[i].org), mk(rays[i].dir), c, cull != 0, &st, uvv);
      hits[i].t = h.hit ? h.dist : CG_MISS_T;
      hits[i].inst = h.inst;
      hits[i].tri = h.tri;
      hits[i]._pad = 0;
      if (uv) {
        uv[2 * i] = uvv[0];
        uv[2 * i + 1] = uvv[1];
      }
      if (second_t) second_t[i] = st;
      if (counts) {
        counts[4 * i + 0] = c.tlas_nodes;
        counts[4 * i + 1] = c.blas_nodes;
        counts[4 * i + 2] = c.tris;
        counts[4 * i + 3] = c.insts;
      }
    }
  });
}

int orc_render(const orc_scene* s, const cg_conf* conf, const cg_camera* cam, uint32_t iter_count,
               cg_history_pixel* history, uint8_t* rgba8, uint32_t* path_hash, uint32_t* ray_count, uint32_t y0,
               uint32_t y1, int threads, int cull) {
  return orc_render_mode(s, conf, cam, iter_count, history, rgba8, path_hash, ray_count, y0, y1, threads, cull, 0);
}

// Sample-stream mode (declared for this build; the reference has one stream
// per pixel): sample i of a pixel draws from its own stream, seeded from the
// pixel's base seed b (the pixel index on the first frame, the carried seed
// later) as step(step(b) ^ (i + 1) * 0x9E3779B9); within the sample the
// draws follow the reference's depth-first order.  The carried seed is
// step(step(b)) (the "i = -1" stream, distinct from every sample's because
// step is a bijection).  The samples' XYZ values are summed in sample order,
// so the image is the same however the samples are spread over workers; the
// path hash folds each sample's own hash in sample order.
uint32_t orc_sample_seed(uint32_t base, uint32_t i) {
  uint32_t t = base;
  step_rng(t);
  t ^= (i + 1u) * 0x9E3779B9u;
  step_rng(t);
  return t;
}
uint32_t orc_sample_carry(uint32_t base) {
  uint32_t t = base;
  step_rng(t);
  step_rng(t);
  return t;
}

int orc_render_mode(const orc_scene* s, const cg_conf* conf, const cg_camera* cam, uint32_t iter_count,
                    cg_history_pixel* history, uint8_t* rgba8, uint32_t* path_hash, uint32_t* ray_count, uint32_t y0,
                    uint32_t y1, int threads, int cull, int rng_mode) {
  const uint32_t W = conf->render_width, H = conf->render_height;
  Params P;
  P.D = static_cast<int>(std::min(std::max(conf->trace_depth, 1u), 64u));
  P.T = conf->max_transport_sample_count;
  P.E = conf->max_emission_sample_count;
  P.DIF = conf->max_diffuse_sample_count;
  P.ROUGH = conf->max_rough_sample_count;
  y1 = std::min(y1, H);
  if (y0 >= y1) return 0;
  parallel_for(y1 - y0, threads, [&](uint32_t row) {
    uint32_t py = y0 + row;
    for (uint32_t px = 0; px < W; ++px) {
      v3 pos, dir;
      camera_ray(*cam, W, H, px, py, pos, dir);
      uint32_t hidx = px + py * W;
      float w = 1.0f / static_cast<float>(iter_count + 1u);
      v3 lum = mk(0.0f, 0.0f, 0.0f);
      uint32_t seed = px + py * W;
      if (iter_count > 0u) {
        lum = mk(history[hidx].xyz) * (1.0f - w);
        seed = history[hidx].rand_seed;
      }
      uint32_t rng = seed;
      step_rng(rng);  // seedRng
      uint32_t hash = 2166136261u, rays = 0;
      // render.wesl:7-18
      v3 acc = mk(0.0f, 0.0f, 0.0f);
      for (uint32_t i = 0; i < conf->sample_count; ++i) {
        uint32_t sample_hash = 2166136261u;
        if (rng_mode == 1) rng = orc_sample_seed(seed, i);
        uint32_t& h = rng_mode == 1 ? sample_hash : hash;
        float lambda, prob;
        sample_wavelength_cdf(rand_float(rng), lambda, prob);
        Tracer tr{*s, P, lambda, rng, h, rays, cull != 0, {}};
        float radiance = tr.trace(pos, dir, 0);
        v3 il = illuminate(lambda, radiance);
        float spp = static_cast<float>(conf->sample_count);
        acc = acc + mk(il.x / prob / spp, il.y / prob / spp, il.z / prob / spp);
        if (rng_mode == 1) hash = hmix(hash, sample_hash);
      }
      lum = lum + acc * w;
      history[hidx].xyz[0] = lum.x;
      history[hidx].xyz[1] = lum.y;
      history[hidx].xyz[2] = lum.z;
      history[hidx].rand_seed = rng_mode == 1 ? orc_sample_carry(seed) : rand_uint(rng);
      if (rgba8) {
        v3 c = xyz_to_srgb(lum);
        uint8_t* o = rgba8 + static_cast<size_t>(hidx) * 4;
        o[0] = unorm8(c.x);
        o[1] = unorm8(c.y);
        o[2] = unorm8(c.z);
        o[3] = 255;
      }
      if (path_hash) path_hash[hidx] = hash;
      if (ray_count) ray_count[hidx] = rays;
    }
  });
  return 0;
}

uint32_t orc_rng_step(uint32_t state) {
  step_rng(state);
  return state;
}
float orc_rand_float(uint32_t* state) { return rand_float(*state); }
float orc_erf(float x) { return erf_approx(x); }
void orc_xyz_match(float l, float* xyz) {
  v3 m = xyz_match(l);
  xyz[0] = m.x;
  xyz[1] = m.y;
  xyz[2] = m.z;
}
float orc_xyz_match_sum_integral(float lo, float hi) { return xyz_match_sum_integral(lo, hi); }
void orc_sample_wavelength(float u, float* lambda, float* prob) { sample_wavelength_cdf(u, *lambda, *prob); }
void orc_illuminate(float lambda, float radiance, float* xyz) {
  v3 m = illuminate(lambda, radiance);
  xyz[0] = m.x;
  xyz[1] = m.y;
  xyz[2] = m.z;
}
float orc_smits(int curve, float lambda) { return smits(curve, lambda); }
float orc_rgb_to_reflectance(const float* rgb, float l) { return rgb_to_reflectance(mk(rgb), l); }
float orc_rgb_to_radiance(const float* rgb, float l) { return rgb_to_radiance(mk(rgb), l); }
float orc_dispersed_ior(float nd, float disp, float l) { return dispersed_ior(nd, disp, l); }
float orc_fresnel(const float* dir, const float* normal, float re, float im) {
  return fresnel(mk(dir), mk(normal), IOR{re, im});
}

// alias_table.h:24-70 — Vose's method in double; probabilities stored as
// float and the float value is what the donor loses.
void orc_alias_build(const double* weights, uint32_t n, cg_alias_entry* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = cg_alias_entry{0u, 0.0f};
  if (n == 0) return;
  std::vector<double> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = std::max(0.0, weights[i]);
  double total = std::accumulate(p.begin(), p.end(), 0.0);
  if (total <= 0.0) return;
  std::vector<uint32_t> large, small;
  for (uint32_t i = 0; i < n; ++i) {
    p[i] = p[i] / total * n;
    if (p[i] >= 1.0)
      large.push_back(i);
    else
      small.push_back(i);
  }
  while (!large.empty() && !small.empty()) {
    uint32_t i = small.back(), j = large.back();
    small.pop_back();
    large.pop_back();
    out[i].prob = static_cast<float>(p[i]);
    out[i].alias = j;
    p[j] -= 1.0 - out[i].prob;
    if (p[j] < 1.0)
      small.push_back(j);
    else
      large.push_back(j);
  }
  while (!large.empty()) {
    uint32_t i = large.back();
    large.pop_back();
    out[i].prob = 1.0f;
    out[i].alias = i;
  }
  while (!small.empty()) {
    uint32_t i = small.back();
    small.pop_back();
    out[i].prob = 1.0f;
    out[i].alias = i;
  }
}
uint32_t orc_alias_pick(const cg_alias_entry* table, uint32_t n, float u1, float u2) {
  uint32_t bin = f2u(u1 * static_cast<float>(n));
  uint32_t b = std::min(bin, n - 1u);
  return u2 < table[b].prob ? bin : table[b].alias;
}
void orc_srgb(const float* xyz, float* rgb) {
  v3 c = xyz_to_srgb(mk(xyz));
  rgb[0] = c.x;
  rgb[1] = c.y;
  rgb[2] = c.z;
}
void orc_camera_ray(const cg_camera* cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py, cg_ray* out) {
  v3 p, d;
  camera_ray(*cam, w, h, px, py, p, d);
  out->org[0] = p.x;
  out->org[1] = p.y;
  out->org[2] = p.z;
  out->_pad0 = 0;
  out->dir[0] = d.x;
  out->dir[1] = d.y;
  out->dir[2] = d.z;
  out->_pad1 = 0;
}
int orc_intersect_triangle(const cg_ray* ray, const float* v0, const float* v1, const float* v2, float* tuv) {
  TriHit h = ray_triangle(mk(ray->org), mk(ray->dir), mk(v0), mk(v1), mk(v2));
  tuv[0] = h.t;
  tuv[1] = h.u;
  tuv[2] = h.v;
  return h.t > 0.0f;
}
int orc_intersect_aabb(const cg_ray* ray, const float* lb, const float* ub) {
  float tn;
  return ray_aabb(mk(ray->org), dir_inv(mk(ray->dir)), lb, ub, &tn) ? 1 : 0;
}
float orc_sample_lobe(int model, const float* curr_dir, const float* shading_norm, float alpha, float ior_re,
                      float ior_im, float transmission, uint32_t* rng_state, float* next_dir) {
  Surface s{};
  s.sn = mk(shading_norm);
  s.gn = s.sn;
  Mat m{};
  m.alpha = alpha;
  m.ior = IOR{ior_re, ior_im};
  m.transmission = transmission;
  m.diffuse = 1.0f;
  v3 d = mk(curr_dir);
  TSample t{mk(0, 0, 0), 0.0f};
  if (model == M_DIFFUSE) t = diffuse_sample(d, s, *rng_state);
  if (model == M_ROUGH_R) t = rough_reflect_sample(d, s, m, *rng_state);
  if (model == M_ROUGH_T) t = rough_trans_sample(d, s, m, *rng_state);
  next_dir[0] = t.dir.x;
  next_dir[1] = t.dir.y;
  next_dir[2] = t.dir.z;
  return t.prob;
}


// ---------------------------------------------------------------------------
// Unit-level pieces of the shading half, for the closed-form and statistical
// checks in tests/test_oracle.py.
}  // extern "C"
namespace {
Mat unit_mat(const orc_mat* m) {
  Mat x{};
  x.diffuse = m->diffuse;
  x.alpha = m->alpha;
  x.ior = IOR{m->ior_re, m->ior_im};
  x.transmission = m->transmission;
  x.volume = m->volume != 0u;
  return x;
}
Surface unit_surface(const float* pos, const float* sn) {
  Surface s{};
  s.pos = pos ? mk(pos) : mk(0, 0, 0);
  s.sn = mk(sn);
  s.gn = s.sn;
  return s;
}
void put3(float* out, v3 v) {
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
}
}  // namespace
extern "C" {

float orc_lobe_sample(int model, const float* curr_dir, const float* shading_norm, const orc_mat* mat,
                      uint32_t* rng_state, float* next_dir) {
  const Surface s = unit_surface(nullptr, shading_norm);
  const Mat m = unit_mat(mat);
  const v3 d = mk(curr_dir);
  TSample t{mk(0, 0, 0), 0.0f};
  switch (model) {
    case M_DIFFUSE: t = diffuse_sample(d, s, *rng_state); break;
    case M_SPEC_R: {  // sample.wesl:274-291 with specularReflectSample
      v3 nd = specular_dir(d, s);
      t = TSample{nd, spec_reflect_pdf(d, nd, s)};
      break;
    }
    case M_ROUGH_R: t = rough_reflect_sample(d, s, m, *rng_state); break;
    case M_SPEC_T: {  // sample.wesl:312-329 with specularTransmitSample
      v3 nd = m.volume ? spec_trans_dir(d, s, m) : d;
      t = TSample{nd, spec_trans_pdf(d, nd, s, m)};
      break;
    }
    case M_ROUGH_T: t = rough_trans_sample(d, s, m, *rng_state); break;
    default: break;
  }
  put3(next_dir, t.dir);
  return t.prob;
}

float orc_lobe_eval(int model, const float* curr_dir, const float* shading_norm, const orc_mat* mat,
                    const float* next_dir, float* dist) {
  const Surface s = unit_surface(nullptr, shading_norm);
  const Mat m = unit_mat(mat);
  const v3 d = mk(curr_dir), n = mk(next_dir);
  float pdf = 0.0f, f = 0.0f;
  switch (model) {
    case M_DIFFUSE: pdf = diffuse_pdf(d, n, s); f = diffuse_dist(d, n, s, m); break;
    case M_SPEC_R: pdf = spec_reflect_pdf(d, n, s); f = spec_reflect_dist(d, n, s, m); break;
    case M_ROUGH_R: pdf = rough_reflect_pdf(d, n, s, m); f = rough_reflect_dist(d, n, s, m); break;
    case M_SPEC_T: pdf = spec_trans_pdf(d, n, s, m); f = spec_trans_dist(d, n, s, m); break;
    case M_ROUGH_T: pdf = rough_trans_pdf(d, n, s, m); f = rough_trans_dist(d, n, s, m); break;
    default: break;
  }
  if (dist) *dist = f;
  return pdf;
}

float orc_ggx_d(const float* micro, const float* macro, float alpha) { return ggx_d(mk(micro), mk(macro), alpha); }
float orc_ggx_g1(float c, float alpha) { return ggx_g1(c, alpha); }
float orc_ggx_g2(float c1, float c2, float alpha) { return ggx_g2(c1, c2, alpha); }
void orc_sample_micro_normal(const float* curr_dir, const float* shading_norm, float alpha, uint32_t* rng_state,
                             float* micro) {
  put3(micro, sample_micro_normal(mk(curr_dir), unit_surface(nullptr, shading_norm), alpha, *rng_state));
}
void orc_tangent_space(const float* n, float* t, float* b) {
  Frame f = tangent_space(mk(n));
  put3(t, f.t);
  put3(b, f.b);
}
void orc_reflect(const float* d, const float* n, float* out) { put3(out, reflect(mk(d), mk(n))); }
void orc_refract(const float* d, const float* n, float eta, float* out) { put3(out, refract(mk(d), mk(n), eta)); }

void orc_sample_counts(const orc_mat* mat, int can, int depth, const cg_conf* conf, uint32_t* out) {
  Params P;
  P.D = std::max(1, std::min(static_cast<int>(conf->trace_depth), 64));
  P.T = conf->max_transport_sample_count;
  P.E = conf->max_emission_sample_count;
  P.DIF = conf->max_diffuse_sample_count;
  P.ROUGH = conf->max_rough_sample_count;
  Counts6 c = sample_counts(can != 0, unit_mat(mat), depth, P);
  out[0] = c.d;
  out[1] = c.sr;
  out[2] = c.rr;
  out[3] = c.st;
  out[4] = c.rt;
  out[5] = c.l;
}

float orc_balance_denominator(const orc_scene* sc, const float* curr_dir, const float* pos, const float* shading_norm,
                              const orc_mat* mat, const uint32_t* counts, int model, float model_pdf,
                              float sampled_light_pdf, const float* next_dir) {
  Counts6 c;
  c.d = counts[0];
  c.sr = counts[1];
  c.rr = counts[2];
  c.st = counts[3];
  c.rt = counts[4];
  c.l = counts[5];
  return balance_denominator(*sc, mk(curr_dir), mk(next_dir), unit_surface(pos, shading_norm), unit_mat(mat), c, model,
                             model_pdf, sampled_light_pdf);
}

float orc_light_pdf(const orc_scene* sc, const float* pos, const float* next_dir) {
  return light_pdf(*sc, mk(next_dir), unit_surface(pos, next_dir));
}

float orc_emission_sample(const orc_scene* sc, float lambda, const float* pos, uint32_t* rng_state, float* next_dir,
                          float* expected_dist, float* radiance, uint32_t* prim) {
  const float up[3] = {0.0f, 0.0f, 1.0f};  // emissionSample reads only the surface position
  ESample e = emission_sample(*sc, lambda, unit_surface(pos, up), *rng_state);
  if (prim) *prim = can_sample(*sc) ? e.prim : 0xFFFFFFFFu;
  put3(next_dir, e.dir);
  *expected_dist = e.expected_dist;
  *radiance = e.radiance;
  return e.prob;
}

void orc_spawn(const float* pos, const float* geom_norm, const float* next_dir, float* out_pos) {
  Surface s{};
  s.pos = mk(pos);
  s.gn = mk(geom_norm);
  s.sn = s.gn;
  v3 p, d;
  spawn(s, mk(next_dir), p, d);
  put3(out_pos, p);
}

}  // extern "C"
