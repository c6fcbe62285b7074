// scene_gen.cpp — seeded procedural scenes standing in for the glTF files the
// reference loads (src/scene.cpp:829-1003), and the ray batch of the
// traversal microbenchmark.  Every draw comes from one PCG32 stream seeded by
// SplitMix64(seed), consumed in the order written here.
//
//   "c0"      ground quad 8x8 m, 64 icospheres (subdivision 2, 320 triangles,
//             one mesh and material each) at uniform positions in a 4x4x2 m
//             box with radius U(0.1, 0.4), one emissive quad (2 triangles,
//             strength 10, 6500 K colour); 20,484 triangles.
//   "c1"      16 instances of one 131,072-triangle octahedral sphere displaced
//             by 3-octave value noise (amplitude 0.1), random rotation,
//             translation and scale U(0.5, 2); 4096 one-triangle emitters with
//             strength log-uniform on [1, 100].  params[0] (optional) scales
//             the sphere's subdivision n = 128.
//   "c2"      one BLAS of params[0] (default 10,000,000) triangles in 1024
//             Gaussian clusters (sigma 0.02) inside the unit cube, edge
//             lengths U(0.001, 0.01).
//   "quad"    one diffuse quad under the sky (a hand-checkable scene).
//   "spheres" params[0] spheres (default 8), for small parity cases.
// "c1" and "spheres" take params[1] > 0 to add four procedural material
// textures (base colour, metallic-roughness, transmission, emission) of side
// params[2] (c1: 2048 each; spheres: 64, with layers of different sizes and
// wrap modes) and reference them from the materials.
// Materials are drawn uniformly from {diffuse albedo U(0.2, 0.9), rough metal
// GGX alpha U(0.05, 0.5), dispersive glass IOR 1.5}; glTF's dispersion for
// Cauchy B = 0.004 um^2 at n_d = 1.5 is 20 B / ((n_d - 1) 523655 nm^2) = 0.30554.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "scene_impl.h"

namespace crystal::graphics {
namespace {

struct Pcg32 {
  uint64_t state = 0, inc = 0;
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Pcg32(uint64_t seed) {
    uint64_t s = seed;
    state = splitmix(s);
    inc = splitmix(s) | 1ull;
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  double uniform() { return (next() >> 8) * (1.0 / 16777216.0); }  // [0, 1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {  // Box-Muller
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct D3 {
  double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 dnorm(D3 a) { return a * (1.0 / std::sqrt(ddot(a, a))); }

// Column-major 4x4 in double.
struct M4 {
  double m[16];
};
M4 identity() {
  M4 r{};
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
  return r;
}
M4 mul(const M4& a, const M4& b) {
  M4 r{};
  for (int c = 0; c < 4; ++c)
    for (int rr = 0; rr < 4; ++rr) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a.m[4 * k + rr] * b.m[4 * c + k];
      r.m[4 * c + rr] = s;
    }
  return r;
}
// Affine inverse in double: [R t]^-1 = [R^-1, -R^-1 t].
M4 affine_inverse(const M4& a) {
  double r[9] = {a.m[0], a.m[4], a.m[8], a.m[1], a.m[5], a.m[9], a.m[2], a.m[6], a.m[10]};  // row-major 3x3
  double det = r[0] * (r[4] * r[8] - r[5] * r[7]) - r[1] * (r[3] * r[8] - r[5] * r[6]) +
               r[2] * (r[3] * r[7] - r[4] * r[6]);
  double inv[9] = {(r[4] * r[8] - r[5] * r[7]) / det, (r[2] * r[7] - r[1] * r[8]) / det,
                   (r[1] * r[5] - r[2] * r[4]) / det, (r[5] * r[6] - r[3] * r[8]) / det,
                   (r[0] * r[8] - r[2] * r[6]) / det, (r[2] * r[3] - r[0] * r[5]) / det,
                   (r[3] * r[7] - r[4] * r[6]) / det, (r[1] * r[6] - r[0] * r[7]) / det,
                   (r[0] * r[4] - r[1] * r[3]) / det};
  M4 o = identity();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.m[4 * j + i] = inv[3 * i + j];
  for (int i = 0; i < 3; ++i)
    o.m[12 + i] = -(inv[3 * i + 0] * a.m[12] + inv[3 * i + 1] * a.m[13] + inv[3 * i + 2] * a.m[14]);
  return o;
}
M4 translate(D3 t) {
  M4 r = identity();
  r.m[12] = t.x;
  r.m[13] = t.y;
  r.m[14] = t.z;
  return r;
}
M4 scale(double s) {
  M4 r = identity();
  r.m[0] = r.m[5] = r.m[10] = s;
  return r;
}
M4 rotation(Pcg32& rng) {  // uniform random rotation (Shoemake)
  double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
  double a = std::sqrt(1 - u1), b = std::sqrt(u1);
  double qx = a * std::sin(6.283185307179586 * u2), qy = a * std::cos(6.283185307179586 * u2);
  double qz = b * std::sin(6.283185307179586 * u3), qw = b * std::cos(6.283185307179586 * u3);
  M4 r = identity();
  r.m[0] = 1 - 2 * (qy * qy + qz * qz);
  r.m[1] = 2 * (qx * qy + qz * qw);
  r.m[2] = 2 * (qx * qz - qy * qw);
  r.m[4] = 2 * (qx * qy - qz * qw);
  r.m[5] = 1 - 2 * (qx * qx + qz * qz);
  r.m[6] = 2 * (qy * qz + qx * qw);
  r.m[8] = 2 * (qx * qz + qy * qw);
  r.m[9] = 2 * (qy * qz - qx * qw);
  r.m[10] = 1 - 2 * (qx * qx + qy * qy);
  return r;
}

struct Builder {
  Scene::Impl& s;
  explicit Builder(Scene::Impl& sc) : s(sc) {
    s.vertices.clear();
    s.indices.clear();
    s.prims.clear();
    s.textures.clear();
    s.materials.assign(1, DefaultMaterial());  // default material at index 0 (scene.cpp:869-872)
  }
  uint32_t material(const cg_material& m) {
    s.materials.push_back(m);
    return static_cast<uint32_t>(s.materials.size() - 1);
  }
  // Append a mesh; returns (vertex_offset, vertex_count, index_offset, index_count).
  std::tuple<uint32_t, uint32_t, uint32_t, uint32_t> mesh(const std::vector<D3>& pos, const std::vector<D3>& nrm,
                                                          const std::vector<uint32_t>& tri) {
    uint32_t vo = static_cast<uint32_t>(s.vertices.size()), io = static_cast<uint32_t>(s.indices.size());
    for (size_t i = 0; i < pos.size(); ++i) {
      cg_vertex v{};
      v.pos[0] = static_cast<float>(pos[i].x);
      v.pos[1] = static_cast<float>(pos[i].y);
      v.pos[2] = static_cast<float>(pos[i].z);
      v.nrm[0] = static_cast<float>(nrm[i].x);
      v.nrm[1] = static_cast<float>(nrm[i].y);
      v.nrm[2] = static_cast<float>(nrm[i].z);
      D3 d = dnorm(pos[i]);
      v.uv0[0] = static_cast<float>(0.5 + std::atan2(d.y, d.x) / 6.283185307179586);
      v.uv0[1] = static_cast<float>(std::acos(std::max(-1.0, std::min(1.0, d.z))) / 3.141592653589793);
      v.uv1[0] = 2.0f * v.uv0[0] - 0.5f;
      v.uv1[1] = 3.0f * v.uv0[1] + 0.25f;
      s.vertices.push_back(v);
    }
    for (uint32_t t : tri) s.indices.push_back(vo + t);
    return {vo, static_cast<uint32_t>(pos.size()), io, static_cast<uint32_t>(tri.size())};
  }
  void prim(std::tuple<uint32_t, uint32_t, uint32_t, uint32_t> m, uint32_t material, const M4& world) {
    cg_primitive p{};
    std::tie(p.vertex_offset, p.vertex_count, p.index_offset, p.index_count) = m;
    p.material_idx = material;
    M4 inv = affine_inverse(world);
    for (int i = 0; i < 16; ++i) {
      p.world[i] = static_cast<float>(world.m[i]);
      p.inv_world[i] = static_cast<float>(inv.m[i]);
    }
    s.prims.push_back(p);
  }
};

// Linear sRGB of a blackbody at `kelvin`, normalised to a maximum of 1:
// Planck's law against the CIE 1931 observer (the Wyman-Sloan-Shirley
// multi-lobe fit), integrated in FP64 over 360-830 nm.
void blackbody_srgb(double kelvin, float out[3]) {
  auto g = [](double l, double mu, double s1, double s2) {
    double t = (l - mu) / (l < mu ? s1 : s2);
    return std::exp(-0.5 * t * t);
  };
  double X = 0, Y = 0, Z = 0;
  for (double l = 360.0; l <= 830.0; l += 0.5) {
    double lm = l * 1e-9;
    double B = 1.0 / (std::pow(lm, 5) * (std::exp(1.438776877e-2 / (lm * kelvin)) - 1.0));
    double x = 1.056 * g(l, 599.8, 37.9, 31.0) + 0.362 * g(l, 442.0, 16.0, 26.7) - 0.065 * g(l, 501.1, 20.4, 26.2);
    double y = 0.821 * g(l, 568.8, 46.9, 40.5) + 0.286 * g(l, 530.9, 16.3, 31.1);
    double z = 1.217 * g(l, 437.0, 11.8, 36.0) + 0.681 * g(l, 459.0, 26.0, 13.8);
    X += B * x;
    Y += B * y;
    Z += B * z;
  }
  double r = 3.2404542 * X - 1.5371385 * Y - 0.4985314 * Z;
  double gg = -0.9692660 * X + 1.8760108 * Y + 0.0415560 * Z;
  double b = 0.0556434 * X - 0.2040259 * Y + 1.0572252 * Z;
  double mx = std::max(r, std::max(gg, b));
  out[0] = static_cast<float>(std::max(r, 0.0) / mx);
  out[1] = static_cast<float>(std::max(gg, 0.0) / mx);
  out[2] = static_cast<float>(std::max(b, 0.0) / mx);
}

constexpr float kGlassDispersion = 0.30554f;

cg_material random_material(Pcg32& rng) {
  cg_material m = DefaultMaterial();
  const uint32_t kind = rng.next() % 3u;
  if (kind == 0) {  // diffuse
    float a = static_cast<float>(rng.uniform(0.2, 0.9));
    m.base_color_factor[0] = m.base_color_factor[1] = m.base_color_factor[2] = a;
    m.metallic_factor = 0.0f;
    m.roughness_factor = 1.0f;
  } else if (kind == 1) {  // rough metal, GGX alpha U(0.05, 0.5)
    double alpha = rng.uniform(0.05, 0.5);
    m.base_color_factor[0] = 0.95f;
    m.base_color_factor[1] = 0.78f;
    m.base_color_factor[2] = 0.45f;
    m.metallic_factor = 1.0f;
    m.roughness_factor = static_cast<float>(std::sqrt(alpha));
  } else {  // dispersive glass
    m.base_color_factor[0] = m.base_color_factor[1] = m.base_color_factor[2] = 1.0f;
    m.metallic_factor = 0.0f;
    m.roughness_factor = 0.0f;
    m.transmission_factor = 1.0f;
    m.thickness_factor = 1.0f;
    m.ior = 1.5f;
    m.dispersion = kGlassDispersion;
  }
  return m;
}

void icosphere(int subdiv, std::vector<D3>& pos, std::vector<uint32_t>& tri) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  pos = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto& p : pos) p = dnorm(p);
  tri = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  for (int s = 0; s < subdiv; ++s) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      auto key = std::minmax(a, b);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      pos.push_back(dnorm((pos[a] + pos[b]) * 0.5));
      uint32_t i = static_cast<uint32_t>(pos.size() - 1);
      mid[key] = i;
      return i;
    };
    std::vector<uint32_t> nt;
    for (size_t f = 0; f < tri.size(); f += 3) {
      uint32_t a = tri[f], b = tri[f + 1], c = tri[f + 2];
      uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      nt.insert(nt.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
    }
    tri.swap(nt);
  }
}

// Value noise on the integer lattice, trilinear with smoothstep weights.
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
  uint64_t h = seed ^ (static_cast<uint64_t>(x) * 0x9E3779B185EBCA87ull) ^
               (static_cast<uint64_t>(y) * 0xC2B2AE3D27D4EB4Full) ^ (static_cast<uint64_t>(z) * 0x165667B19E3779F9ull);
  h = Pcg32::splitmix(h);
  return (h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
double value_noise(D3 p, uint64_t seed) {
  double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
  int64_t ix = static_cast<int64_t>(fx), iy = static_cast<int64_t>(fy), iz = static_cast<int64_t>(fz);
  double tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
  auto sm = [](double t) { return t * t * (3 - 2 * t); };
  tx = sm(tx);
  ty = sm(ty);
  tz = sm(tz);
  double c[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
  auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
  double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
  double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
  return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
}

// Octahedral sphere with 8 n^2 triangles on the lattice |a| + |b| + |c| = n,
// radially displaced by 3-octave value noise of amplitude `amp`; smooth
// area-weighted normals.
void noise_sphere(int n, double amp, uint64_t seed, std::vector<D3>& pos, std::vector<D3>& nrm,
                  std::vector<uint32_t>& tri) {
  std::map<std::tuple<int, int, int>, uint32_t> index;
  auto vertex = [&](int a, int b, int c) {
    auto key = std::make_tuple(a, b, c);
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    D3 d = dnorm(D3{static_cast<double>(a), static_cast<double>(b), static_cast<double>(c)});
    double noise = 0, amp_sum = 0, f = 2.0, w = 1.0;
    for (int o = 0; o < 3; ++o) {
      noise += w * value_noise(d * f, seed + static_cast<uint64_t>(o));
      amp_sum += w;
      f *= 2.0;
      w *= 0.5;
    }
    pos.push_back(d * (1.0 + amp * noise / amp_sum));
    uint32_t i = static_cast<uint32_t>(pos.size() - 1);
    index[key] = i;
    return i;
  };
  for (int sx = -1; sx <= 1; sx += 2)
    for (int sy = -1; sy <= 1; sy += 2)
      for (int sz = -1; sz <= 1; sz += 2) {
        bool flip = (sx * sy * sz) < 0;
        for (int i = 0; i < n; ++i)
          for (int j = 0; j + i < n; ++j) {
            auto v = [&](int a, int b) { return vertex(sx * a, sy * b, sz * (n - a - b)); };
            uint32_t p0 = v(i, j), p1 = v(i + 1, j), p2 = v(i, j + 1);
            if (flip) std::swap(p1, p2);
            tri.insert(tri.end(), {p0, p1, p2});
            if (i + j < n - 1) {
              uint32_t q0 = v(i + 1, j), q1 = v(i + 1, j + 1), q2 = v(i, j + 1);
              if (flip) std::swap(q1, q2);
              tri.insert(tri.end(), {q0, q1, q2});
            }
          }
      }
  nrm.assign(pos.size(), D3{0, 0, 0});
  for (size_t f = 0; f < tri.size(); f += 3) {
    D3 c = dcross(pos[tri[f + 1]] - pos[tri[f]], pos[tri[f + 2]] - pos[tri[f]]);
    for (int k = 0; k < 3; ++k) nrm[tri[f + k]] = nrm[tri[f + k]] + c;
  }
  for (auto& v : nrm) v = dnorm(v);
}

void quad(double half, double z, bool down, std::vector<D3>& pos, std::vector<D3>& nrm, std::vector<uint32_t>& tri) {
  pos = {{-half, -half, z}, {half, -half, z}, {half, half, z}, {-half, half, z}};
  D3 n{0, 0, down ? -1.0 : 1.0};
  nrm.assign(4, n);
  tri = down ? std::vector<uint32_t>{0, 2, 1, 0, 3, 2} : std::vector<uint32_t>{0, 1, 2, 0, 2, 3};
}

void set_camera(Scene::Impl& s, D3 pos, D3 target, uint32_t w, uint32_t h) {
  D3 d = dnorm(target - pos);
  s.camera = cg_camera{};
  s.camera.position[0] = static_cast<float>(pos.x);
  s.camera.position[1] = static_cast<float>(pos.y);
  s.camera.position[2] = static_cast<float>(pos.z);
  s.camera.direction[0] = static_cast<float>(d.x);
  s.camera.direction[1] = static_cast<float>(d.y);
  s.camera.direction[2] = static_cast<float>(d.z);
  s.camera.viewport[0] = 0.5f * static_cast<float>(w) / static_cast<float>(h);
  s.camera.viewport[1] = 0.5f;
}
void set_conf(Scene::Impl& s, uint32_t w, uint32_t h, uint32_t spp, uint32_t depth) {
  s.conf = cg_conf{1920, 1080, w, h, spp, depth, 3, 1, 1, 1};
}

// Procedural material textures: an 8x8 checker of per-tile values with
// per-texel hash noise on top (so that neighbouring texels differ and the
// bilinear weights matter).  kind 0 base colour, 1 metallic-roughness (G
// roughness, B metallic mask), 2 transmission (R), 3 emission.
uint32_t mix32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xFF51AFD7ED558CCDull;
  x ^= x >> 33;
  x *= 0xC4CEB9FE1A85EC53ull;
  x ^= x >> 33;
  return static_cast<uint32_t>(x);
}
HostTexture make_texture(int kind, uint32_t w, uint32_t h, uint64_t seed, int32_t wrap_s, int32_t wrap_t) {
  HostTexture t;
  t.width = w;
  t.height = h;
  t.wrap_s = wrap_s;
  t.wrap_t = wrap_t;
  t.rgba8.resize(static_cast<size_t>(w) * h * 4);
  const uint64_t key = seed * 0x9E3779B97F4A7C15ull + static_cast<uint64_t>(kind) * 0x632BE59BD9B4E019ull;
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      const uint32_t tile = mix32(key ^ ((static_cast<uint64_t>(x * 8u / w) << 8) | (y * 8u / h)));
      const uint32_t n = mix32(key + (static_cast<uint64_t>(y) << 32 | x));
      auto ch = [&](int sh, int lo, int span) { return static_cast<uint8_t>(lo + ((tile >> sh) & 0xFFu) * span / 255 + (n >> sh) % 24u); };
      uint8_t* px = &t.rgba8[(static_cast<size_t>(y) * w + x) * 4];
      switch (kind) {
        case 0: px[0] = ch(0, 40, 190); px[1] = ch(8, 40, 190); px[2] = ch(16, 40, 190); px[3] = 255; break;
        case 1: px[0] = 255; px[1] = ch(0, 50, 180); px[2] = (tile & 0x100u) ? 255 : static_cast<uint8_t>(n % 16u); px[3] = 255; break;
        case 2: px[0] = ch(0, 120, 110); px[1] = px[2] = 0; px[3] = 255; break;
        default: px[0] = ch(0, 60, 170); px[1] = ch(8, 60, 170); px[2] = ch(16, 60, 170); px[3] = 255; break;
      }
    }
  return t;
}
cg_tex_info tex_ref(int index, uint32_t tex_coord) { return cg_tex_info{index, tex_coord, 1.0f, 1.0f}; }
// Texture the scene's materials with layers 0-3: base colour on diffuse and
// metal, metallic-roughness (through uv1) on metal, transmission on glass,
// emission (through uv1) on emitters.
void texture_materials(Scene::Impl& s) {
  for (size_t i = 1; i < s.materials.size(); ++i) {
    cg_material& m = s.materials[i];
    const bool emitter = m.emissive_factor[0] > 0.0f || m.emissive_factor[1] > 0.0f || m.emissive_factor[2] > 0.0f;
    if (emitter) {
      m.tex[CG_TEX_EMISSIVE] = tex_ref(3, 1);
    } else if (m.transmission_factor > 0.0f) {
      m.tex[CG_TEX_TRANSMISSION] = tex_ref(2, 0);
    } else {
      m.tex[CG_TEX_BASE_COLOR] = tex_ref(0, 0);
      if (m.metallic_factor > 0.0f) m.tex[CG_TEX_METALLIC_ROUGHNESS] = tex_ref(1, 1);
    }
  }
}

}  // namespace

cg_material DefaultMaterial() {  // src/scene.cpp:659-706
  cg_material m{};
  m.base_color_factor[0] = m.base_color_factor[1] = m.base_color_factor[2] = m.base_color_factor[3] = 1.0f;
  m.emissive_strength = 1.0f;
  m.metallic_factor = 1.0f;
  m.roughness_factor = 1.0f;
  m.alpha_cutoff = 0.5f;
  m.alpha_mode = 0;
  m.ior = 1.5f;
  m.attenuation_color[0] = m.attenuation_color[1] = m.attenuation_color[2] = 1.0f;
  m.attenuation_distance = INFINITY;
  m.iridescence_ior = 1.3f;
  m.iridescence_thickness_minimum = 100.0f;
  m.iridescence_thickness_maximum = 400.0f;
  m.specular_color_factor[0] = m.specular_color_factor[1] = m.specular_color_factor[2] = 1.0f;
  m.specular_factor = 1.0f;
  for (auto& t : m.tex) t = cg_tex_info{-1, 0, 1.0f, 1.0f};
  return m;
}

std::string GenerateScene(const std::string& name, uint64_t seed, const double* params, int nparams, Scene::Impl& s) {
  Pcg32 rng(seed);
  Builder B(s);
  s.env_rgba32f.clear();
  s.env_w = s.env_h = 0;
  auto param = [&](int i, double dflt) { return (params && i < nparams && params[i] > 0) ? params[i] : dflt; };
  std::vector<D3> pos, nrm;
  std::vector<uint32_t> tri;
  if (name == "c0") {
    cg_material ground = DefaultMaterial();
    ground.base_color_factor[0] = ground.base_color_factor[1] = ground.base_color_factor[2] = 0.5f;
    ground.metallic_factor = 0.0f;
    quad(4.0, 0.0, false, pos, nrm, tri);
    B.prim(B.mesh(pos, nrm, tri), B.material(ground), identity());
    std::vector<D3> ip;
    std::vector<uint32_t> it;
    icosphere(2, ip, it);
    for (int i = 0; i < 64; ++i) {
      D3 c{rng.uniform(-2.0, 2.0), rng.uniform(-2.0, 2.0), rng.uniform(0.0, 2.0)};
      double r = rng.uniform(0.1, 0.4);
      std::vector<D3> p2(ip.size());
      for (size_t k = 0; k < ip.size(); ++k) p2[k] = ip[k] * r;
      uint32_t mat = B.material(random_material(rng));
      B.prim(B.mesh(p2, ip, it), mat, translate(c));
    }
    cg_material light = DefaultMaterial();
    blackbody_srgb(6500.0, light.emissive_factor);
    light.emissive_strength = 10.0f;
    light.base_color_factor[0] = light.base_color_factor[1] = light.base_color_factor[2] = 0.0f;
    light.metallic_factor = 0.0f;
    quad(0.75, 3.5, true, pos, nrm, tri);
    B.prim(B.mesh(pos, nrm, tri), B.material(light), identity());
    set_camera(s, D3{0.0, -6.5, 2.6}, D3{0.0, 0.0, 0.7}, 320, 180);
    set_conf(s, 320, 180, 1, 2);
  } else if (name == "c1") {
    int n = static_cast<int>(128 * param(0, 1.0));
    noise_sphere(std::max(n, 1), 0.1, seed, pos, nrm, tri);
    auto sphere = B.mesh(pos, nrm, tri);
    for (int i = 0; i < 16; ++i) {
      M4 R = rotation(rng);
      D3 t{rng.uniform(-5.0, 5.0), rng.uniform(-5.0, 5.0), rng.uniform(-1.5, 1.5)};
      double sc = rng.uniform(0.5, 2.0);
      uint32_t mat = B.material(random_material(rng));
      B.prim(sphere, mat, mul(translate(t), mul(R, scale(sc))));
    }
    float white[3];
    blackbody_srgb(6500.0, white);
    for (int i = 0; i < 4096; ++i) {
      D3 c{rng.uniform(-9.0, 9.0), rng.uniform(-9.0, 9.0), rng.uniform(5.0, 8.0)};
      D3 a = dnorm(D3{rng.normal(), rng.normal(), rng.normal()});
      D3 b = dnorm(dcross(a, D3{rng.normal(), rng.normal(), rng.normal()}));
      double e = rng.uniform(0.1, 0.3);
      std::vector<D3> tp{c, c + a * e, c + b * e};
      D3 fn = dnorm(dcross(tp[1] - tp[0], tp[2] - tp[0]));
      std::vector<D3> tn(3, fn);
      cg_material m = DefaultMaterial();
      for (int k = 0; k < 3; ++k) m.emissive_factor[k] = white[k];
      m.emissive_strength = static_cast<float>(std::pow(10.0, rng.uniform(0.0, 2.0)));
      m.base_color_factor[0] = m.base_color_factor[1] = m.base_color_factor[2] = 0.0f;
      m.metallic_factor = 0.0f;
      B.prim(B.mesh(tp, tn, {0, 1, 2}), B.material(m), identity());
    }
    if (param(1, 0.0) > 0.0) {
      const uint32_t ts = static_cast<uint32_t>(param(2, 2048.0));
      for (int k = 0; k < 4; ++k) s.textures.push_back(make_texture(k, ts, ts, seed, CG_WRAP_REPEAT, CG_WRAP_REPEAT));
      texture_materials(s);
    }
    set_camera(s, D3{0.0, -15.0, 5.0}, D3{0.0, 0.0, 0.0}, 1280, 720);
    set_conf(s, 1280, 720, 16, 8);
  } else if (name == "c2") {
    const uint64_t count = static_cast<uint64_t>(param(0, 10000000.0));
    std::vector<D3> centers(1024);
    for (auto& c : centers) c = D3{rng.uniform(), rng.uniform(), rng.uniform()};
    pos.reserve(count * 3);
    nrm.reserve(count * 3);
    tri.reserve(count * 3);
    for (uint64_t i = 0; i < count; ++i) {
      const D3& cc = centers[rng.next() % 1024u];
      D3 p0{cc.x + 0.02 * rng.normal(), cc.y + 0.02 * rng.normal(), cc.z + 0.02 * rng.normal()};
      p0 = D3{std::min(1.0, std::max(0.0, p0.x)), std::min(1.0, std::max(0.0, p0.y)), std::min(1.0, std::max(0.0, p0.z))};
      D3 u = dnorm(D3{rng.normal(), rng.normal(), rng.normal()});
      D3 w = dnorm(dcross(u, D3{rng.normal(), rng.normal(), rng.normal()}));
      double l1 = rng.uniform(0.001, 0.01), l2 = rng.uniform(0.001, 0.01);
      // second edge at 0.9 rad from the first, in the plane of u and w
      D3 p1 = p0 + u * l1, p2 = p0 + (u * std::cos(0.9) + w * std::sin(0.9)) * l2;
      D3 fn = dnorm(dcross(p1 - p0, p2 - p0));
      uint32_t base = static_cast<uint32_t>(pos.size());
      pos.insert(pos.end(), {p0, p1, p2});
      nrm.insert(nrm.end(), {fn, fn, fn});
      tri.insert(tri.end(), {base, base + 1, base + 2});
    }
    cg_material m = DefaultMaterial();
    m.metallic_factor = 0.0f;
    m.base_color_factor[0] = m.base_color_factor[1] = m.base_color_factor[2] = 0.6f;
    B.prim(B.mesh(pos, nrm, tri), B.material(m), identity());
    set_camera(s, D3{0.5, -1.2, 0.5}, D3{0.5, 0.5, 0.5}, 4096, 2048);
    set_conf(s, 4096, 2048, 1, 1);
  } else if (name == "quad") {
    cg_material m = DefaultMaterial();
    m.base_color_factor[0] = 0.8f;
    m.base_color_factor[1] = 0.6f;
    m.base_color_factor[2] = 0.4f;
    m.metallic_factor = 0.0f;
    quad(1.0, 0.0, false, pos, nrm, tri);
    B.prim(B.mesh(pos, nrm, tri), B.material(m), identity());
    set_camera(s, D3{0.0, -2.0, 2.0}, D3{0.0, 0.0, 0.0}, 64, 64);
    set_conf(s, 64, 64, 4, 2);
  } else if (name == "spheres") {
    int count = static_cast<int>(param(0, 8.0));
    cg_material ground = DefaultMaterial();
    ground.base_color_factor[0] = ground.base_color_factor[1] = ground.base_color_factor[2] = 0.5f;
    ground.metallic_factor = 0.0f;
    quad(3.0, 0.0, false, pos, nrm, tri);
    B.prim(B.mesh(pos, nrm, tri), B.material(ground), identity());
    std::vector<D3> ip;
    std::vector<uint32_t> it;
    icosphere(2, ip, it);
    for (int i = 0; i < count; ++i) {
      D3 c{rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(0.2, 1.2)};
      double r = rng.uniform(0.15, 0.45);
      uint32_t mat = B.material(random_material(rng));
      std::vector<D3> p2(ip.size());
      for (size_t k = 0; k < ip.size(); ++k) p2[k] = ip[k] * r;
      B.prim(B.mesh(p2, ip, it), mat, translate(c));
    }
    cg_material light = DefaultMaterial();
    blackbody_srgb(6500.0, light.emissive_factor);
    light.emissive_strength = 8.0f;
    light.base_color_factor[0] = light.base_color_factor[1] = light.base_color_factor[2] = 0.0f;
    light.metallic_factor = 0.0f;
    quad(0.5, 2.5, true, pos, nrm, tri);
    B.prim(B.mesh(pos, nrm, tri), B.material(light), identity());
    if (param(1, 0.0) > 0.0) {  // layers of different sizes and wrap modes
      const uint32_t ts = static_cast<uint32_t>(param(2, 64.0));
      s.textures.push_back(make_texture(0, ts, ts, seed, CG_WRAP_REPEAT, CG_WRAP_REPEAT));
      s.textures.push_back(make_texture(1, std::max(1u, ts / 2), ts + 3, seed, CG_WRAP_CLAMP_TO_EDGE, CG_WRAP_REPEAT));
      s.textures.push_back(make_texture(2, 7, 5, seed, CG_WRAP_MIRRORED_REPEAT, CG_WRAP_MIRRORED_REPEAT));
      s.textures.push_back(make_texture(3, ts, std::max(1u, ts / 4), seed, CG_WRAP_MIRRORED_REPEAT, CG_WRAP_CLAMP_TO_EDGE));
      texture_materials(s);
      s.materials[1].tex[CG_TEX_BASE_COLOR] = tex_ref(0, 1);  // the ground, through uv1
    }
    set_camera(s, D3{0.0, -4.5, 2.0}, D3{0.0, 0.0, 0.5}, 96, 64);
    set_conf(s, 96, 64, 4, 4);
  } else {
    return "unknown procedural scene '" + name + "'";
  }
  uint64_t tag_seed = seed;
  for (int i = 0; params && i < nparams; ++i) {
    uint64_t bits;
    std::memcpy(&bits, &params[i], sizeof bits);
    tag_seed = tag_seed * 0x100000001B3ull ^ bits;
  }
  s.tag = Pcg32::splitmix(tag_seed) ^ std::hash<std::string>{}(name);
  return "";
}

void GenerateRays(const Scene::Impl& s, uint64_t seed, uint32_t n, cg_ray* out) {
  Pcg32 rng(seed ^ 0x5EEDull);
  const uint32_t half = n / 2;
  // Coherent half: pinhole rays on a grid (aspect 2:1), in pixel order.
  uint32_t gw = 1, gh = 1;
  while (static_cast<uint64_t>(gw) * gh < half) {
    if (gw <= gh)
      gw *= 2;
    else
      gh *= 2;
  }
  if (gw < gh) std::swap(gw, gh);
  const float* cp = s.camera.position;
  const float* cd = s.camera.direction;
  D3 d{cd[0], cd[1], cd[2]};
  D3 dx = dnorm(dcross(d, D3{0, 0, 1})) * 0.5;
  D3 dy = dnorm(dcross(d, dx)) * 0.5;
  for (uint32_t i = 0; i < half; ++i) {
    uint32_t px = i % gw, py = i / gw;
    double cx = (px + 0.5) / gw * 2.0 - 1.0, cy = (py + 0.5) / gh * 2.0 - 1.0;
    D3 r = dnorm(d + dx * cx + dy * cy);
    cg_ray& o = out[i];
    o = cg_ray{};
    o.org[0] = cp[0];
    o.org[1] = cp[1];
    o.org[2] = cp[2];
    o.dir[0] = static_cast<float>(r.x);
    o.dir[1] = static_cast<float>(r.y);
    o.dir[2] = static_cast<float>(r.z);
  }
  // Incoherent half: origins on random triangles (offset 1e-4 along the
  // normal), cosine-weighted directions about the side they leave from.
  const uint32_t tris = static_cast<uint32_t>(s.indices.size() / 3);
  for (uint32_t i = half; i < n; ++i) {
    uint32_t t = rng.next() % tris;
    const float* a = s.vertices[s.indices[3 * t]].pos;
    const float* b = s.vertices[s.indices[3 * t + 1]].pos;
    const float* c = s.vertices[s.indices[3 * t + 2]].pos;
    double u = rng.uniform(), v = rng.uniform();
    if (u + v > 1) {
      u = 1 - u;
      v = 1 - v;
    }
    D3 A{a[0], a[1], a[2]}, Bv{b[0], b[1], b[2]}, C{c[0], c[1], c[2]};
    D3 p = A + (Bv - A) * u + (C - A) * v;
    D3 nn = dnorm(dcross(Bv - A, C - A));
    if (rng.next() & 1u) nn = nn * -1.0;
    D3 h = std::fabs(nn.x) > 0.9 ? D3{0, 1, 0} : D3{1, 0, 0};
    D3 tt = dnorm(dcross(h, nn)), bb = dcross(nn, tt);
    double r = std::sqrt(rng.uniform()), phi = 6.283185307179586 * rng.uniform();
    D3 dir = dnorm(tt * (r * std::cos(phi)) + bb * (r * std::sin(phi)) + nn * std::sqrt(std::max(0.0, 1 - r * r)));
    p = p + nn * 1e-4;
    cg_ray& o = out[i];
    o = cg_ray{};
    o.org[0] = static_cast<float>(p.x);
    o.org[1] = static_cast<float>(p.y);
    o.org[2] = static_cast<float>(p.z);
    o.dir[0] = static_cast<float>(dir.x);
    o.dir[1] = static_cast<float>(dir.y);
    o.dir[2] = static_cast<float>(dir.z);
  }
}

}  // namespace crystal::graphics
