// scene_prep.cpp — SceneData build of the crystal::graphics host layer:
// BLAS/TLAS median-split trees, emissive triangle buffer, alias table and
// environment texture, in the reference's buffer formats
// (src/pathtracing/{blas,tlas,emissive_buffer,alias_table,scene_data}.h).
//
// The BLAS build is a task tree: every interior node's descendants occupy a
// contiguous range whose size follows from the triangle count alone, so
// subtrees are built on separate threads and still receive exactly the node
// numbers of the reference's serial preorder build (blas.h:171-252).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>
#include <utility>

#include "scene_impl.h"

namespace crystal::graphics {
namespace {

struct V3 {
  float x, y, z;
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 ld(const float* p) { return V3{p[0], p[1], p[2]}; }
inline V3 vmin(V3 a, V3 b) { return V3{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return V3{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }

// Emissive texture factor at a triangle's centroid uv
// (src/pathtracing/emissive_buffer.h:135-212): nearest texel of the decoded
// image under the texture's own wrap modes.  A NaN coordinate reads texel 0.
float wrap_coord(float c, int32_t wrap) {  // emissive_buffer.h:191-204
  float w;
  if (wrap == CG_WRAP_CLAMP_TO_EDGE) {
    w = c;
  } else if (wrap == CG_WRAP_MIRRORED_REPEAT) {
    const float fl = std::floor(c);
    w = c - fl;
    if (std::fmod(fl, 2.0f) != 0.0f) w = 1.0f - w;
  } else {
    w = c - std::floor(c);
  }
  if (!(w >= 0.0f)) w = 0.0f;  // NaN
  return std::min(w, 0.999999f);
}
V3 emissive_texture_factor(const Scene::Impl& sc, const cg_material& m, const cg_vertex& a, const cg_vertex& b,
                           const cg_vertex& c) {
  const cg_tex_info& ti = m.tex[CG_TEX_EMISSIVE];
  if (ti.index < 0 || static_cast<size_t>(ti.index) >= sc.textures.size()) return V3{1.0f, 1.0f, 1.0f};
  const HostTexture& t = sc.textures[static_cast<size_t>(ti.index)];
  const float* ua = ti.tex_coord == 1 ? a.uv1 : a.uv0;
  const float* ub = ti.tex_coord == 1 ? b.uv1 : b.uv0;
  const float* uc = ti.tex_coord == 1 ? c.uv1 : c.uv0;
  const float u = ((ua[0] + ub[0]) + uc[0]) / 3.0f, v = ((ua[1] + ub[1]) + uc[1]) / 3.0f;
  const uint32_t x = std::min(t.width - 1, static_cast<uint32_t>(wrap_coord(u, t.wrap_s) * static_cast<float>(t.width)));
  const uint32_t y = std::min(t.height - 1, static_cast<uint32_t>(wrap_coord(v, t.wrap_t) * static_cast<float>(t.height)));
  const uint8_t* px = &t.rgba8[(static_cast<size_t>(y) * t.width + x) * 4];
  return V3{static_cast<float>(px[0]) / 255.0f, static_cast<float>(px[1]) / 255.0f, static_cast<float>(px[2]) / 255.0f};
}
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline void st(float* p, V3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}
// glm's mat4 * vec4(p, 1) for a column-major matrix, pairwise:
// (m0 * x + m1 * y) + (m2 * z + m3 * 1).
inline V3 glm_point(const float* m, V3 p) {
  V3 a = add(mul(ld(m), p.x), mul(ld(m + 4), p.y));
  V3 b = add(mul(ld(m + 8), p.z), mul(ld(m + 12), 1.0f));
  return add(a, b);
}
inline int longest_axis(V3 range) {  // blas.h:216-219, tlas.h:120-123
  int axis = 0;
  if (range.y > range.x) axis = 1;
  if (range.z > range[axis]) axis = 2;
  return axis;
}

// Interior nodes of a median-split tree over n triangles with leaves of at
// most 2: I(n) = 0 for n <= 2, else 1 + I(floor(n/2)) + I(ceil(n/2)).  Each
// level holds only sizes a and a + 1, so this walks levels, not nodes.
uint64_t interior_nodes(uint64_t n) {
  uint64_t total = 0, a = n, ca = 1, cb = 0;
  for (;;) {
    uint64_t here = (a > 2 ? ca : 0) + (a + 1 > 2 ? cb : 0);
    if (here == 0) break;
    total += here;
    uint64_t f = a / 2, na = 0, nb = 0;
    auto put = [&](uint64_t size, uint64_t count) {
      if (size == f)
        na += count;
      else
        nb += count;
    };
    if (a > 2) {
      put(a / 2, ca);
      put(a - a / 2, ca);
    }
    if (a + 1 > 2) {
      put((a + 1) / 2, cb);
      put((a + 1) - (a + 1) / 2, cb);
    }
    a = f;
    ca = na;
    cb = nb;
  }
  return total;
}

struct TriInfo {
  V3 center;
  uint32_t index;
};

struct BlasBuilder {
  const cg_primitive& prim;
  const std::vector<cg_vertex>& verts;
  const std::vector<uint32_t>& idx;
  std::vector<TriInfo>& info;
  cg_blas_node* nodes;
  uint32_t base;        // global index of this BLAS's root
  uint32_t tri_offset;  // global slot of this BLAS's first triangle

  V3 pos(uint32_t t, int k) const { return ld(verts[idx[prim.index_offset + t * 3 + k]].pos); }

  // node: local index (0 = root); next: first local index of its descendants.
  // Returns the node's unpadded box.  Boxes are unions of exact minima and
  // maxima, so building them bottom-up gives the same floats as scanning every
  // triangle of the node (blas.h:187-201) at O(n) instead of O(n log n).
  void build(uint32_t node, uint32_t l, uint32_t r, uint32_t next, int par_depth, V3& lb, V3& ub) {
    const uint32_t count = r - l;
    const V3 pad{0.001f, 0.001f, 0.001f};
    cg_blas_node& n = nodes[node];
    if (count <= 2) {
      lb = V3{FLT_MAX, FLT_MAX, FLT_MAX};
      ub = V3{-FLT_MAX, -FLT_MAX, -FLT_MAX};
      for (uint32_t i = l; i < r; ++i) {
        V3 v0 = pos(info[i].index, 0), v1 = pos(info[i].index, 1), v2 = pos(info[i].index, 2);
        lb = vmin(lb, vmin(vmin(v0, v1), v2));
        ub = vmax(ub, vmax(vmax(v0, v1), v2));
      }
      st(n.lb, sub(lb, pad));
      st(n.ub, add(ub, pad));
      n.child = tri_offset + l;
      n.triangle_count = count;
      return;
    }
    V3 clb{FLT_MAX, FLT_MAX, FLT_MAX}, cub{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = l; i < r; ++i) {
      clb = vmin(clb, info[i].center);
      cub = vmax(cub, info[i].center);
    }
    const int axis = longest_axis(sub(cub, clb));
    const uint32_t mid = l + (r - l) / 2;
    std::nth_element(info.begin() + l, info.begin() + mid, info.begin() + r,
                     [axis](const TriInfo& a, const TriInfo& b) { return a.center[axis] < b.center[axis]; });
    const uint32_t c = next;
    n.child = base + c;
    n.triangle_count = 0;
    const uint32_t left_next = next + 2;
    const uint32_t right_next = next + 2 + static_cast<uint32_t>(2 * interior_nodes(mid - l));
    V3 llb, lub, rlb, rub;
    if (par_depth > 0 && count > 65536) {
      std::thread t([&] { build(c, l, mid, left_next, par_depth - 1, llb, lub); });
      build(c + 1, mid, r, right_next, par_depth - 1, rlb, rub);
      t.join();
    } else {
      build(c, l, mid, left_next, 0, llb, lub);
      build(c + 1, mid, r, right_next, 0, rlb, rub);
    }
    lb = vmin(llb, rlb);
    ub = vmax(lub, rub);
    st(nodes[node].lb, sub(lb, pad));
    st(nodes[node].ub, add(ub, pad));
  }
};

}  // namespace

std::vector<cg_alias_entry> BuildAliasTable(const std::vector<double>& weights) {
  const uint32_t n = static_cast<uint32_t>(weights.size());
  std::vector<cg_alias_entry> out(n, cg_alias_entry{0u, 0.0f});
  if (n == 0) return out;
  std::vector<double> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = std::max(0.0, weights[i]);
  const double total = std::accumulate(p.begin(), p.end(), 0.0);
  if (total <= 0.0) return out;
  std::vector<uint32_t> large, small;
  large.reserve(n);
  small.reserve(n);
  for (uint32_t i = 0; i < n; ++i) {
    p[i] = p[i] / total * n;
    (p[i] >= 1.0 ? large : small).push_back(i);
  }
  while (!large.empty() && !small.empty()) {
    const uint32_t i = small.back(), j = large.back();
    small.pop_back();
    large.pop_back();
    out[i].prob = static_cast<float>(p[i]);
    out[i].alias = j;
    p[j] -= 1.0 - out[i].prob;  // the stored (float) probability is what j gives away
    (p[j] < 1.0 ? small : large).push_back(j);
  }
  for (auto* rest : {&large, &small})
    while (!rest->empty()) {
      const uint32_t i = rest->back();
      rest->pop_back();
      out[i].prob = 1.0f;
      out[i].alias = i;
    }
  return out;
}

cg_scene_data SceneData::View() const {
  cg_scene_data d{};
  d.tlas_nodes = tlas.data();
  d.tlas_node_count = static_cast<uint32_t>(tlas.size());
  d.instances = instances.data();
  d.instance_count = static_cast<uint32_t>(instances.size());
  d.blas_nodes = blas.data();
  d.blas_node_count = static_cast<uint32_t>(blas.size());
  d.indices = indices.data();
  d.index_count = static_cast<uint32_t>(indices.size());
  d.tri_ids = tri_ids.data();
  d.vertices = scene->vertices.data();
  d.vertex_count = static_cast<uint32_t>(scene->vertices.size());
  d.materials = scene->materials.data();
  d.material_count = static_cast<uint32_t>(scene->materials.size());
  d.emissive = emissive.data();
  d.emissive_count = static_cast<uint32_t>(emissive.size());
  d.alias = alias.data();
  d.alias_count = static_cast<uint32_t>(alias.size());
  d.env_rgba32f = env.data();
  d.env_width = env_w;
  d.env_height = env_h;
  d.tex_rgba8 = tex.empty() ? nullptr : tex.data();
  d.tex_width = tex_w;
  d.tex_height = tex_h;
  d.tex_layers = tex_layers;
  return d;
}

std::vector<uint8_t> ResizeRgba8Nearest(const HostTexture& t, uint32_t w, uint32_t h) {
  if (t.width == w && t.height == h) return t.rgba8;
  std::vector<uint8_t> out(static_cast<size_t>(w) * h * 4);
  for (uint32_t y = 0; y < h; ++y) {
    const uint32_t sy = std::min(t.height - 1, static_cast<uint32_t>(static_cast<uint64_t>(y) * t.height / h));
    for (uint32_t x = 0; x < w; ++x) {
      const uint32_t sx = std::min(t.width - 1, static_cast<uint32_t>(static_cast<uint64_t>(x) * t.width / w));
      std::memcpy(&out[(static_cast<size_t>(y) * w + x) * 4], &t.rgba8[(static_cast<size_t>(sy) * t.width + sx) * 4], 4);
    }
  }
  return out;
}

SceneData BuildSceneData(const Scene::Impl& sc, bool share_blas, int threads) {
  SceneData sd;
  sd.scene = &sc;
  const auto& verts = sc.vertices;
  const auto& idx = sc.indices;
  const uint32_t np = static_cast<uint32_t>(sc.prims.size());
  int par_depth = 0;
  while ((1 << par_depth) < threads && par_depth < 6) ++par_depth;

  // BLAS per primitive object (blas.h:94-169).
  std::vector<uint32_t> roots(np);
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> built;  // (index_offset, index_count) -> root
  for (uint32_t p = 0; p < np; ++p) {
    const cg_primitive& pr = sc.prims[p];
    const auto key = std::make_pair(pr.index_offset, pr.index_count);
    if (share_blas) {
      auto it = built.find(key);
      if (it != built.end()) {
        roots[p] = it->second;
        continue;
      }
    }
    const uint32_t tcount = pr.index_count / 3;
    std::vector<TriInfo> info(tcount);
    for (uint32_t t = 0; t < tcount; ++t) {
      const uint32_t o = pr.index_offset + t * 3;
      V3 v0 = ld(verts[idx[o]].pos), v1 = ld(verts[idx[o + 1]].pos), v2 = ld(verts[idx[o + 2]].pos);
      V3 c = add(add(v0, v1), v2);
      info[t] = TriInfo{V3{c.x / 3.0f, c.y / 3.0f, c.z / 3.0f}, t};
    }
    const uint32_t base = static_cast<uint32_t>(sd.blas.size());
    const uint32_t offset = static_cast<uint32_t>(sd.indices.size());
    const uint64_t nodes = 1 + 2 * interior_nodes(tcount);
    cg_blas_node init{};
    init.lb[0] = init.lb[1] = init.lb[2] = FLT_MAX;
    init.ub[0] = init.ub[1] = init.ub[2] = FLT_MIN;
    sd.blas.resize(base + nodes, init);
    BlasBuilder b{pr, verts, idx, info, sd.blas.data() + base, base, offset};
    V3 rlb, rub;
    b.build(0, 0, tcount, 1, par_depth, rlb, rub);
    for (const TriInfo& t : info) {
      const uint32_t o = pr.index_offset + t.index * 3;
      sd.indices.push_back(cg_index{idx[o], idx[o + 1], idx[o + 2], 0u});
      sd.tri_ids.push_back(t.index);
    }
    roots[p] = base;
    built[key] = base;
  }

  // TLAS (tlas.h:34-140).
  struct Bounded {
    V3 lb, ub, center;
    uint32_t prim;
  };
  std::vector<Bounded> bp(np);
  sd.instances.resize(np);
  for (uint32_t p = 0; p < np; ++p) {
    const cg_primitive& pr = sc.prims[p];
    cg_instance& in = sd.instances[p];
    std::memcpy(in.inv_trans, pr.inv_world, sizeof(in.inv_trans));
    in.blas_root_idx = roots[p];
    in.material_idx = pr.material_idx;
    in._pad[0] = in._pad[1] = 0;
    // AABB{primitive, vertices} starts from ub = FLT_MIN (aabb.h:17-33).
    V3 alb{FLT_MAX, FLT_MAX, FLT_MAX}, aub{FLT_MIN, FLT_MIN, FLT_MIN};
    for (uint32_t i = pr.vertex_offset; i < pr.vertex_offset + pr.vertex_count; ++i) {
      alb = vmin(alb, ld(verts[i].pos));
      aub = vmax(aub, ld(verts[i].pos));
    }
    const float* m = pr.world;  // trans_aabb (tlas.h:38-55)
    float lb[3] = {m[12], m[13], m[14]}, ub[3] = {m[12], m[13], m[14]};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const float a = m[4 * j + i] * alb[j], b = m[4 * j + i] * aub[j];
        lb[i] += a < b ? a : b;
        ub[i] += a < b ? b : a;
      }
    bp[p].lb = ld(lb);
    bp[p].ub = ld(ub);
    bp[p].center = mul(add(bp[p].lb, bp[p].ub), 0.5f);
    bp[p].prim = p;
  }
  if (np > 0) {
    sd.tlas.reserve(2 * static_cast<size_t>(np));
    cg_tlas_node root{};
    sd.tlas.push_back(root);
    struct Item {
      uint32_t node, l, r;
    };
    // Preorder with an explicit stack (left subtree completes before right).
    std::vector<Item> stack{{0u, 0u, np}};
    while (!stack.empty()) {
      Item it = stack.back();
      stack.pop_back();
      V3 lb{FLT_MAX, FLT_MAX, FLT_MAX}, ub{-FLT_MAX, -FLT_MAX, -FLT_MAX};
      for (uint32_t i = it.l; i < it.r; ++i) {
        lb = vmin(lb, bp[i].lb);
        ub = vmax(ub, bp[i].ub);
      }
      V3 pad{0.001f, 0.001f, 0.001f};
      st(sd.tlas[it.node].lb, sub(lb, pad));
      st(sd.tlas[it.node].ub, add(ub, pad));
      if (it.r - it.l <= 1) {
        sd.tlas[it.node].instance_idx = bp[it.l].prim;
        sd.tlas[it.node].child = 0;
        continue;
      }
      V3 clb{FLT_MAX, FLT_MAX, FLT_MAX}, cub{FLT_MIN, FLT_MIN, FLT_MIN};  // AABB{} default
      for (uint32_t i = it.l; i < it.r; ++i) {
        clb = vmin(clb, bp[i].center);
        cub = vmax(cub, bp[i].center);
      }
      const int axis = longest_axis(sub(cub, clb));
      const uint32_t mid = it.l + (it.r - it.l) / 2;
      std::nth_element(bp.begin() + it.l, bp.begin() + mid, bp.begin() + it.r,
                       [axis](const Bounded& x, const Bounded& y) { return x.center[axis] < y.center[axis]; });
      const uint32_t c = static_cast<uint32_t>(sd.tlas.size());
      sd.tlas[it.node].child = c;
      sd.tlas[it.node].instance_idx = 0;
      sd.tlas.push_back(root);
      sd.tlas.push_back(root);
      stack.push_back({c + 1, mid, it.r});
      stack.push_back({c, it.l, mid});
    }
  }

  // Emissive buffer (emissive_buffer.h:44-100).
  float total = 0.0f;
  auto lum = [](V3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; };
  for (uint32_t p = 0; p < np; ++p) {
    const cg_primitive& pr = sc.prims[p];
    if (pr.material_idx >= sc.materials.size()) continue;
    const cg_material& m = sc.materials[pr.material_idx];
    const V3 base{std::fmax(m.emissive_factor[0] * m.emissive_strength, 0.0f),
                  std::fmax(m.emissive_factor[1] * m.emissive_strength, 0.0f),
                  std::fmax(m.emissive_factor[2] * m.emissive_strength, 0.0f)};
    if (lum(base) <= 0.0f) continue;
    for (uint32_t i = 0; i + 2 < pr.index_count; i += 3) {
      const uint32_t o = pr.index_offset + i;
      const V3 p0 = glm_point(pr.world, ld(verts[idx[o]].pos));
      const V3 p1 = glm_point(pr.world, ld(verts[idx[o + 1]].pos));
      const V3 p2 = glm_point(pr.world, ld(verts[idx[o + 2]].pos));
      const V3 cr = cross(sub(p1, p0), sub(p2, p0));
      const float area = std::sqrt(dot(cr, cr)) * 0.5f;
      if (area <= 0.0f) continue;
      const V3 f = emissive_texture_factor(sc, m, verts[idx[o]], verts[idx[o + 1]], verts[idx[o + 2]]);
      const V3 em{base.x * f.x, base.y * f.y, base.z * f.z};
      const float w = lum(em);
      if (w <= 0.0f) continue;
      cg_emissive_prim e{};
      st(e.v[0], p0);
      st(e.v[1], p1);
      st(e.v[2], p2);
      st(e.emission, em);
      e.area = area;
      e.weight = area * w;
      sd.emissive.push_back(e);
      total += e.weight;
    }
  }
  if (total > 0.0f)
    for (auto& e : sd.emissive) e.weight /= total;
  std::vector<double> weights;
  weights.reserve(sd.emissive.size());
  for (const auto& e : sd.emissive) weights.push_back(static_cast<double>(e.weight));
  sd.alias = BuildAliasTable(weights);

  // Material texture array (src/wgpu/resource.cpp:471-546).
  if (!sc.textures.empty()) {
    sd.tex_w = sd.tex_h = 1;
    for (const auto& t : sc.textures) {
      sd.tex_w = std::max(sd.tex_w, t.width);
      sd.tex_h = std::max(sd.tex_h, t.height);
    }
    sd.tex_layers = static_cast<uint32_t>(sc.textures.size());
    const size_t layer = static_cast<size_t>(sd.tex_w) * sd.tex_h * 4;
    sd.tex.resize(layer * sd.tex_layers);
    for (uint32_t l = 0; l < sd.tex_layers; ++l) {
      std::vector<uint8_t> r = ResizeRgba8Nearest(sc.textures[l], sd.tex_w, sd.tex_h);
      std::memcpy(&sd.tex[layer * l], r.data(), layer);
    }
  }

  // Environment (scene_data.h:57-98).
  if (!sc.env_rgba32f.empty() && sc.env_w && sc.env_h) {
    sd.env = sc.env_rgba32f;
    sd.env_w = sc.env_w;
    sd.env_h = sc.env_h;
  } else {
    sd.env_w = 512;
    sd.env_h = 256;
    sd.env.resize(512 * 256 * 4);
    for (uint32_t y = 0; y < 256; ++y) {
      const float t = static_cast<float>(y) / static_cast<float>(256 - 1);
      // glm::mix(a, b, t) = a * (1 - t) + b * t
      const float c[3] = {0.7f * (1.0f - t) + 0.05f * t, 0.8f * (1.0f - t) + 0.05f * t,
                          1.0f * (1.0f - t) + 0.06f * t};
      for (uint32_t x = 0; x < 512; ++x) {
        float* px = &sd.env[(static_cast<size_t>(y) * 512 + x) * 4];
        px[0] = c[0];
        px[1] = c[1];
        px[2] = c[2];
        px[3] = 1.0f;
      }
    }
  }
  return sd;
}

}  // namespace crystal::graphics
