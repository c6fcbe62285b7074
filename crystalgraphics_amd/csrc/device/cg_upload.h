// cg_upload.h — host-side construction of the device layout (cg_layout.h)
// from the reference's flattened SceneData (cg_scene.h), and the wavelength
// sampling table.
#ifndef CG_UPLOAD_H_
#define CG_UPLOAD_H_

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/cg_device.h"
#include "cg_core.h"
#include "cg_layout.h"
#include "cg_lbvh.h"
#include "cg_quant.h"

namespace cg {

inline int set_error(std::string& err, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  err = buf;
  return code;
}

// fn(begin, end) over [0, n) in up to 16 chunks on threads (one call below
// 2^16 items).  For loops whose iterations are independent.
template <typename F>
inline void par_ranges(size_t n, F&& fn) {
  const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < (size_t{1} << 16) || hw == 1) {
    fn(size_t{0}, n);
    return;
  }
  const size_t chunk = (n + hw - 1) / hw;
  std::vector<std::thread> pool;
  for (size_t t = 0; t < hw; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&fn, b, e] { fn(b, e); });
  }
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Wavelength table: the reference's 16-step bisection over the CMF-sum CDF
// (physics/wavelength.wesl:20-45, physics/cie.wesl:1-140) evaluated on the
// host once, in binary32 with the same operation order, so the device walk
// `i = 2i + (cdf[i] < target)` reproduces its outcome exactly.
namespace cie {

constexpr float kSqrtHalfPi = 1.2533141373155001f;
constexpr float kInvSqrt2 = 0.7071067811865475f;
inline float erf_as(float x) {  // cie.wesl:132-140
  float sign = x >= 0.0f ? 1.0f : -1.0f;
  float ax = std::fabs(x);
  float t = 1.0f / (1.0f + 0.3275911f * ax);
  float poly =
      (((((1.061405429f * t - 1.453152027f) * t) + 1.421413741f) * t - 0.284496736f) * t + 0.254829592f) * t;
  return sign * (1.0f - poly * std::exp(-ax * ax));
}
inline float gauss(float l, float c, float lw, float rw) {  // cie.wesl:95-104
  float w = l < c ? lw : rw;
  float s = (l - c) / w;
  return std::exp(-0.5f * s * s);
}
inline float anti(float l, float c, float lw, float rw) {  // cie.wesl:117-130
  if (l < c) return kSqrtHalfPi * lw * (1.0f + erf_as((l - c) * kInvSqrt2 / lw));
  return kSqrtHalfPi * lw + kSqrtHalfPi * rw * erf_as((l - c) * kInvSqrt2 / rw);
}
inline float integ(float lo, float hi, float c, float lw, float rw) { return anti(hi, c, lw, rw) - anti(lo, c, lw, rw); }
struct L4 {
  float a, c, lw, rw;
};
const L4 kX[3] = {{0.362f, 442.0f, 1.0f / 0.0624f, 1.0f / 0.0374f},
                  {1.056f, 599.8f, 1.0f / 0.0264f, 1.0f / 0.0323f},
                  {0.065f, 501.1f, 1.0f / 0.0490f, 1.0f / 0.0382f}};
const L4 kY[2] = {{0.821f, 568.8f, 1.0f / 0.0213f, 1.0f / 0.0247f},
                  {0.286f, 530.9f, 1.0f / 0.0613f, 1.0f / 0.0322f}};
const L4 kZ[2] = {{1.217f, 437.0f, 1.0f / 0.0845f, 1.0f / 0.0278f},
                  {0.681f, 459.0f, 1.0f / 0.0385f, 1.0f / 0.0725f}};
inline void match(float l, float& x, float& y, float& z) {  // cie.wesl:4-30
  x = kX[0].a * gauss(l, kX[0].c, kX[0].lw, kX[0].rw) + kX[1].a * gauss(l, kX[1].c, kX[1].lw, kX[1].rw) -
      kX[2].a * gauss(l, kX[2].c, kX[2].lw, kX[2].rw);
  y = kY[0].a * gauss(l, kY[0].c, kY[0].lw, kY[0].rw) + kY[1].a * gauss(l, kY[1].c, kY[1].lw, kY[1].rw);
  z = kZ[0].a * gauss(l, kZ[0].c, kZ[0].lw, kZ[0].rw) + kZ[1].a * gauss(l, kZ[1].c, kZ[1].lw, kZ[1].rw);
}
inline float sum_integral(float lo, float hi) {  // cie.wesl:37-92
  float x = kX[0].a * integ(lo, hi, kX[0].c, kX[0].lw, kX[0].rw) +
            kX[1].a * integ(lo, hi, kX[1].c, kX[1].lw, kX[1].rw) -
            kX[2].a * integ(lo, hi, kX[2].c, kX[2].lw, kX[2].rw);
  float y = kY[0].a * integ(lo, hi, kY[0].c, kY[0].lw, kY[0].rw) +
            kY[1].a * integ(lo, hi, kY[1].c, kY[1].lw, kY[1].rw);
  float z = kZ[0].a * integ(lo, hi, kZ[0].c, kZ[0].lw, kZ[0].rw) +
            kZ[1].a * integ(lo, hi, kZ[1].c, kZ[1].lw, kZ[1].rw);
  return x + y + z;
}
}  // namespace cie

struct WlTable {
  std::vector<float> cdf;  // heap, index 1..65535
  std::vector<WlLeaf> leaf;
  float avg_integral;      // xyzMatchSumIntegral(380, 720) / 3 (luminance.wesl:14)
};
inline void build_wl_table(WlTable& t) {
  t.cdf.assign(CG_WL_LEAVES, 0.0f);
  t.leaf.assign(CG_WL_LEAVES, WlLeaf{});
  const float total = cie::sum_integral(380.0f, 720.0f);
  t.avg_integral = total / 3.0f;
  struct Item {
    uint32_t node;
    float low, high;
  };
  std::vector<Item> work{{1u, 380.0f, 720.0f}};
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    float mid = 0.5f * (it.low + it.high);
    if (it.node >= CG_WL_LEAVES) {
      // after 16 decisions: wavelength = 0.5 * (low + high)
      WlLeaf& L = t.leaf[it.node - CG_WL_LEAVES];
      float x, y, z;
      cie::match(mid, x, y, z);
      L.lambda = mid;
      L.prob = std::fmax(((x + y) + z) / total, 1e-6f);
      L.x = x;
      L.y = y;
      L.z = z;
      continue;
    }
    t.cdf[it.node] = cie::sum_integral(380.0f, mid) / total;
    work.push_back({2u * it.node, it.low, mid});       // cdf >= target: high = mid
    work.push_back({2u * it.node + 1u, mid, it.high});  // cdf <  target: low = mid
  }
}

// ---------------------------------------------------------------------------
// Upload: reference SceneData -> device layout.
inline void box_of(const float* lb, const float* ub, float* lo, float* hi) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = lb[k];
    hi[k] = ub[k];
  }
}

struct Layout {
  std::vector<Node8> bvh;       // scene BVH (TLAS and BLASes), 8-wide, exact boxes
  std::vector<Node8Q> bvhq;     // the same nodes with quantised boxes (what is uploaded)
  std::vector<PairNode> nodes;  // binary trees while building; then the emitter BVH
  std::vector<TriRec> tris;
  std::vector<InstRec> insts;
  std::vector<uint4> vidx;
  std::vector<VAttr> vattr;
  std::vector<MatRec> mats;
  std::vector<EmitRec> emit;
  std::vector<float2> alias;
  std::vector<uint32_t> emit_order;
  uint32_t emit_root = CG_REF_NONE;
  uint32_t root = CG_REF_NONE;
  uint32_t stack_depth = 0;
};

// Wavelength-independent records shared by both acceleration layouts:
// materials, vertex attributes, emitters, alias table; plus input checks.
inline int build_common(const cg_scene_data* sd, Layout& L, std::string& err) {
  // Materials.  A texture index past the array reads its last layer (WGSL
  // clamps array indices); with no texture array every lookup reads white,
  // which leaves the factor unchanged, so the slot is dropped.
  if (sd->tex_layers > 0 && (!sd->tex_rgba8 || sd->tex_width == 0 || sd->tex_height == 0))
    return set_error(err, CG_ERR_INVALID, "texture array has %u layers but no image data", sd->tex_layers);
  if (sd->tex_layers > 0 && (sd->tex_width > 16384u || sd->tex_height > 16384u))
    return set_error(err, CG_ERR_UNSUPPORTED, "texture array of %ux%u texels (max 16384x16384)", sd->tex_width,
                     sd->tex_height);
  if (sd->tex_layers >= CG_TEX_NONE)
    return set_error(err, CG_ERR_UNSUPPORTED, "%u texture layers (max %u)", sd->tex_layers, CG_TEX_NONE - 1u);
  auto layer = [&](const cg_tex_info& t) -> uint32_t {
    if (t.index < 0 || sd->tex_layers == 0) return CG_TEX_NONE;
    return std::min(static_cast<uint32_t>(t.index), sd->tex_layers - 1u);
  };
  for (uint32_t i = 0; i < sd->material_count; ++i) {
    const cg_material& m = sd->materials[i];
    MatRec r{};
    for (int k = 0; k < 3; ++k) r.base[k] = m.base_color_factor[k];
    r.metallic = m.metallic_factor;
    r.roughness = m.roughness_factor;
    r.transmission = m.transmission_factor;
    r.ior = m.ior;
    r.dispersion = m.dispersion;
    for (int k = 0; k < 3; ++k) r.emission[k] = m.emissive_factor[k];
    r.strength = m.emissive_strength;
    r.flags = m.thickness_factor > 0.0f ? CG_MAT_VOLUME : 0u;
    if (m.tex[CG_TEX_BASE_COLOR].tex_coord == 1u) r.flags |= CG_MAT_UV1_BASE;
    if (m.tex[CG_TEX_METALLIC_ROUGHNESS].tex_coord == 1u) r.flags |= CG_MAT_UV1_MR;
    if (m.tex[CG_TEX_TRANSMISSION].tex_coord == 1u) r.flags |= CG_MAT_UV1_TRANS;
    if (m.tex[CG_TEX_EMISSIVE].tex_coord == 1u) r.flags |= CG_MAT_UV1_EMIT;
    r.tex_base_mr = layer(m.tex[CG_TEX_BASE_COLOR]) | layer(m.tex[CG_TEX_METALLIC_ROUGHNESS]) << 16;
    r.tex_trans_emit = layer(m.tex[CG_TEX_TRANSMISSION]) | layer(m.tex[CG_TEX_EMISSIVE]) << 16;
    L.mats.push_back(r);
  }
  if (L.mats.empty()) return set_error(err, CG_ERR_INVALID, "scene has no materials");
  if (sd->env_width == 0 || sd->env_height == 0 || !sd->env_rgba32f)
    return set_error(err, CG_ERR_INVALID, "scene has no environment texture");
  if (sd->alias_count != sd->emissive_count)
    return set_error(err, CG_ERR_INVALID, "alias table (%u) and emissive buffer (%u) differ in length", sd->alias_count,
                sd->emissive_count);
  if (sd->index_count >= CG_REF_SLOT_MASK) return set_error(err, CG_ERR_INVALID, "too many triangles");

  // Vertices: attributes read for the final hit only.
  L.vattr.resize(sd->vertex_count);
  par_ranges(sd->vertex_count, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const cg_vertex& v = sd->vertices[i];
      VAttr& a = L.vattr[i];
      for (int k = 0; k < 3; ++k) a.nrm[k] = v.nrm[k];
      a.uv0[0] = v.uv0[0];
      a.uv0[1] = v.uv0[1];
      a.uv1[0] = v.uv1[0];
      a.uv1[1] = v.uv1[1];
      a._pad = 0.0f;
    }
  });
  std::atomic<uint32_t> bad{~0u};  // lowest slot with an index past the vertex buffer
  par_ranges(sd->index_count, [&](size_t b, size_t e) {
    for (size_t s = b; s < e; ++s) {
      const cg_index& ix = sd->indices[s];
      if (ix.i0 >= sd->vertex_count || ix.i1 >= sd->vertex_count || ix.i2 >= sd->vertex_count) {
        uint32_t cur = bad.load();
        while (s < cur && !bad.compare_exchange_weak(cur, static_cast<uint32_t>(s))) {
        }
        return;
      }
    }
  });
  if (bad.load() != ~0u)
    return set_error(err, CG_ERR_INVALID, "triangle slot %u indexes past the vertex buffer", bad.load());
  // Emitters and alias table.
  for (uint32_t i = 0; i < sd->emissive_count; ++i) {
    const cg_emissive_prim& e = sd->emissive[i];
    EmitRec r{};
    for (int k = 0; k < 3; ++k) {
      r.v0[k] = e.v[0][k];
      r.v1[k] = e.v[1][k];
      r.v2[k] = e.v[2][k];
      r.rgb[k] = e.emission[k];
    }
    f3 p0 = ld3(e.v[0]), p1 = ld3(e.v[1]), p2 = ld3(e.v[2]);
    f3 ln = normalize(cross(p1 - p0, p2 - p0));
    r.ln_x = ln.x;
    r.ln_y = ln.y;
    r.ln_z = ln.z;
    r.area_pdf = fmaxf(e.weight / e.area, 1e-3f);
    if (!std::isfinite(r.area_pdf) || !std::isfinite(ln.x))
      return set_error(err, CG_ERR_INVALID, "emissive triangle %u is degenerate", i);
    L.emit.push_back(r);
  }
  for (uint32_t i = 0; i < sd->alias_count; ++i) {
    float2 a;
    uint32_t al = sd->alias[i].alias;
    std::memcpy(&a.x, &al, 4);
    a.y = sd->alias[i].prob;
    L.alias.push_back(a);
  }
  return CG_OK;
}

inline TriRec make_tri(const cg_scene_data* sd, uint32_t s, uint32_t inst) {
  const cg_index& ix = sd->indices[s];
  const float* p0 = sd->vertices[ix.i0].pos;
  const float* p1 = sd->vertices[ix.i1].pos;
  const float* p2 = sd->vertices[ix.i2].pos;
  TriRec t;
  for (int k = 0; k < 3; ++k) {
    t.v0[k] = p0[k];
    t.e1[k] = p1[k] - p0[k];
    t.e2[k] = p2[k] - p0[k];
  }
  t.tri_id = sd->tri_ids ? sd->tri_ids[s] : s;
  t.inst = inst;
  t._pad = 0u;
  return t;
}

// The reference's own trees (median split, leaves of 1-2 triangles), one pair
// node per interior node: TLAS pairs first, BLAS pairs after.
inline int build_reference_accel(const cg_scene_data* sd, Layout& L, std::string& err) {
  L.tris.resize(sd->index_count);
  L.vidx.resize(sd->index_count);
  for (uint32_t s = 0; s < sd->index_count; ++s) {
    L.tris[s] = make_tri(sd, s, CG_MISS_ID);
    const cg_index& ix = sd->indices[s];
    L.vidx[s] = make_uint4(ix.i0, ix.i1, ix.i2, 0u);
  }
  // BLAS: one pair node per interior node (triangle_count == 0).
  const uint32_t nb = sd->blas_node_count;
  std::vector<uint32_t> blas_ref(nb, CG_REF_NONE);
  uint32_t pair_count = 0;
  for (uint32_t i = 0; i < nb; ++i)
    if (sd->blas_nodes[i].triangle_count == 0u) blas_ref[i] = pair_count++;
  for (uint32_t i = 0; i < nb; ++i) {
    const cg_blas_node& n = sd->blas_nodes[i];
    if (n.triangle_count == 0u) continue;
    if (n.triangle_count > 2u || n.child + n.triangle_count > sd->index_count)
      return set_error(err, CG_ERR_INVALID, "BLAS leaf %u is malformed", i);
    blas_ref[i] = CG_MAKE_LEAF(n.child, n.triangle_count);
  }
  // TLAS pair nodes come first; BLAS pair nodes follow at offset `tp`.
  const uint32_t nt = sd->tlas_node_count;
  std::vector<uint32_t> tlas_pair(nt, CG_REF_NONE);
  uint32_t tp = 0;
  for (uint32_t i = 0; i < nt; ++i)
    if (sd->tlas_nodes[i].child != 0u) tlas_pair[i] = tp++;
  auto bref = [&](uint32_t i) -> uint32_t {
    uint32_t r = blas_ref[i];
    return (r & CG_REF_LEAF) ? r : r + tp;
  };
  auto tref = [&](uint32_t i) -> uint32_t {
    const cg_tlas_node& n = sd->tlas_nodes[i];
    return n.child != 0u ? tlas_pair[i] : (CG_REF_INST | n.instance_idx);
  };
  L.nodes.resize(static_cast<size_t>(tp) + pair_count);
  for (uint32_t i = 0; i < nt; ++i) {
    const cg_tlas_node& n = sd->tlas_nodes[i];
    if (n.child == 0u) {
      if (n.instance_idx >= sd->instance_count)
        return set_error(err, CG_ERR_INVALID, "TLAS leaf %u names instance %u of %u", i, n.instance_idx,
                    sd->instance_count);
      continue;
    }
    if (n.child + 1u >= nt) return set_error(err, CG_ERR_INVALID, "TLAS node %u has children past the end", i);
    PairNode& p = L.nodes[tlas_pair[i]];
    const cg_tlas_node& l = sd->tlas_nodes[n.child];
    const cg_tlas_node& r = sd->tlas_nodes[n.child + 1u];
    box_of(l.lb, l.ub, p.lo_l, p.hi_l);
    box_of(r.lb, r.ub, p.lo_r, p.hi_r);
    p.ref_l = tref(n.child);
    p.ref_r = tref(n.child + 1u);
    p._pad[0] = p._pad[1] = 0u;
  }
  for (uint32_t i = 0; i < nb; ++i) {
    const cg_blas_node& n = sd->blas_nodes[i];
    if (n.triangle_count != 0u) continue;
    if (n.child + 1u >= nb) return set_error(err, CG_ERR_INVALID, "BLAS node %u has children past the end", i);
    PairNode& p = L.nodes[tp + blas_ref[i]];
    const cg_blas_node& l = sd->blas_nodes[n.child];
    const cg_blas_node& r = sd->blas_nodes[n.child + 1u];
    box_of(l.lb, l.ub, p.lo_l, p.hi_l);
    box_of(r.lb, r.ub, p.lo_r, p.hi_r);
    p.ref_l = bref(n.child);
    p.ref_r = bref(n.child + 1u);
    p._pad[0] = p._pad[1] = 0u;
  }
  // Instances.
  L.insts.resize(sd->instance_count);
  for (uint32_t i = 0; i < sd->instance_count; ++i) {
    const cg_instance& in = sd->instances[i];
    InstRec& r = L.insts[i];
    std::memcpy(r.inv, in.inv_trans, sizeof(r.inv));
    if (in.blas_root_idx >= nb) return set_error(err, CG_ERR_INVALID, "instance %u names BLAS root %u", i, in.blas_root_idx);
    r.root = bref(in.blas_root_idx);
    r.material = in.material_idx;
    r._pad[0] = r._pad[1] = 0u;
  }
  L.root = nt ? tref(0) : CG_REF_NONE;
  // Stack depth: longest TLAS path + longest BLAS path + the marker.
  auto depth_of = [](uint32_t n_nodes, auto child_of) {
    std::vector<uint32_t> depth(n_nodes, 0u);
    uint32_t best = 0;
    for (uint32_t i = 0; i < n_nodes; ++i) {  // parents precede children
      uint32_t c = child_of(i);
      if (c != 0u && c + 1u < n_nodes) {
        depth[c] = depth[c + 1u] = depth[i] + 1u;
      }
      best = std::max(best, depth[i]);
    }
    return best;
  };
  uint32_t dt = depth_of(nt, [&](uint32_t i) { return sd->tlas_nodes[i].child; });
  uint32_t dbl = 0;
  {
    std::vector<uint32_t> depth(nb, 0u);
    for (uint32_t i = 0; i < nb; ++i) {
      const cg_blas_node& n = sd->blas_nodes[i];
      if (n.triangle_count == 0u && n.child + 1u < nb) depth[n.child] = depth[n.child + 1u] = depth[i] + 1u;
      dbl = std::max(dbl, depth[i]);
    }
  }
  L.stack_depth = dt + dbl + 2u;
  return CG_OK;
}

// ---------------------------------------------------------------------------
// Binned SAH build (16 bins per axis) over primitive boxes; leaves of at most
// `max_leaf` primitives.  Nodes are stored parents first, left subtree before
// right (depth-first), so a pair node's left child follows it in memory.
struct SahTree {
  struct Prim {
    float lo[3], hi[3], c[3];
  };
  struct Node {
    float lo[3], hi[3];
    uint32_t left = 0, right = 0, first = 0, count = 0;  // leaf iff count > 0
  };
  std::vector<Node> nodes;
  std::vector<uint32_t> order;  // primitive ids in leaf order
  uint32_t depth = 0;

  static double area(const float* lo, const float* hi) {
    double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }

  // Subtrees over disjoint ranges of the primitive array are built on
  // separate threads once a node holds more than kParallel primitives; nodes
  // come from a pre-sized array through an atomic counter (a node's children
  // are adjacent).  Primitive records are partitioned in place, so every pass
  // over a node reads memory sequentially; one pass bins all three axes.
  void build(const std::vector<Prim>& prims, uint32_t max_leaf) {
    const uint32_t n = static_cast<uint32_t>(prims.size());
    nodes.clear();
    order.clear();
    depth = 0;
    if (n == 0) return;
    std::vector<Ref> refs(n);
    for (uint32_t i = 0; i < n; ++i) refs[i] = Ref{prims[i], i};
    nodes.resize(2 * static_cast<size_t>(n));
    std::atomic<uint32_t> next{1u};
    std::atomic<uint32_t> max_depth{0u};
    build_rec(refs, max_leaf, 0u, 0u, n, 0u, next, max_depth, 3);
    nodes.resize(next.load());
    depth = max_depth.load();
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = refs[i].id;
  }

 private:
  struct Ref {
    Prim p;
    uint32_t id;
  };
  static constexpr uint32_t kParallel = 1u << 16;
  static constexpr int kBins = 16;
  struct Bins {
    uint32_t cnt[3][kBins];
    float lo[3][kBins][3], hi[3][kBins][3];
    void clear() {
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < kBins; ++b) {
          cnt[a][b] = 0;
          for (int k = 0; k < 3; ++k) {
            lo[a][b][k] = INFINITY;
            hi[a][b][k] = -INFINITY;
          }
        }
    }
  };
  void build_rec(std::vector<Ref>& refs, uint32_t max_leaf, uint32_t node, uint32_t first, uint32_t count, uint32_t d,
                 std::atomic<uint32_t>& next, std::atomic<uint32_t>& max_depth, int par) {
    for (;;) {
      uint32_t cur = max_depth.load();
      while (d > cur && !max_depth.compare_exchange_weak(cur, d)) {
      }
      const int chunks = count > (1u << 20) ? 8 : 1;
      auto for_chunks = [&](auto&& fn) {
        if (chunks == 1) {
          fn(0, first, first + count);
          return;
        }
        std::vector<std::thread> pool;
        for (int c = 0; c < chunks; ++c) {
          uint32_t b = first + static_cast<uint32_t>(uint64_t(count) * c / chunks);
          uint32_t e = first + static_cast<uint32_t>(uint64_t(count) * (c + 1) / chunks);
          pool.emplace_back([&fn, c, b, e] { fn(c, b, e); });
        }
        for (auto& t : pool) t.join();
      };
      // bounds and centroid bounds
      float plo[8][3], phi[8][3], pclo[8][3], pchi[8][3];
      for_chunks([&](int c, uint32_t b, uint32_t e) {
        for (int k = 0; k < 3; ++k) {
          plo[c][k] = pclo[c][k] = INFINITY;
          phi[c][k] = pchi[c][k] = -INFINITY;
        }
        for (uint32_t i = b; i < e; ++i) {
          const Prim& p = refs[i].p;
          for (int k = 0; k < 3; ++k) {
            plo[c][k] = std::min(plo[c][k], p.lo[k]);
            phi[c][k] = std::max(phi[c][k], p.hi[k]);
            pclo[c][k] = std::min(pclo[c][k], p.c[k]);
            pchi[c][k] = std::max(pchi[c][k], p.c[k]);
          }
        }
      });
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int c = 0; c < chunks; ++c)
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], plo[c][k]);
          hi[k] = std::max(hi[k], phi[c][k]);
          clo[k] = std::min(clo[k], pclo[c][k]);
          chi[k] = std::max(chi[k], pchi[c][k]);
        }
      Node& nd = nodes[node];
      for (int k = 0; k < 3; ++k) {
        nd.lo[k] = lo[k];
        nd.hi[k] = hi[k];
      }
      if (count <= 1) {
        nd.first = first;
        nd.count = count;
        return;
      }
      // bin all three axes in one pass
      double scale[3];
      for (int a = 0; a < 3; ++a) scale[a] = chi[a] > clo[a] ? kBins / double(chi[a] - clo[a]) : 0.0;
      Bins local;
      std::vector<Bins> heap(chunks > 1 ? chunks : 0);
      Bins* part = chunks > 1 ? heap.data() : &local;
      for_chunks([&](int c, uint32_t b0, uint32_t e0) {
        Bins& B = part[c];
        B.clear();
        for (uint32_t i = b0; i < e0; ++i) {
          const Prim& p = refs[i].p;
          for (int a = 0; a < 3; ++a) {
            int b = std::min(kBins - 1, static_cast<int>((p.c[a] - clo[a]) * scale[a]));
            B.cnt[a][b]++;
            for (int k = 0; k < 3; ++k) {
              B.lo[a][b][k] = std::min(B.lo[a][b][k], p.lo[k]);
              B.hi[a][b][k] = std::max(B.hi[a][b][k], p.hi[k]);
            }
          }
        }
      });
      Bins& B = part[0];
      for (int c = 1; c < chunks; ++c)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < kBins; ++b) {
            B.cnt[a][b] += part[c].cnt[a][b];
            for (int k = 0; k < 3; ++k) {
              B.lo[a][b][k] = std::min(B.lo[a][b][k], part[c].lo[a][b][k]);
              B.hi[a][b][k] = std::max(B.hi[a][b][k], part[c].hi[a][b][k]);
            }
          }
      double best_cost = INFINITY;
      int best_axis = -1, best_plane = 0;
      for (int a = 0; a < 3; ++a) {
        if (scale[a] == 0.0) continue;
        double right_area[kBins];
        uint32_t right_cnt[kBins];
        float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t rc = 0;
        for (int b = kBins - 1; b > 0; --b) {
          rc += B.cnt[a][b];
          for (int k = 0; k < 3; ++k) {
            rlo[k] = std::min(rlo[k], B.lo[a][b][k]);
            rhi[k] = std::max(rhi[k], B.hi[a][b][k]);
          }
          right_area[b] = area(rlo, rhi);
          right_cnt[b] = rc;
        }
        float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t lc = 0;
        for (int b = 0; b < kBins - 1; ++b) {
          lc += B.cnt[a][b];
          for (int k = 0; k < 3; ++k) {
            llo[k] = std::min(llo[k], B.lo[a][b][k]);
            lhi[k] = std::max(lhi[k], B.hi[a][b][k]);
          }
          if (lc == 0 || right_cnt[b + 1] == 0) continue;
          double cst = area(llo, lhi) * lc + right_area[b + 1] * right_cnt[b + 1];
          if (cst < best_cost) {
            best_cost = cst;
            best_axis = a;
            best_plane = b;
          }
        }
      }
      const double parent_area = std::max(area(lo, hi), 1e-30);
      const double split_cost = 1.0 + best_cost / parent_area;  // traversal 1, intersection 1 per primitive
      if (count <= max_leaf && (best_axis < 0 || double(count) <= split_cost)) {
        nd.first = first;
        nd.count = count;
        return;
      }
      uint32_t mid;
      if (best_axis < 0) {  // all centroids equal: split by count
        mid = first + count / 2;
      } else {
        const double sc = scale[best_axis];
        const float c0 = clo[best_axis];
        auto it = std::partition(refs.begin() + first, refs.begin() + first + count, [&](const Ref& r) {
          int b = std::min(kBins - 1, static_cast<int>((r.p.c[best_axis] - c0) * sc));
          return b <= best_plane;
        });
        mid = static_cast<uint32_t>(it - refs.begin());
        if (mid == first || mid == first + count) mid = first + count / 2;
      }
      const uint32_t left = next.fetch_add(2u);
      nodes[node].left = left;
      nodes[node].right = left + 1;
      const uint32_t lcount = mid - first, rcount = first + count - mid;
      if (par > 0 && count > kParallel) {
        std::thread t([&, left, first, lcount, d, par] {
          build_rec(refs, max_leaf, left, first, lcount, d + 1, next, max_depth, par - 1);
        });
        build_rec(refs, max_leaf, left + 1, mid, rcount, d + 1, next, max_depth, par - 1);
        t.join();
        return;
      }
      build_rec(refs, max_leaf, left, first, lcount, d + 1, next, max_depth, 0);
      node = left + 1;  // continue with the right child in this frame
      first = mid;
      count = rcount;
      ++d;
    }
  }
};

// Pair nodes for a SAH tree: one per interior node, appended to `out`.
// Leaves become BLAS leaf refs (slot_base + first) or, for a TLAS
// (`inst_ids` given), instance refs.  Returns the root ref.
inline uint32_t emit_pairs(const SahTree& t, std::vector<PairNode>& out, uint32_t slot_base,
                           const std::vector<uint32_t>* inst_ids, float pad) {
  if (t.nodes.empty()) return CG_REF_NONE;
  std::vector<uint32_t> pair_of(t.nodes.size(), CG_REF_NONE);
  const uint32_t base = static_cast<uint32_t>(out.size());
  uint32_t next = base;
  for (size_t i = 0; i < t.nodes.size(); ++i)
    if (t.nodes[i].count == 0) pair_of[i] = next++;
  out.resize(next);
  auto ref = [&](uint32_t i) -> uint32_t {
    const SahTree::Node& n = t.nodes[i];
    if (n.count == 0) return pair_of[i];
    if (inst_ids) return CG_REF_INST | (*inst_ids)[t.order[n.first]];
    return CG_MAKE_LEAF(slot_base + n.first, n.count);
  };
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    const SahTree::Node& n = t.nodes[i];
    if (n.count != 0) continue;
    PairNode& p = out[pair_of[i]];
    const SahTree::Node& l = t.nodes[n.left];
    const SahTree::Node& r = t.nodes[n.right];
    for (int k = 0; k < 3; ++k) {
      p.lo_l[k] = l.lo[k] - pad;
      p.hi_l[k] = l.hi[k] + pad;
      p.lo_r[k] = r.lo[k] - pad;
      p.hi_r[k] = r.hi[k] + pad;
    }
    p.ref_l = ref(n.left);
    p.ref_r = ref(n.right);
    p._pad[0] = p._pad[1] = 0u;
  }
  return ref(0);
}

// Rebuilt acceleration structure (same closest hits, ties aside):
//  * every distinct BLAS rebuilt with binned SAH, binary leaves of up to
//    `blas_leaf` triangles (the wide collapse merges them into leaves of up to
//    CG_LEAF_MAX where that is cheaper);
//  * instances whose inverse transform is the identity and whose BLAS holds at
//    most 16 triangles (e.g. one-triangle emitters) merged into one world-space
//    BLAS whose triangles carry their instance index -- their local rays equal
//    the world rays, so the same triangle tests run on the same rays;
//  * a SAH TLAS over the remaining instances and the merged BLAS.
// A BLAS left to the linear-BVH build (CG_ACCEL_DEVICE): its triangles sit in
// L.tris / L.vidx at [slot_base, slot_base + boxes.size()) in source order
// until the build sorts them; instances refer to it through the placeholder
// root CG_REF_INST | job index (an instance ref is never a BLAS root) until
// its nodes are placed.
struct BlasJob {
  uint32_t slot_base = 0;
  std::vector<PrimBox> boxes;
  float pad = 0.0f;
  uint32_t root = CG_REF_NONE;   // global root ref once built
  uint32_t nodes = 0;            // wide nodes
  uint32_t stack = 0;            // deepest stack the BLAS needs
};

inline int build_sah_accel(const cg_scene_data* sd, Layout& L, std::string& err,
                           std::vector<BlasJob>* jobs = nullptr, uint32_t blas_leaf = 2u) {
  const uint32_t nb = sd->blas_node_count, ni = sd->instance_count, nt = sd->tlas_node_count;
  // World boxes of the instances: the reference TLAS leaves.
  std::vector<int> have_box(ni, 0);
  std::vector<SahTree::Prim> inst_box(ni);
  for (uint32_t i = 0; i < nt; ++i) {
    const cg_tlas_node& n = sd->tlas_nodes[i];
    if (n.child != 0u) {
      if (n.child + 1u >= nt) return set_error(err, CG_ERR_INVALID, "TLAS node %u has children past the end", i);
      continue;
    }
    if (n.instance_idx >= ni)
      return set_error(err, CG_ERR_INVALID, "TLAS leaf %u names instance %u of %u", i, n.instance_idx, ni);
    SahTree::Prim& p = inst_box[n.instance_idx];
    for (int k = 0; k < 3; ++k) {
      p.lo[k] = n.lb[k];
      p.hi[k] = n.ub[k];
      p.c[k] = 0.5f * (n.lb[k] + n.ub[k]);
    }
    have_box[n.instance_idx] = 1;
  }
  // Triangle slots of each BLAS root (walk the reference tree).
  auto slots_of = [&](uint32_t root, std::vector<uint32_t>& out) -> bool {
    std::vector<uint32_t> st{root};
    while (!st.empty()) {
      uint32_t i = st.back();
      st.pop_back();
      if (i >= nb) return false;
      const cg_blas_node& n = sd->blas_nodes[i];
      if (n.triangle_count == 0u) {
        if (n.child + 1u >= nb) return false;
        st.push_back(n.child + 1u);
        st.push_back(n.child);
      } else {
        if (n.triangle_count > 2u || n.child + n.triangle_count > sd->index_count) return false;
        for (uint32_t k = 0; k < n.triangle_count; ++k) out.push_back(n.child + k);
      }
    }
    return true;
  };
  auto is_identity = [](const float* m) {
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 4; ++r)
        if (m[4 * c + r] != (c == r ? 1.0f : 0.0f)) return false;
    return true;
  };
  // A BLAS for the device build: source-order triangles, a placeholder root;
  // four triangles or fewer make one leaf.
  auto add_job = [&](const std::vector<SahTree::Prim>& prims, float pad) -> uint32_t {
    const uint32_t base = static_cast<uint32_t>(L.tris.size()) - static_cast<uint32_t>(prims.size());
    if (prims.size() <= CG_LEAF_MAX) return CG_MAKE_LEAF(base, static_cast<uint32_t>(prims.size()));
    BlasJob j;
    j.slot_base = base;
    j.pad = pad;
    j.boxes.resize(prims.size());
    par_ranges(prims.size(), [&](size_t b, size_t e) {
      for (size_t k = b; k < e; ++k)
        for (int a = 0; a < 3; ++a) {
          j.boxes[k].lo[a] = prims[k].lo[a];
          j.boxes[k].hi[a] = prims[k].hi[a];
        }
    });
    jobs->push_back(std::move(j));
    return CG_REF_INST | static_cast<uint32_t>(jobs->size() - 1);
  };
  auto tri_prim = [&](uint32_t s) {
    const cg_index& ix = sd->indices[s];
    const float* p[3] = {sd->vertices[ix.i0].pos, sd->vertices[ix.i1].pos, sd->vertices[ix.i2].pos};
    SahTree::Prim q;
    for (int k = 0; k < 3; ++k) {
      q.lo[k] = std::min(std::min(p[0][k], p[1][k]), p[2][k]);
      q.hi[k] = std::max(std::max(p[0][k], p[1][k]), p[2][k]);
      q.c[k] = 0.5f * (q.lo[k] + q.hi[k]);
    }
    return q;
  };
  // Box padding: a small multiple of the coordinates' magnitude keeps every
  // box test conservative against the rounding of the triangle test.
  auto pad_of = [](const std::vector<SahTree::Prim>& prims) {
    std::vector<float> part(16, 1.0f);  // a maximum: the same in any grouping
    std::atomic<uint32_t> next{0};
    par_ranges(prims.size(), [&](size_t b, size_t e) {
      float m = 1.0f;
      for (size_t i = b; i < e; ++i)
        for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs(prims[i].lo[k]), std::fabs(prims[i].hi[k])));
      part[next.fetch_add(1) % 16] = m;
    });
    float m = 1.0f;
    for (float v : part) m = std::max(m, v);
    return 4e-5f * m;
  };

  L.insts.resize(ni);
  std::vector<uint32_t> merged, kept;
  std::vector<std::vector<uint32_t>> root_slots;
  std::vector<int> root_index(nb, -1);
  for (uint32_t i = 0; i < ni; ++i) {
    const cg_instance& in = sd->instances[i];
    InstRec& r = L.insts[i];
    std::memcpy(r.inv, in.inv_trans, sizeof(r.inv));
    r.material = in.material_idx;
    r._pad[0] = r._pad[1] = 0u;
    r.root = CG_REF_NONE;
    if (in.blas_root_idx >= nb) return set_error(err, CG_ERR_INVALID, "instance %u names BLAS root %u", i, in.blas_root_idx);
    if (!have_box[i]) continue;  // not reachable from the TLAS
    if (root_index[in.blas_root_idx] < 0) {
      root_index[in.blas_root_idx] = static_cast<int>(root_slots.size());
      root_slots.emplace_back();
      if (!slots_of(in.blas_root_idx, root_slots.back()))
        return set_error(err, CG_ERR_INVALID, "BLAS rooted at %u is malformed", in.blas_root_idx);
    }
    const auto& sl = root_slots[root_index[in.blas_root_idx]];
    if (is_identity(in.inv_trans) && sl.size() <= 16u)
      merged.push_back(i);
    else
      kept.push_back(i);
  }
  uint32_t max_blas_depth = 0;
  // Distinct BLAS of kept instances.
  std::vector<uint32_t> built_root(root_slots.size(), CG_REF_NONE);
  for (uint32_t i : kept) {
    const int ri = root_index[sd->instances[i].blas_root_idx];
    if (built_root[ri] == CG_REF_NONE && !root_slots[ri].empty()) {
      const auto& sl = root_slots[ri];
      std::vector<SahTree::Prim> prims(sl.size());
      par_ranges(sl.size(), [&](size_t b, size_t e) {
        for (size_t k = b; k < e; ++k) prims[k] = tri_prim(sl[k]);
      });
      // triangle records in the order `order` gives (identity when null)
      auto append = [&](const std::vector<uint32_t>* order) {
        const size_t base = L.tris.size();
        L.tris.resize(base + sl.size());
        L.vidx.resize(base + sl.size());
        par_ranges(sl.size(), [&](size_t b, size_t e) {
          for (size_t k = b; k < e; ++k) {
            const uint32_t sidx = sl[order ? (*order)[k] : k];
            L.tris[base + k] = make_tri(sd, sidx, CG_MISS_ID);
            const cg_index& ix = sd->indices[sidx];
            L.vidx[base + k] = make_uint4(ix.i0, ix.i1, ix.i2, 0u);
          }
        });
      };
      if (jobs) {
        append(nullptr);
        built_root[ri] = add_job(prims, pad_of(prims));
        L.insts[i].root = built_root[ri];
        continue;
      }
      SahTree t;
      t.build(prims, blas_leaf);
      const uint32_t slot_base = static_cast<uint32_t>(L.tris.size());
      append(&t.order);
      built_root[ri] = 0;  // placeholder until pair nodes are placed
      const uint32_t r = emit_pairs(t, L.nodes, slot_base, nullptr, pad_of(prims));
      built_root[ri] = r;
      max_blas_depth = std::max(max_blas_depth, t.depth);
    }
    L.insts[i].root = built_root[ri];
  }
  // Merged world BLAS.
  std::vector<SahTree::Prim> tlas_prims;
  std::vector<uint32_t> tlas_ids;
  for (uint32_t i : kept) {
    if (L.insts[i].root == CG_REF_NONE) continue;
    tlas_prims.push_back(inst_box[i]);
    tlas_ids.push_back(i);
  }
  if (!merged.empty()) {
    std::vector<SahTree::Prim> prims;
    std::vector<std::pair<uint32_t, uint32_t>> tri_inst;  // (slot, instance)
    for (uint32_t i : merged)
      for (uint32_t s : root_slots[root_index[sd->instances[i].blas_root_idx]]) {
        prims.push_back(tri_prim(s));
        tri_inst.emplace_back(s, i);
      }
    const float pad = pad_of(prims);
    InstRec w{};
    for (int k = 0; k < 16; ++k) w.inv[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    w.material = 0u;
    SahTree::Prim box;
    if (jobs) {
      for (const auto& ti : tri_inst) {
        L.tris.push_back(make_tri(sd, ti.first, ti.second));
        const cg_index& ix = sd->indices[ti.first];
        L.vidx.push_back(make_uint4(ix.i0, ix.i1, ix.i2, 0u));
      }
      w.root = add_job(prims, pad);
      for (int k = 0; k < 3; ++k) {
        box.lo[k] = INFINITY;
        box.hi[k] = -INFINITY;
        for (const auto& p : prims) {
          box.lo[k] = std::min(box.lo[k], p.lo[k]);
          box.hi[k] = std::max(box.hi[k], p.hi[k]);
        }
      }
    } else {
      SahTree t;
      t.build(prims, blas_leaf);
      const uint32_t slot_base = static_cast<uint32_t>(L.tris.size());
      for (uint32_t id : t.order) {
        L.tris.push_back(make_tri(sd, tri_inst[id].first, tri_inst[id].second));
        const cg_index& ix = sd->indices[tri_inst[id].first];
        L.vidx.push_back(make_uint4(ix.i0, ix.i1, ix.i2, 0u));
      }
      w.root = emit_pairs(t, L.nodes, slot_base, nullptr, pad);
      for (int k = 0; k < 3; ++k) {
        box.lo[k] = t.nodes[0].lo[k];
        box.hi[k] = t.nodes[0].hi[k];
      }
      max_blas_depth = std::max(max_blas_depth, t.depth);
    }
    for (int k = 0; k < 3; ++k) {
      box.lo[k] -= pad;
      box.hi[k] += pad;
      box.c[k] = 0.5f * (box.lo[k] + box.hi[k]);
    }
    const uint32_t wi = static_cast<uint32_t>(L.insts.size());
    L.insts.push_back(w);
    tlas_prims.push_back(box);
    tlas_ids.push_back(wi);
  }
  if (L.tris.size() >= CG_REF_SLOT_MASK) return set_error(err, CG_ERR_INVALID, "too many triangles");
  // TLAS over instance boxes (already padded), one instance per leaf.
  SahTree tl;
  tl.build(tlas_prims, 1u);
  L.root = emit_pairs(tl, L.nodes, 0u, &tlas_ids, 0.0f);
  L.stack_depth = tl.depth + max_blas_depth + 2u;
  return CG_OK;
}

// BVH over the emitters for the light pdf's search (appended pair nodes,
// leaves index emit_order).  Boxes are padded like the triangle boxes.
inline void build_emitter_bvh(Layout& L) {
  if (L.emit.size() < 2) return;
  std::vector<SahTree::Prim> prims(L.emit.size());
  float m = 1.0f;
  for (size_t i = 0; i < L.emit.size(); ++i) {
    const EmitRec& e = L.emit[i];
    for (int k = 0; k < 3; ++k) {
      prims[i].lo[k] = std::min(std::min(e.v0[k], e.v1[k]), e.v2[k]);
      prims[i].hi[k] = std::max(std::max(e.v0[k], e.v1[k]), e.v2[k]);
      prims[i].c[k] = 0.5f * (prims[i].lo[k] + prims[i].hi[k]);
      m = std::max(m, std::max(std::fabs(prims[i].lo[k]), std::fabs(prims[i].hi[k])));
    }
  }
  SahTree t;
  t.build(prims, CG_LEAF_MAX);
  L.emit_order = t.order;
  L.emit_root = emit_pairs(t, L.nodes, 0u, nullptr, 4e-5f * m);
  L.stack_depth = std::max(L.stack_depth, t.depth + 1u);
}

// ---------------------------------------------------------------------------
// 8-wide BVH from a binary one: each wide node takes an interior binary
// node's two children and repeatedly replaces the interior child of largest
// surface area by its own two children, up to eight.  Every child box is a
// box of the binary tree, bit for bit.  A traversal visits about a third as
// many nodes as the binary walk, each one 128-byte line.  Nodes are stored
// parent first, depth first.
struct CollapseWide {
  const std::vector<PairNode>& pairs;
  std::vector<Node8>& out;
  std::vector<uint32_t> map;
  explicit CollapseWide(const std::vector<PairNode>& p, std::vector<Node8>& o)
      : pairs(p), out(o), map(p.size(), CG_REF_NONE) {}
  static bool interior(uint32_t r) { return r != CG_REF_NONE && (r & (CG_REF_LEAF | CG_REF_INST)) == 0u; }
  static float half_area(const float* lo, const float* hi) {
    const float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    return x * y + y * z + z * x;
  }
  struct Child {
    uint32_t ref;
    float lo[3], hi[3];
  };
  static Child left(const PairNode& p) {
    return Child{p.ref_l, {p.lo_l[0], p.lo_l[1], p.lo_l[2]}, {p.hi_l[0], p.hi_l[1], p.hi_l[2]}};
  }
  static Child right(const PairNode& p) {
    return Child{p.ref_r, {p.lo_r[0], p.lo_r[1], p.lo_r[2]}, {p.hi_r[0], p.hi_r[1], p.hi_r[2]}};
  }
  uint32_t convert(uint32_t r) {
    if (!interior(r)) return r;
    if (map[r] != CG_REF_NONE) return map[r];
    Child ch[CG_WIDE];
    int n = 2;
    ch[0] = left(pairs[r]);
    ch[1] = right(pairs[r]);
    while (n < CG_WIDE) {
      int best = -1;
      float ba = -1.0f;
      for (int i = 0; i < n; ++i)
        if (interior(ch[i].ref)) {
          const float a = half_area(ch[i].lo, ch[i].hi);
          if (a > ba) {
            ba = a;
            best = i;
          }
        }
      if (best < 0) break;
      const PairNode& q = pairs[ch[best].ref];
      ch[best] = left(q);
      ch[n++] = right(q);
    }
    const uint32_t idx = static_cast<uint32_t>(out.size());
    map[r] = idx;
    out.emplace_back();
    Node8 nd;
    for (int i = 0; i < CG_WIDE; ++i) {
      const bool used = i < n;
      for (int a = 0; a < 3; ++a) {
        nd.lo[a][i] = used ? ch[i].lo[a] : __builtin_inff();
        nd.hi[a][i] = used ? ch[i].hi[a] : -__builtin_inff();
      }
      nd.ref[i] = CG_REF_NONE;
    }
    for (int i = 0; i < n; ++i) nd.ref[i] = convert(ch[i].ref);
    out[idx] = nd;
    return idx;
  }
};

// 8-wide BVH from a binary one by the SAH-optimal collapse (Ylitie, Karras and
// Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs", HPG 2017), priced for this traversal: a node visit is one step, a
// leaf one step per two triangles, entering an instance one step, each
// weighted by the area of the box whose hit makes the ray take it.  For a
// binary node n and a budget of i = 1..8 slots, Entry::cost[i - 1] is the
// least cost of n's subtree given as at most i children of the wide node
// above: n itself (a wide node, or one leaf when its subtree is at most
// CG_LEAF_MAX triangles in consecutive slots), or n's two subtrees sharing the
// i slots.  Every child box is a box of the binary tree, bit for bit; nodes
// are stored parent first, depth first.
struct CollapseWideSah {
  static constexpr float kNode = 1.0f, kTriStep = 1.0f, kInst = 1.0f;
  struct Entry {
    float cost[CG_WIDE];
    uint8_t kl[CG_WIDE], kr[CG_WIDE];  // the subtrees' budgets; both 0: n itself
    uint8_t il, ir;                    // the subtrees' budgets when n is a wide node
    uint8_t leaf, done;                // n itself is one leaf / entry computed
    uint32_t first, count;             // the subtree's slot run (count 0: none)
    float area;                        // half area of n's box (held by its parent)
  };
  const std::vector<PairNode>& pairs;
  std::vector<Node8>& out;
  std::vector<Entry> e;
  std::vector<uint32_t> map;
  CollapseWideSah(const std::vector<PairNode>& p, std::vector<Node8>& o)
      : pairs(p), out(o), e(p.size()), map(p.size(), CG_REF_NONE) {
    for (Entry& x : e) x.done = 0;
  }
  static bool interior(uint32_t r) { return CollapseWide::interior(r); }
  static float leaf_steps(uint32_t count) { return static_cast<float>((count + 1u) / 2u); }
  float child_cost(uint32_t r, const float* lo, const float* hi, int budget) const {
    if (interior(r)) return e[r].cost[budget - 1];
    if (r == CG_REF_NONE) return 0.0f;
    const float a = CollapseWide::half_area(lo, hi);
    if ((r & CG_REF_LEAF) != 0u) return a * kTriStep * leaf_steps(((r >> CG_REF_COUNT_SHIFT) & 3u) + 1u);
    return a * kInst;
  }
  void run_of(uint32_t r, uint32_t& first, uint32_t& count) const {
    first = 0u;
    count = 0u;
    if (interior(r)) {
      first = e[r].first;
      count = e[r].count;
    } else if (r != CG_REF_NONE && (r & CG_REF_LEAF) != 0u) {
      first = r & CG_REF_SLOT_MASK;
      count = ((r >> CG_REF_COUNT_SHIFT) & 3u) + 1u;
    }
  }
  void evaluate(uint32_t n) {
    const PairNode& p = pairs[n];
    Entry& x = e[n];
    uint32_t fl, cl, fr, cr;
    run_of(p.ref_l, fl, cl);
    run_of(p.ref_r, fr, cr);
    x.first = 0u;
    x.count = 0u;
    if (cl != 0u && cr != 0u && cl + cr <= CG_LEAF_MAX) {
      if (fl + cl == fr) {
        x.first = fl;
        x.count = cl + cr;
      } else if (fr + cr == fl) {
        x.first = fr;
        x.count = cl + cr;
      }
    }
    float dist[CG_WIDE + 1];
    uint8_t dl[CG_WIDE + 1], dr[CG_WIDE + 1];
    for (int j = 2; j <= CG_WIDE; ++j) {
      dist[j] = INFINITY;
      dl[j] = 1u;  // a valid split even when every cost overflows or is NaN
      dr[j] = static_cast<uint8_t>(j - 1);
      for (int k = 1; k < j; ++k) {
        const float c = child_cost(p.ref_l, p.lo_l, p.hi_l, k) + child_cost(p.ref_r, p.lo_r, p.hi_r, j - k);
        if (c < dist[j]) {
          dist[j] = c;
          dl[j] = static_cast<uint8_t>(k);
          dr[j] = static_cast<uint8_t>(j - k);
        }
      }
    }
    const float internal = x.area * kNode + dist[CG_WIDE];
    const float leaf = x.count != 0u ? x.area * kTriStep * leaf_steps(x.count) : INFINITY;
    x.il = dl[CG_WIDE];
    x.ir = dr[CG_WIDE];
    x.leaf = x.count != 0u && leaf <= internal ? 1u : 0u;  // (both costs may be infinite)
    x.cost[0] = std::min(internal, leaf);
    x.kl[0] = x.kr[0] = 0u;
    for (int i = 2; i <= CG_WIDE; ++i) {
      if (dist[i] < x.cost[i - 2]) {
        x.cost[i - 1] = dist[i];
        x.kl[i - 1] = dl[i];
        x.kr[i - 1] = dr[i];
      } else {
        x.cost[i - 1] = x.cost[i - 2];
        x.kl[i - 1] = x.kl[i - 2];
        x.kr[i - 1] = x.kr[i - 2];
      }
    }
    x.done = 1u;
  }
  // Children before parents, from one root (an explicit stack: binary trees
  // over one triangle per leaf run deep).
  void solve(uint32_t root) {
    if (!interior(root) || e[root].done) return;
    std::vector<std::pair<uint32_t, bool>> st{{root, false}};
    while (!st.empty()) {
      const auto [n, expanded] = st.back();
      st.pop_back();
      if (e[n].done) continue;
      if (expanded) {
        evaluate(n);
        continue;
      }
      st.push_back({n, true});
      const PairNode& p = pairs[n];
      if (interior(p.ref_l)) {
        e[p.ref_l].area = CollapseWide::half_area(p.lo_l, p.hi_l);
        if (!e[p.ref_l].done) st.push_back({p.ref_l, false});
      }
      if (interior(p.ref_r)) {
        e[p.ref_r].area = CollapseWide::half_area(p.lo_r, p.hi_r);
        if (!e[p.ref_r].done) st.push_back({p.ref_r, false});
      }
    }
  }
  void gather(uint32_t r, const float* lo, const float* hi, int budget, CollapseWide::Child* ch, int& n) const {
    if (interior(r)) {
      const Entry& x = e[r];
      if (x.kl[budget - 1] != 0u) {
        const PairNode& p = pairs[r];
        gather(p.ref_l, p.lo_l, p.hi_l, x.kl[budget - 1], ch, n);
        gather(p.ref_r, p.lo_r, p.hi_r, x.kr[budget - 1], ch, n);
        return;
      }
      if (x.leaf) r = CG_MAKE_LEAF(x.first, x.count);
    }
    ch[n++] = CollapseWide::Child{r, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
  }
  uint32_t convert(uint32_t r) {
    if (!interior(r)) return r;
    if (map[r] != CG_REF_NONE) return map[r];
    solve(r);
    CollapseWide::Child ch[CG_WIDE];
    int n = 0;
    const PairNode& p = pairs[r];
    gather(p.ref_l, p.lo_l, p.hi_l, e[r].il, ch, n);
    gather(p.ref_r, p.lo_r, p.hi_r, e[r].ir, ch, n);
    const uint32_t idx = static_cast<uint32_t>(out.size());
    map[r] = idx;
    out.emplace_back();
    Node8 nd;
    for (int i = 0; i < CG_WIDE; ++i) {
      const bool used = i < n;
      for (int a = 0; a < 3; ++a) {
        nd.lo[a][i] = used ? ch[i].lo[a] : __builtin_inff();
        nd.hi[a][i] = used ? ch[i].hi[a] : -__builtin_inff();
      }
      nd.ref[i] = CG_REF_NONE;
    }
    for (int i = 0; i < n; ++i) nd.ref[i] = convert(ch[i].ref);
    out[idx] = nd;
    return idx;
  }
};

// Deepest traversal stack of a wide tree: a node with k children pushes at
// most k - 1 of them and descends into the nearest; entering an instance
// pushes one marker.
inline uint32_t stack_bound_wide(const std::vector<Node8>& bvh, const std::vector<InstRec>& insts, uint32_t r,
                                 std::vector<uint32_t>& memo) {
  if (r == CG_REF_NONE || (r & CG_REF_LEAF) != 0u) return 0u;
  if ((r & CG_REF_INST) != 0u) {
    const InstRec& in = insts[r & CG_REF_INDEX_MASK];
    if (in.root == CG_REF_NONE) return 0u;
    if ((in.root & (CG_REF_LEAF | CG_REF_INST)) == CG_REF_INST) return 1u;  // a BLAS still to be built: bound added later
    return 1u + stack_bound_wide(bvh, insts, in.root, memo);
  }
  if (memo[r] != ~0u) return memo[r];
  const Node8& n = bvh[r];
  uint32_t k = 0, deepest = 0;
  for (int i = 0; i < CG_WIDE; ++i)
    if (n.ref[i] != CG_REF_NONE) {
      ++k;
      deepest = std::max(deepest, stack_bound_wide(bvh, insts, n.ref[i], memo));
    }
  return memo[r] = (k > 0 ? k - 1u : 0u) + deepest;
}

template <class Collapse = CollapseWideSah>
inline void collapse_scene_bvh(Layout& L) {
  L.bvh.clear();
  L.bvh.reserve(L.nodes.size() / 4 + 1);
  {
    Collapse c(L.nodes, L.bvh);
    for (InstRec& in : L.insts) in.root = c.convert(in.root);
    L.root = c.convert(L.root);
  }
  std::vector<uint32_t> memo(L.bvh.size(), ~0u);
  L.stack_depth = stack_bound_wide(L.bvh, L.insts, L.root, memo);
  L.nodes.clear();
  L.nodes.shrink_to_fit();
}

inline void quantize_scene_bvh(Layout& L) {
  L.bvhq.resize(L.bvh.size());
  const size_t n = L.bvh.size();
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t chunk = (n + hw - 1) / hw;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < hw; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&L, b, e] {
      for (size_t i = b; i < e; ++i) L.bvhq[i] = quantize_node(L.bvh[i]);
    });
  }
  for (auto& th : pool) th.join();
}

// The Morton grid of a BLAS: box centres -> [0, 1024) per axis.
inline void lbvh_grid(const BlasJob& j, float* cmin, float* cscale) {
  for (int a = 0; a < 3; ++a) {
    float lo = INFINITY, hi = -INFINITY;
    for (const PrimBox& b : j.boxes) {
      const float c = 0.5f * (b.lo[a] + b.hi[a]);
      lo = std::min(lo, c);
      hi = std::max(hi, c);
    }
    cmin[a] = lo;
    cscale[a] = hi > lo ? 1024.0f / (hi - lo) : 0.0f;
  }
}

// The linear-BVH build of one BLAS on the host, running the device build's
// per-thread steps in loops.  Appends its wide nodes (global indices from
// node_base) to `nodes` (and the exact ones to `exact`), and returns the
// sorted order of its triangles.
inline std::vector<uint32_t> lbvh_build_host(BlasJob& j, uint32_t node_base, std::vector<Node8Q>& nodes,
                                             std::vector<Node8>& exact) {
  const uint32_t n = static_cast<uint32_t>(j.boxes.size());
  std::vector<uint32_t> keys(n), ids(n), left(n - 1), right(n - 1), first(n - 1), count(n - 1),
      parent(2 * n - 1, kLbvhNone), flag(n - 1, 0u);
  std::vector<PrimBox> box(n - 1);
  std::vector<Price> price(n - 1);
  std::vector<Node8Q> wq(n - 1);
  std::vector<Node8> wx(n - 1);
  std::vector<uint2> front{make_uint2(0u, 0u)}, next;
  uint32_t next_n = 0, wide_n = 1, maxc = 0;
  LbvhArrays A{};
  A.n = n;
  A.prims = j.boxes.data();
  A.pad = j.pad;
  lbvh_grid(j, A.cmin, A.cscale);
  A.keys = keys.data();
  A.ids = ids.data();
  A.left = left.data();
  A.right = right.data();
  A.first = first.data();
  A.count = count.data();
  A.parent = parent.data();
  A.box = box.data();
  A.flag = flag.data();
  A.price = price.data();
  A.front_out_n = &next_n;
  A.wide_n = &wide_n;
  A.max_children = &maxc;
  A.nodes = wq.data();
  A.exact = wx.data();
  A.node_base = node_base;
  A.slot_base = j.slot_base;
  for (uint32_t i = 0; i < n; ++i) lbvh_morton(A, i);
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) { return keys[x] < keys[y]; });
  std::vector<uint32_t> sk(n);
  for (uint32_t i = 0; i < n; ++i) sk[i] = keys[perm[i]];
  keys.swap(sk);
  ids = perm;
  A.keys = keys.data();
  A.ids = ids.data();
  for (uint32_t i = 0; i + 1 < n; ++i) lbvh_node(A, i);
  for (uint32_t i = 0; i < n; ++i) lbvh_bounds(A, i);
  uint32_t stack = 0;
  while (!front.empty()) {
    next.assign(front.size() * CG_WIDE, make_uint2(0u, 0u));
    next_n = 0;
    maxc = 0;
    A.front_in = front.data();
    A.front_out = next.data();
    for (const uint2& e : front) lbvh_collapse(A, e);
    stack += maxc - 1u;
    front.assign(next.begin(), next.begin() + next_n);
  }
  nodes.insert(nodes.end(), wq.begin(), wq.begin() + wide_n);
  exact.insert(exact.end(), wx.begin(), wx.begin() + wide_n);
  j.root = node_base;
  j.nodes = wide_n;
  j.stack = stack;
  return ids;
}

// Places built BLAS: instance roots, and the stack bound (TLAS part plus the
// deepest BLAS).
inline void lbvh_finish(Layout& L, const std::vector<BlasJob>& jobs) {
  uint32_t deepest = 0;
  for (const BlasJob& j : jobs) deepest = std::max(deepest, j.stack);
  for (InstRec& in : L.insts)
    if ((in.root & (CG_REF_LEAF | CG_REF_INST)) == CG_REF_INST) in.root = jobs[in.root & CG_REF_INDEX_MASK].root;
  L.stack_depth += deepest;
}

// The whole layout with every BLAS built by the linear-BVH steps on the host
// (the device build's check; cg_scene_upload runs the same steps as kernels).
inline int build_layout_lbvh_host(const cg_scene_data* sd, Layout& L, std::string& err) {
  int rc = build_common(sd, L, err);
  if (rc != CG_OK) return rc;
  std::vector<BlasJob> jobs;
  rc = build_sah_accel(sd, L, err, &jobs);
  if (rc != CG_OK) return rc;
  collapse_scene_bvh(L);
  quantize_scene_bvh(L);
  for (BlasJob& j : jobs) {
    const std::vector<uint32_t> order = lbvh_build_host(j, static_cast<uint32_t>(L.bvhq.size()), L.bvhq, L.bvh);
    std::vector<TriRec> t(order.size());
    std::vector<uint4> v(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
      t[k] = L.tris[j.slot_base + order[k]];
      v[k] = L.vidx[j.slot_base + order[k]];
    }
    std::copy(t.begin(), t.end(), L.tris.begin() + j.slot_base);
    std::copy(v.begin(), v.end(), L.vidx.begin() + j.slot_base);
  }
  lbvh_finish(L, jobs);
  build_emitter_bvh(L);
  return CG_OK;
}

// accel 0: the reference's trees; accel 1: rebuilt (SAH, merged world BLAS).
// Either is traversed as an 8-wide BVH of quantised nodes.
inline int build_layout(const cg_scene_data* sd, Layout& L, std::string& err, int accel = 1) {
  if (accel == 2) return build_layout_lbvh_host(sd, L, err);
  int rc = build_common(sd, L, err);
  if (rc != CG_OK) return rc;
  rc = accel == 0 ? build_reference_accel(sd, L, err) : build_sah_accel(sd, L, err);
  if (rc != CG_OK) return rc;
  collapse_scene_bvh(L);
  quantize_scene_bvh(L);
  build_emitter_bvh(L);
  return CG_OK;
}

}  // namespace cg

#endif  // CG_UPLOAD_H_
