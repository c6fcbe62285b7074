// cg_upload.h — host-side construction of the device layout (cg_layout.h)
// from the reference's flattened SceneData (cg_scene.h), and the wavelength
// sampling table.
#ifndef CG_UPLOAD_H_
#define CG_UPLOAD_H_

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/cg_device.h"
#include "cg_core.h"
#include "cg_layout.h"

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
  std::vector<PairNode> nodes;
  std::vector<TriRec> tris;
  std::vector<InstRec> insts;
  std::vector<uint4> vidx;
  std::vector<VAttr> vattr;
  std::vector<MatRec> mats;
  std::vector<EmitRec> emit;
  std::vector<float2> alias;
  uint32_t root = CG_REF_NONE;
  uint32_t stack_depth = 0;
};

inline int build_layout(const cg_scene_data* sd, Layout& L, std::string& err) {
  // Materials: textures are not implemented by the kernels.
  for (uint32_t i = 0; i < sd->material_count; ++i) {
    const cg_material& m = sd->materials[i];
    for (int t : {CG_TEX_BASE_COLOR, CG_TEX_METALLIC_ROUGHNESS, CG_TEX_EMISSIVE, CG_TEX_TRANSMISSION})
      if (m.tex[t].index >= 0)
        return set_error(err, CG_ERR_UNSUPPORTED, "material %u uses texture slot %d: material textures are not supported",
                    i, t);
    MatRec r{};
    for (int k = 0; k < 3; ++k) r.base[k] = m.base_color_factor[k];
    r.metallic = fminf(fmaxf(m.metallic_factor, 0.0f), 1.0f);
    r.roughness = fminf(fmaxf(m.roughness_factor, 0.0f), 1.0f);
    r.transmission = fminf(fmaxf(m.transmission_factor, 0.0f), 1.0f);
    r.ior = m.ior;
    r.dispersion = m.dispersion;
    for (int k = 0; k < 3; ++k) r.emission[k] = fmaxf(m.emissive_factor[k] * m.emissive_strength, 0.0f);
    r.volume = m.thickness_factor > 0.0f ? 1u : 0u;
    r.alpha = r.roughness * r.roughness;
    L.mats.push_back(r);
  }
  if (L.mats.empty()) return set_error(err, CG_ERR_INVALID, "scene has no materials");
  if (sd->env_width == 0 || sd->env_height == 0 || !sd->env_rgba32f)
    return set_error(err, CG_ERR_INVALID, "scene has no environment texture");
  if (sd->alias_count != sd->emissive_count)
    return set_error(err, CG_ERR_INVALID, "alias table (%u) and emissive buffer (%u) differ in length", sd->alias_count,
                sd->emissive_count);
  if (sd->index_count >= (1u << 30) - 1u) return set_error(err, CG_ERR_INVALID, "too many triangles");

  // Vertices: attributes read for the final hit only.
  L.vattr.resize(sd->vertex_count);
  for (uint32_t i = 0; i < sd->vertex_count; ++i) {
    const cg_vertex& v = sd->vertices[i];
    VAttr& a = L.vattr[i];
    for (int k = 0; k < 3; ++k) a.nrm[k] = v.nrm[k];
    a.uv0[0] = v.uv0[0];
    a.uv0[1] = v.uv0[1];
    a.uv1[0] = v.uv1[0];
    a.uv1[1] = v.uv1[1];
    a._pad = 0.0f;
  }
  // Triangles in BVH order.
  L.tris.resize(sd->index_count);
  L.vidx.resize(sd->index_count);
  for (uint32_t s = 0; s < sd->index_count; ++s) {
    const cg_index& ix = sd->indices[s];
    if (ix.i0 >= sd->vertex_count || ix.i1 >= sd->vertex_count || ix.i2 >= sd->vertex_count)
      return set_error(err, CG_ERR_INVALID, "triangle slot %u indexes past the vertex buffer", s);
    const float* p0 = sd->vertices[ix.i0].pos;
    const float* p1 = sd->vertices[ix.i1].pos;
    const float* p2 = sd->vertices[ix.i2].pos;
    TriRec& t = L.tris[s];
    for (int k = 0; k < 3; ++k) {
      t.v0[k] = p0[k];
      t.e1[k] = p1[k] - p0[k];
      t.e2[k] = p2[k] - p0[k];
    }
    t.tri_id = sd->tri_ids ? sd->tri_ids[s] : s;
    t._pad[0] = t._pad[1] = 0u;
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
    blas_ref[i] = CG_REF_LEAF | (n.triangle_count == 2u ? CG_REF_LEAF2 : 0u) | n.child;
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


}  // namespace cg

#endif  // CG_UPLOAD_H_
