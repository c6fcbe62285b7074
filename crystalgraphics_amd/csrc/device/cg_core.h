// cg_core.h — device functions of the spectral path tracer.
//
// Float discipline: every expression is evaluated in the order the
// reference's WGSL writes it, one IEEE binary32 operation at a time
// (the library is compiled with -ffp-contract=off), min/max are the
// NaN-ignoring fminf/fmaxf, f32->u32 conversions saturate, and runtime array
// indices are clamped the way WebGPU robustness clamps them.  That makes
// traversal, intersection, RNG, wavelength sampling, alias sampling, Smits
// lifts, Fresnel, GGX terms and lobe sampling bit-identical to a CPU
// restatement; only acos/atan2 (the sky texel, a discrete choice) may differ
// in the last ulp.  The sRGB target's power is taken in double precision and
// rounded once to binary32, as the host does it.
#ifndef CG_CORE_H_
#define CG_CORE_H_

#include <stdint.h>

#include "../../../include/cg_scene.h"
#include "cg_layout.h"

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define CG_HD __host__ __device__ __forceinline__
#define CG_SQRT sqrtf
#define CG_FABS fabsf
#else
#include <cmath>
#define CG_HD inline
#define CG_SQRT std::sqrt
#define CG_FABS std::fabs
#endif

#pragma clang fp contract(off)

namespace cg {

// ---------------------------------------------------------------------------
// Vectors (evaluation order: left to right).
struct f3 {
  float x, y, z;
};
struct f2 {
  float x, y;
};
CG_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
CG_HD f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
CG_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CG_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CG_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
CG_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
CG_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
CG_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
CG_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
CG_HD f2 operator+(f2 a, f2 b) { return f2{a.x + b.x, a.y + b.y}; }
CG_HD f2 operator*(float s, f2 a) { return f2{s * a.x, s * a.y}; }
CG_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CG_HD f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
CG_HD float length(f3 a) { return CG_SQRT(dot(a, a)); }
CG_HD f3 normalize(f3 a) {
  float l = length(a);
  return f3{a.x / l, a.y / l, a.z / l};
}
CG_HD float fmn(float a, float b) { return fminf(a, b); }
CG_HD float fmx(float a, float b) { return fmaxf(a, b); }
CG_HD float clampf(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }
CG_HD f3 vmin(f3 a, f3 b) { return f3{fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
CG_HD f3 vmax(f3 a, f3 b) { return f3{fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }
CG_HD uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(x);
}
CG_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
CG_HD uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
CG_HD f3 reflect(f3 e1, f3 e2) {
  float s = 2.0f * dot(e2, e1);
  return e1 - e2 * s;
}
CG_HD f3 refract(f3 e1, f3 e2, float e3) {
  float d = dot(e2, e1);
  float k = 1.0f - e3 * e3 * (1.0f - d * d);
  if (k < 0.0f) return mk3(0.0f, 0.0f, 0.0f);
  return e1 * e3 - e2 * (e3 * d + CG_SQRT(k));
}
CG_HD float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// Column-major 4x4: (M * vec4(p, 1)).xyz, (M * vec4(d, 0)).xyz and
// (transpose(M) * vec4(n, 0)).xyz, summed column by column.
CG_HD f3 mat_point(const float* m, f3 p) {
  return ((ld3(m) * p.x + ld3(m + 4) * p.y) + ld3(m + 8) * p.z) + ld3(m + 12) * 1.0f;
}
CG_HD f3 mat_dir(const float* m, f3 d) {
  return ((ld3(m) * d.x + ld3(m + 4) * d.y) + ld3(m + 8) * d.z) + ld3(m + 12) * 0.0f;
}
CG_HD f3 mat_t_dir(const float* m, f3 n) {
  float r0 = ((m[0] * n.x + m[1] * n.y) + m[2] * n.z) + m[3] * 0.0f;
  float r1 = ((m[4] * n.x + m[5] * n.y) + m[6] * n.z) + m[7] * 0.0f;
  float r2 = ((m[8] * n.x + m[9] * n.y) + m[10] * n.z) + m[11] * 0.0f;
  return mk3(r0, r1, r2);
}

// ---------------------------------------------------------------------------
// Per-pixel RNG (shader/compute/math/rng.wesl:1-21).
CG_HD void rng_step(uint32_t& s) {
  s = s * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  s = (w >> 22u) ^ w;
}
CG_HD float rng_float(uint32_t& s) {
  rng_step(s);
  return static_cast<float>(s) / 4294967296.0f;
}
// Sample-stream mode (declared for this build; the reference has one stream
// per pixel): sample i of a pixel draws from the stream seeded
// step(step(b) ^ (i + 1) * 0x9E3779B9) from the pixel's base seed b, and the
// pixel carries step(step(b)) to the next frame (the "i = -1" stream; step is
// a bijection, so it differs from every sample's).
CG_HD uint32_t sample_seed(uint32_t base, uint32_t i) {
  uint32_t t = base;
  rng_step(t);
  t ^= (i + 1u) * 0x9E3779B9u;
  rng_step(t);
  return t;
}
CG_HD uint32_t sample_carry(uint32_t base) {
  uint32_t t = base;
  rng_step(t);
  rng_step(t);
  return t;
}

constexpr float kPi = 3.1415926535f;
constexpr float kMax = 1e6f;

// ---------------------------------------------------------------------------
// Ray/box (math/geometric.wesl:5-26) and ray/triangle
// (structural/bvh.wesl:188-221).
CG_HD f3 dir_inv(f3 d) {
  float x = d.x == 0.0f ? 1e-7f : d.x;
  float y = d.y == 0.0f ? 1e-7f : d.y;
  float z = d.z == 0.0f ? 1e-7f : d.z;
  return mk3(1.0f / x, 1.0f / y, 1.0f / z);
}
// Slab test; returns t_near, or +inf when the reference's test rejects.
CG_HD float box_enter(f3 o, f3 inv, const float* lo, const float* hi) {
  float t0x = (lo[0] - o.x) * inv.x, t1x = (hi[0] - o.x) * inv.x;
  float t0y = (lo[1] - o.y) * inv.y, t1y = (hi[1] - o.y) * inv.y;
  float t0z = (lo[2] - o.z) * inv.z, t1z = (hi[2] - o.z) * inv.z;
  float tn = fmx(fmx(fmn(t0x, t1x), fmn(t0y, t1y)), fmn(t0z, t1z));
  float tf = fmn(fmn(fmx(t0x, t1x), fmx(t0y, t1y)), fmx(t0z, t1z));
  return (tf >= tn && tf > 0.0f) ? tn : __builtin_inff();
}
struct TriT {
  float t, u, v;
};
// Moller-Trumbore with the edges precomputed (e1 = v1 - v0, e2 = v2 - v0 are
// the same subtractions the reference performs per test).
CG_HD TriT tri_test(f3 o, f3 d, f3 v0, f3 e1, f3 e2) {
  TriT r{-1.0f, 0.0f, 0.0f};
  f3 h = cross(d, e2);
  float a = dot(e1, h);
  if (a > -1e-6f && a < 1e-6f) return r;
  float f = 1.0f / a;
  f3 s = o - v0;
  float u = f * dot(s, h);
  if (u < 0.0f || u > 1.0f) return r;
  f3 q = cross(s, e1);
  float v = f * dot(d, q);
  if (v < 0.0f || u + v > 1.0f) return r;
  float t = f * dot(e2, q);
  if (t > 1e-6f) {
    r.t = t;
    r.u = u;
    r.v = v;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Scene view used by the kernels.
struct DevScene {
  const Node8Q* bvh;       // scene BVH, 8-wide, quantised boxes
  const PairNode* nodes;   // emitter BVH
  const TriRec* tris;
  const InstRec* insts;
  const uint4* tri_vidx;   // per slot: source vertex indices
  const VAttr* vattr;
  const MatRec* mats;
  uint32_t mat_count;
  const EmitRec* emit;
  uint32_t emit_count;
  const float2* alias;     // {bitcast alias index, prob}
  uint32_t alias_count;
  const uint32_t* emit_order;  // emitter BVH leaf slot -> emitter index
  uint32_t emit_root;          // emitter BVH root ref (CG_REF_NONE: none)
  const float4* env;
  uint32_t env_w, env_h;
  uint32_t root;           // TLAS root ref
  const uint32_t* tex;     // material texture array, RGBA8 texels (r in the low byte)
  uint32_t tex_w, tex_h;
};

struct TravStats {
  uint32_t nodes, tris, insts;
};

struct HitRec {
  float t, u, v;
  uint32_t inst, slot;
};

// Traversal stack: the first `short_n` entries of each lane live in `lds`
// (`stride`-interleaved, one 8-byte entry per access), deeper entries in
// `ovf` (device memory, `ovf_stride`-interleaved; sized from the tree's stack
// bound).  An entry is a node ref and the bits of the distance at which the
// ray enters that node.
constexpr int kShortStack = 16;  // default LDS entries per lane
struct TStack {
  uint2* lds;
  int stride;
  int short_n;
  uint2* ovf;
  uint32_t ovf_stride;
  int ovf_cap;  // overflow entries per lane (the bound says none are needed beyond)
  uint32_t* err;  // set to 1 if an entry is dropped past ovf_cap (nullable)
  CG_HD void push(int& sp, uint32_t r, float tt) const {
    if (sp < short_n) {
      lds[sp * stride] = make_uint2(r, __float_as_uint(tt));
    } else if (sp - short_n < ovf_cap) {
      ovf[static_cast<size_t>(sp - short_n) * ovf_stride] = make_uint2(r, __float_as_uint(tt));
    } else if (err) {
      *err = 1u;  // the stack bound was exceeded: the walk is incomplete
    }
    ++sp;
  }
  CG_HD void pop(int& sp, uint32_t& r, float& tt) const {
    --sp;
    if (sp < short_n) {
      const uint2 e = lds[sp * stride];
      r = e.x;
      tt = __uint_as_float(e.y);
    } else if (sp - short_n < ovf_cap) {
      const uint2 e = ovf[static_cast<size_t>(sp - short_n) * ovf_stride];
      r = e.x;
      tt = __uint_as_float(e.y);
    } else {
      r = CG_REF_NONE;
      tt = __builtin_inff();
    }
  }
};

CG_HD void cswap(float& ta, uint32_t& ra, float& tb, uint32_t& rb) {
  if (tb < ta) {
    float t = ta;
    ta = tb;
    tb = t;
    uint32_t r = ra;
    ra = rb;
    rb = r;
  }
}

// Two-level closest-hit (or any-hit) traversal of the 8-wide BVH: the hit
// children of a node sorted by entry distance, the nearest visited next and
// the others pushed with their distances; popped entries entered beyond the
// best hit so far are dropped; a marker on the stack returns from a BLAS to
// the TLAS.  The closest hit equals the reference's unordered walk
// (bvh.wesl:56-186) except between triangles hit at exactly the same
// distance.
// The scene pointers a traversal reads (per lane: lanes of one wave may
// trace rays of different scenes).
struct TravScene {
  const Node8Q* bvh;
  const TriRec* tris;
  const InstRec* insts;
  uint32_t root;
};
CG_HD TravScene trav_scene(const DevScene& S) { return TravScene{S.bvh, S.tris, S.insts, S.root}; }

// A traversal in progress, resumable one visit at a time.
struct TravRay {
  f3 wo, wd, winv;  // world-space ray
  f3 o, d, inv;     // ray in the space of the BLAS being walked
  HitRec best;
  uint32_t ref;     // next node, instance or leaf to visit
  uint32_t cur_inst;
  int sp;
  bool any;         // any-hit query: finished at the first hit below tmax
};

// Start a traversal; false when the scene is empty (the ray misses).
CG_HD bool trav_begin(const TravScene& S, TravRay& r, f3 wo, f3 wd, float tmax, bool any) {
  r.best = HitRec{tmax, 0.0f, 0.0f, CG_MISS_ID, CG_MISS_ID};
  r.ref = S.root;
  r.wo = wo;
  r.wd = wd;
  r.winv = dir_inv(wd);
  r.o = wo;
  r.d = wd;
  r.inv = r.winv;
  r.cur_inst = CG_MISS_ID;
  r.sp = 0;
  r.any = any;
  return r.ref != CG_REF_NONE;
}

// Two-level closest-hit (or any-hit) traversal of the 8-wide BVH, one visit
// per call: at a node the hit children are sorted by entry distance, the
// nearest is visited next and the others are pushed with their distances;
// popped entries entered beyond the best hit so far are dropped; a marker on
// the stack returns from a BLAS to the TLAS.  Returns true when finished
// (r.best holds the result).  The closest hit equals the reference's
// unordered walk (bvh.wesl:56-186) except between triangles hit at exactly
// the same distance.  kMode 0: closest hit; 1: any hit; 2: any hit where
// r.any is set (one launch serves a queue holding rays of both kinds).
template <int kMode>
CG_HD bool trav_step(const TravScene& S, TravRay& r, const TStack& stk, TravStats& st) {
  const float kInf = __builtin_inff();
  const uint32_t ref = r.ref;
  const bool at_node = (ref & (CG_REF_LEAF | CG_REF_INST)) == 0u;
  const bool at_leaf = (ref & CG_REF_LEAF) != 0u;
  const uint32_t slot = ref & CG_REF_SLOT_MASK;
  const uint32_t cnt = ((ref >> CG_REF_COUNT_SHIFT) & 3u) + 1u;
  // One load phase whatever the visit is, so a wave whose lanes sit at nodes,
  // instances and leaves waits for memory once per step: a node's first 96
  // bytes, an instance record (80 bytes), or the leaf's first two triangle
  // records (48 bytes each).  Words past a record's end re-read its start.
  const uint4* p = at_node ? reinterpret_cast<const uint4*>(S.bvh + ref)
                   : at_leaf ? reinterpret_cast<const uint4*>(S.tris + slot)
                             : reinterpret_cast<const uint4*>(S.insts + (ref & CG_REF_INDEX_MASK));
  const bool wide = at_node || (at_leaf && cnt >= 2u);  // six words
  const uint4* p3 = (wide || !at_leaf) ? p + 3 : p;
  const uint4* p5 = wide ? p + 5 : p;
  uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p3[0], w4 = p3[1], w5 = p5[0];
  // All six words arrive as six 16-byte loads issued together: left to
  // itself the compiler narrows some to single words and re-reads them inside
  // the branch that uses them, a second memory round trip per visit.
  asm volatile("" : "+v"(w0.x), "+v"(w0.y), "+v"(w0.z), "+v"(w0.w), "+v"(w1.x), "+v"(w1.y), "+v"(w1.z), "+v"(w1.w),
               "+v"(w2.x), "+v"(w2.y), "+v"(w2.z), "+v"(w2.w), "+v"(w3.x), "+v"(w3.y), "+v"(w3.z), "+v"(w3.w),
               "+v"(w4.x), "+v"(w4.y), "+v"(w4.z), "+v"(w4.w), "+v"(w5.x), "+v"(w5.y), "+v"(w5.z), "+v"(w5.w));
  if (at_node) {
    // 128-B quantised node: origin + exponents, 8-bit child bounds
    // (lo x/y/z, hi x/y/z; 8 bytes each, byte i = child i), refs.
    st.nodes++;
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const float sx = __uint_as_float((w0.w & 0xFFu) << 23), sy = __uint_as_float(((w0.w >> 8) & 0xFFu) << 23),
                sz = __uint_as_float(((w0.w >> 16) & 0xFFu) << 23);
    float t[CG_WIDE];
    uint32_t c[CG_WIDE] = {w4.x, w4.y, w4.z, w4.w, w5.x, w5.y, w5.z, w5.w};
    // The slab distance of a grid line q is (origin + q * step - o) * inv =
    // q * (step * inv) + (origin - o) * inv: one multiply-add per bound.  Its
    // rounding differs from the unfolded form by a few ulps, far inside the
    // boxes' outward padding, so no box holding a hit is culled and the
    // closest hits are unchanged.
    const float ax = sx * r.inv.x, ay = sy * r.inv.y, az = sz * r.inv.z;
    const float bx = (ox - r.o.x) * r.inv.x, by = (oy - r.o.y) * r.inv.y, bz = (oz - r.o.z) * r.inv.z;
    // q * a + b is monotone in q, rising where the ray's slope is positive:
    // there the low bound is entered and the high bound left, elsewhere the
    // reverse.  Choosing the bytes by the slope's sign once per node gives
    // the same entry and exit distances as a min and a max per child.
    const bool fx = ax >= 0.0f, fy = ay >= 0.0f, fz = az >= 0.0f;
    const uint32_t nx0 = fx ? w1.x : w2.z, nx1 = fx ? w1.y : w2.w, mx0 = fx ? w2.z : w1.x, mx1 = fx ? w2.w : w1.y;
    const uint32_t ny0 = fy ? w1.z : w3.x, ny1 = fy ? w1.w : w3.y, my0 = fy ? w3.x : w1.z, my1 = fy ? w3.y : w1.w;
    const uint32_t nz0 = fz ? w2.x : w3.z, nz1 = fz ? w2.y : w3.w, mz0 = fz ? w3.z : w2.x, mz1 = fz ? w3.w : w2.y;
#define CG_QB(w, i) static_cast<float>(((w) >> (8 * ((i) & 3))) & 0xFFu)
#pragma unroll
    for (int i = 0; i < CG_WIDE; ++i) {
      const float tn = fmx(fmx(fmaf(CG_QB(i < 4 ? nx0 : nx1, i), ax, bx), fmaf(CG_QB(i < 4 ? ny0 : ny1, i), ay, by)),
                           fmaf(CG_QB(i < 4 ? nz0 : nz1, i), az, bz));
      const float tf = fmn(fmn(fmaf(CG_QB(i < 4 ? mx0 : mx1, i), ax, bx), fmaf(CG_QB(i < 4 ? my0 : my1, i), ay, by)),
                           fmaf(CG_QB(i < 4 ? mz0 : mz1, i), az, bz));
      t[i] = (tf >= tn && tf > 0.0f && tn <= r.best.t && c[i] != CG_REF_NONE) ? tn : kInf;
    }
#undef CG_QB
    // nearest first: an 8-input sorting network (19 compare-exchanges)
    cswap(t[0], c[0], t[2], c[2]);
    cswap(t[1], c[1], t[3], c[3]);
    cswap(t[4], c[4], t[6], c[6]);
    cswap(t[5], c[5], t[7], c[7]);
    cswap(t[0], c[0], t[4], c[4]);
    cswap(t[1], c[1], t[5], c[5]);
    cswap(t[2], c[2], t[6], c[6]);
    cswap(t[3], c[3], t[7], c[7]);
    cswap(t[0], c[0], t[1], c[1]);
    cswap(t[2], c[2], t[3], c[3]);
    cswap(t[4], c[4], t[5], c[5]);
    cswap(t[6], c[6], t[7], c[7]);
    cswap(t[2], c[2], t[4], c[4]);
    cswap(t[3], c[3], t[5], c[5]);
    cswap(t[1], c[1], t[4], c[4]);
    cswap(t[3], c[3], t[6], c[6]);
    cswap(t[1], c[1], t[2], c[2]);
    cswap(t[3], c[3], t[4], c[4]);
    cswap(t[5], c[5], t[6], c[6]);
#pragma unroll
    for (int i = CG_WIDE - 1; i >= 1; --i)
      if (t[i] < kInf) stk.push(r.sp, c[i], t[i]);
    if (t[0] < kInf) {
      r.ref = c[0];
      return false;
    }
  } else if (!at_leaf) {
    // InstRec: the inverse transform's columns (words 0-3), then root.
    const uint32_t ii = ref & CG_REF_INDEX_MASK;
    st.insts++;
    if (w4.x != CG_REF_NONE) {
      const float inv[16] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z), __uint_as_float(w0.w),
                             __uint_as_float(w1.x), __uint_as_float(w1.y), __uint_as_float(w1.z), __uint_as_float(w1.w),
                             __uint_as_float(w2.x), __uint_as_float(w2.y), __uint_as_float(w2.z), __uint_as_float(w2.w),
                             __uint_as_float(w3.x), __uint_as_float(w3.y), __uint_as_float(w3.z), __uint_as_float(w3.w)};
      r.o = mat_point(inv, r.wo);
      r.d = mat_dir(inv, r.wd);
      r.inv = dir_inv(r.d);
      r.cur_inst = ii;
      stk.push(r.sp, CG_REF_MARKER, 0.0f);
      r.ref = w4.x;
      return false;
    }
  } else {
    // TriRec: v0 (3 words), e1 (3), e2 (3), tri_id, inst, pad; two per load.
    const uint4 a[2][3] = {{w0, w1, w2}, {w3, w4, w5}};
#pragma unroll
    for (uint32_t i = 0; i < 2u; ++i) {
      if (i >= cnt) break;
      const uint4 x = a[i][0], y = a[i][1], z = a[i][2];
      st.tris++;
      TriT h = tri_test(r.o, r.d, f3{__uint_as_float(x.x), __uint_as_float(x.y), __uint_as_float(x.z)},
                        f3{__uint_as_float(x.w), __uint_as_float(y.x), __uint_as_float(y.y)},
                        f3{__uint_as_float(y.z), __uint_as_float(y.w), __uint_as_float(z.x)});
      if (h.t > 0.0f && h.t < r.best.t) {
        r.best.t = h.t;
        r.best.u = h.u;
        r.best.v = h.v;
        r.best.inst = z.z != CG_MISS_ID ? z.z : r.cur_inst;
        r.best.slot = slot + i;
        if (kMode == 1 || (kMode == 2 && r.any)) return true;
      }
    }
    if (cnt > 2u) {  // the leaf's other triangles: the next step
      r.ref = CG_MAKE_LEAF(slot + 2u, cnt - 2u);
      return false;
    }
  }
  // pop
  for (;;) {
    if (r.sp == 0) return true;
    float te;
    uint32_t nr;
    stk.pop(r.sp, nr, te);
    if (nr == CG_REF_MARKER) {
      r.o = r.wo;
      r.d = r.wd;
      r.inv = r.winv;
      r.cur_inst = CG_MISS_ID;
      continue;
    }
    if (te <= r.best.t) {
      r.ref = nr;
      return false;
    }
  }
}

template <int kMode>
CG_HD HitRec traverse_m(const DevScene& S, f3 wo, f3 wd, float tmax, bool any, const TStack& stk, TravStats& st) {
  const TravScene ts = trav_scene(S);
  TravRay r;
  if (!trav_begin(ts, r, wo, wd, tmax, any)) return r.best;
  while (!trav_step<kMode>(ts, r, stk, st)) {
  }
  return r.best;
}

template <bool kAnyHit>
CG_HD HitRec traverse(const DevScene& S, f3 wo, f3 wd, float tmax, const TStack& stk, TravStats& st) {
  return traverse_m<kAnyHit ? 1 : 0>(S, wo, wd, tmax, kAnyHit, stk, st);
}

// Surface at a hit (bvh.wesl:103-129 and the per-hit attribute fetch of
// bvh.wesl:164-180), computed once for the final closest hit.
struct Surface {
  f3 pos, sn, gn;
  uint32_t material;
  f2 uv0, uv1;
};
CG_HD Surface surface_at(const DevScene& S, f3 wo, f3 wd, const HitRec& h) {
  const InstRec* in = S.insts + h.inst;
  const TriRec* tr = S.tris + h.slot;
  uint4 vi = S.tri_vidx[h.slot];
  const VAttr& a0 = S.vattr[vi.x];
  const VAttr& a1 = S.vattr[vi.y];
  const VAttr& a2 = S.vattr[vi.z];
  float w0 = 1.0f - h.u - h.v;
  f3 ln = normalize((w0 * ld3(a0.nrm) + h.u * ld3(a1.nrm)) + h.v * ld3(a2.nrm));
  f3 lgn = normalize(cross(ld3(tr->e1), ld3(tr->e2)));
  f2 u00{a0.uv0[0], a0.uv0[1]}, u01{a1.uv0[0], a1.uv0[1]}, u02{a2.uv0[0], a2.uv0[1]};
  f2 u10{a0.uv1[0], a0.uv1[1]}, u11{a1.uv1[0], a1.uv1[1]}, u12{a2.uv1[0], a2.uv1[1]};
  Surface s;
  f3 wgn = normalize(mat_t_dir(in->inv, lgn));
  f3 wsn = normalize(mat_t_dir(in->inv, ln));
  s.pos = wo + h.t * wd;
  s.sn = dot(wsn, wgn) >= 0.0f ? wsn : -wsn;
  s.gn = wgn;
  s.material = in->material;
  s.uv0 = (w0 * u00 + h.u * u01) + h.v * u02;
  s.uv1 = (w0 * u10 + h.u * u11) + h.v * u12;
  return s;
}

// ---------------------------------------------------------------------------
// Spectral pieces.  Smits (1999) basis as the reference tabulates it
// (shading/smits.wesl:3-124): bin = u32(10 * (l - 380) / 340) with the upper
// neighbour clamped to the last sample.
struct SmitsTables {
  float bins[10];
  float curve[7][10];
};
#if defined(__HIPCC__) || defined(__HIP__)
__constant__
#endif
static const SmitsTables kSmits = {
    {380.0f, 417.7778f, 455.5556f, 493.3333f, 531.1111f, 568.8889f, 606.6667f, 644.4444f, 682.2222f, 720.0f},
    {{1.0f, 1.0f, 0.9999f, 0.9993f, 0.9992f, 0.9998f, 1.0f, 1.0f, 1.0f, 1.0f},
     {0.1012f, 0.0515f, 0.0f, 0.0f, 0.0f, 0.0f, 0.8325f, 1.0149f, 1.0149f, 1.0149f},
     {0.0f, 0.0f, 0.0273f, 0.7937f, 1.0f, 0.9418f, 0.1719f, 0.0f, 0.0f, 0.0025f},
     {1.0f, 1.0f, 0.8916f, 0.3323f, 0.0f, 0.0f, 0.0003f, 0.0369f, 0.0483f, 0.0496f},
     {1.0f, 1.0f, 0.9685f, 0.2229f, 0.0f, 0.0458f, 0.8369f, 1.0f, 1.0f, 0.9959f},
     {0.971f, 0.9426f, 1.0007f, 1.0007f, 1.0007f, 1.0007f, 0.1564f, 0.0f, 0.0f, 0.0f},
     {0.0001f, 0.0f, 0.1088f, 0.6651f, 1.0f, 1.0f, 0.9996f, 0.9586f, 0.9685f, 0.984f}}};
enum { kWhite = 0, kRed, kGreen, kBlue, kMagenta, kCyan, kYellow };

// The seven curves share the wavelength's bin and mix weight.  A wavelength
// at or beyond an end of the table reads that end's sample with weight 0:
// (1 - 0) * c + 0 * c' = c exactly for the table's finite values.
struct SmitsBin {
  uint32_t b0;  // lower sample (0..9); the upper one is min(b0 + 1, 9)
  float mix;
};
CG_HD SmitsBin smits_bin(float l) {
  if (l <= 380.0f) return SmitsBin{0u, 0.0f};
  if (l >= 720.0f) return SmitsBin{9u, 0.0f};
  const uint32_t bin = f2u(10.0f * (l - 380.0f) / 340.0f);
  const uint32_t b0 = umin(bin, 9u), b1 = umin(bin + 1u, 9u);
  return SmitsBin{b0, (l - kSmits.bins[b0]) / (kSmits.bins[b1] - kSmits.bins[b0])};
}
// The curves, copied to LDS by every kernel that shades (smits_lds_init) so
// a lookup is one 32-bit LDS address per lane, not a pointer held per lane.
__shared__ float cg_smits_lds[7 * 10];
__device__ __forceinline__ void smits_lds_init() {
  for (uint32_t i = threadIdx.x; i < 70u; i += blockDim.x) cg_smits_lds[i] = kSmits.curve[i / 10u][i % 10u];
  __syncthreads();
}
// (Read through a volatile pointer: the reads stay where they are used
// instead of being hoisted into registers held across the sample.)
__device__ __forceinline__ float smits_at(int c, const SmitsBin& b) {
  const volatile float* t = cg_smits_lds + c * 10;
  const uint32_t b1 = umin(b.b0 + 1u, 9u);
  return (1.0f - b.mix) * t[b.b0] + b.mix * t[b1];
}
// material.wesl:264-297 / environment.wesl:27-61 / pdf/emission.wesl:84-117.
CG_HD float rgb_lift(f3 rgb, const SmitsBin& b) {
  float s = 0.0f;
  float r = rgb.x, g = rgb.y, bl = rgb.z;
  if (r <= g && g <= bl) {
    s += r * smits_at(kWhite, b);
    s += (g - r) * smits_at(kCyan, b);
    s += (bl - g) * smits_at(kBlue, b);
  } else if (g <= r && g <= bl) {
    s += g * smits_at(kWhite, b);
    if (r <= bl) {
      s += (r - g) * smits_at(kMagenta, b);
      s += (bl - r) * smits_at(kBlue, b);
    } else {
      s += (bl - g) * smits_at(kMagenta, b);
      s += (r - bl) * smits_at(kRed, b);
    }
  } else {
    s += bl * smits_at(kWhite, b);
    if (r <= g) {
      s += (r - bl) * smits_at(kYellow, b);
      s += (g - r) * smits_at(kGreen, b);
    } else {
      s += (g - bl) * smits_at(kYellow, b);
      s += (r - g) * smits_at(kRed, b);
    }
  }
  return s;
}
CG_HD float rgb_reflectance(f3 rgb, const SmitsBin& b) { return clampf(0.94f * rgb_lift(rgb, b), 0.0f, 1.0f); }
CG_HD float rgb_radiance(f3 rgb, const SmitsBin& b) { return fmx(rgb_lift(rgb, b), 0.0f); }

// ---------------------------------------------------------------------------
// Material at one wavelength (shading/material.wesl:19-110).
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
CG_HD float dispersed_ior(float nd, float disp, float l) {
  if (disp <= 0.0f) return nd;
  float inv_abbe = disp / 20.0f;
  return fmx(nd + (nd - 1.0f) * inv_abbe * (523655.0f / (l * l) - 1.5168f), 1.0f);
}

// Bilinear lookup in one layer of the material texture array at level 0 with
// repeat addressing: the sampler of src/wgpu/resource.cpp:589-606 as
// textureSampleLevel applies it (material.wesl:170-171), in declared fp32
// semantics.  The coordinate is wrapped into [0, 1] first (NaN and infinities
// read as 0), texel centres sit at (i + 0.5) / size, texels read as k / 255,
// and the four texels blend as two lerps across and one down, so a constant
// region reads back exactly.  Sizes are at most 16384, so x0, y0 stay in
// [-1, size - 1].
struct Texel4 {
  float r, g, b, a;
};
CG_HD float wrap01(float c) {
  const float f = c - floorf(c);
  return f >= 0.0f ? f : 0.0f;
}
CG_HD Texel4 unorm8x4(uint32_t t) {
  return Texel4{static_cast<float>(t & 255u) / 255.0f, static_cast<float>((t >> 8) & 255u) / 255.0f,
                static_cast<float>((t >> 16) & 255u) / 255.0f, static_cast<float>(t >> 24) / 255.0f};
}
CG_HD float lerp_decl(float p, float q, float w) { return p + w * (q - p); }
CG_HD Texel4 tex_sample(const uint32_t* tex, uint32_t w, uint32_t h, uint32_t layer, f2 uv) {
  const float x = wrap01(uv.x) * static_cast<float>(w) - 0.5f;
  const float y = wrap01(uv.y) * static_cast<float>(h) - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float a = x - fx, b = y - fy;
  const uint32_t x0 = fx < 0.0f ? w - 1u : umin(static_cast<uint32_t>(fx), w - 1u);
  const uint32_t y0 = fy < 0.0f ? h - 1u : umin(static_cast<uint32_t>(fy), h - 1u);
  const uint32_t x1 = x0 + 1u == w ? 0u : x0 + 1u;
  const uint32_t y1 = y0 + 1u == h ? 0u : y0 + 1u;
  const uint32_t* L = tex + static_cast<size_t>(layer) * w * h;
  const Texel4 t00 = unorm8x4(L[static_cast<size_t>(y0) * w + x0]), t10 = unorm8x4(L[static_cast<size_t>(y0) * w + x1]);
  const Texel4 t01 = unorm8x4(L[static_cast<size_t>(y1) * w + x0]), t11 = unorm8x4(L[static_cast<size_t>(y1) * w + x1]);
  auto ch = [&](float p00, float p10, float p01, float p11) {
    return lerp_decl(lerp_decl(p00, p10, a), lerp_decl(p01, p11, a), b);
  };
  return Texel4{ch(t00.r, t10.r, t01.r, t11.r), ch(t00.g, t10.g, t01.g, t11.g), ch(t00.b, t10.b, t01.b, t11.b),
                ch(t00.a, t10.a, t01.a, t11.a)};
}

// Material decode at one wavelength with its texture lookups
// (material.wesl:19-53 and the getters :181-262): factors times texture
// samples, then clamps.  The metallic-roughness texture is sampled once for
// both getters (they read the same texel).
CG_HD Mat decode(const DevScene& S, const MatRec& m, float l, const SmitsBin& b, f2 uv0, f2 uv1) {
  f3 base = ld3(m.base);
  float rough = m.roughness, metal = m.metallic, trans = m.transmission;
  f3 em = ld3(m.emission);
  const uint32_t lb = m.tex_base_mr & 0xFFFFu, lmr = m.tex_base_mr >> 16;
  const uint32_t lt = m.tex_trans_emit & 0xFFFFu, le = m.tex_trans_emit >> 16;
  if (lb != CG_TEX_NONE) {
    Texel4 t = tex_sample(S.tex, S.tex_w, S.tex_h, lb, (m.flags & CG_MAT_UV1_BASE) ? uv1 : uv0);
    base = f3{base.x * t.r, base.y * t.g, base.z * t.b};
  }
  if (lmr != CG_TEX_NONE) {
    Texel4 t = tex_sample(S.tex, S.tex_w, S.tex_h, lmr, (m.flags & CG_MAT_UV1_MR) ? uv1 : uv0);
    rough = rough * t.g;
    metal = metal * t.b;
  }
  if (lt != CG_TEX_NONE) {
    Texel4 t = tex_sample(S.tex, S.tex_w, S.tex_h, lt, (m.flags & CG_MAT_UV1_TRANS) ? uv1 : uv0);
    trans = trans * t.r;
  }
  if (le != CG_TEX_NONE) {
    Texel4 t = tex_sample(S.tex, S.tex_w, S.tex_h, le, (m.flags & CG_MAT_UV1_EMIT) ? uv1 : uv0);
    em = f3{em.x * t.r, em.y * t.g, em.z * t.b};
  }
  rough = clampf(rough, 0.0f, 1.0f);
  metal = clampf(metal, 0.0f, 1.0f);
  trans = clampf(trans, 0.0f, 1.0f);
  em = f3{fmx(em.x * m.strength, 0.0f), fmx(em.y * m.strength, 0.0f), fmx(em.z * m.strength, 0.0f)};
  float refl = rgb_reflectance(base, b);
  float real_ior = (metal == 0.0f && trans > 0.0f) ? dispersed_ior(m.ior, m.dispersion, l) : m.ior;
  float np1 = real_ior + 1.0f;
  float im = CG_SQRT(clampf((4.0f * real_ior) / (1.0f - refl) - np1 * np1, 0.0f, 10.0f));
  Mat r;
  r.diffuse = (1.0f - metal) * (1.0f - trans) * refl;
  r.alpha = rough * rough;
  r.ior = IOR{real_ior, metal * im};
  r.transmission = trans;
  r.volume = (m.flags & CG_MAT_VOLUME) != 0u;
  r.emission = rgb_reflectance(em, b);
  return r;
}

// shading/util.wesl:4-81
CG_HD f3 oriented(f3 d, f3 n) { return dot(d, n) > 0.0f ? -n : n; }
CG_HD IOR inv_ior(IOR i) {
  float nsq = i.re * i.re + i.im * i.im;
  return IOR{i.re / nsq, -i.im / nsq};
}
CG_HD IOR rel_ior(f3 sn, f3 d, const Mat& m) { return dot(sn, d) < 0.0f ? inv_ior(m.ior) : m.ior; }
CG_HD float fresnel(f3 d, f3 n, IOR rel) {
  float ci = clampf(dot(-d, n), 0.0f, 1.0f);
  if (CG_FABS(rel.im) > 1e-6f) {
    float re = rel.re, im = rel.im;
    float rsq = re * re, isq = im * im, csq = ci * ci;
    float nsq = rsq + isq;
    float trc = 2.0f * re * ci;
    float perp = (nsq - trc + csq) / fmx(nsq + trc + csq, 1e-6f);
    float par = (nsq * csq - trc + 1.0f) / fmx(nsq * csq + trc + 1.0f, 1e-6f);
    return clampf(0.5f * (perp + par), 0.0f, 1.0f);
  }
  float r = rel.re;
  float ts2 = r * r * fmx(0.0f, 1.0f - ci * ci);
  if (ts2 >= 1.0f) return 1.0f;
  float tc = CG_SQRT(fmx(0.0f, 1.0f - ts2));
  float perp = (r * ci - tc) / (r * ci + tc);
  float par = (ci - r * tc) / (ci + r * tc);
  return clampf(0.5f * (perp * perp + par * par), 0.0f, 1.0f);
}

// math/geometric.wesl:28-33
struct Frame {
  f3 t, b, n;
};
CG_HD Frame tangent_space(f3 n) {
  f3 helper = CG_FABS(n.x) > 0.9f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
  f3 t = normalize(cross(helper, n));
  return Frame{t, cross(n, t), n};
}
CG_HD f3 to_world(const Frame& f, f3 v) { return (f.t * v.x + f.b * v.y) + f.n * v.z; }
CG_HD f3 to_local(const Frame& f, f3 v) { return mk3(dot(f.t, v), dot(f.b, v), dot(f.n, v)); }

// sin and cos of a sampling angle in [0, 2 pi], as declared for this path:
// Cody-Waite reduction by pi/2 (three-part constant) and Taylor polynomials of
// degree 9 (sin) and 10 (cos) on [-pi/4, pi/4], evaluated in this order.
// Accuracy is about one ulp, well inside WGSL's 2^-11 bound for sin/cos.
CG_HD void sincos_decl(float x, float& s, float& c) {
  const float fk = floorf(x * 0x1.45f306p-1f + 0.5f);
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
// Lobes (shading/pdf/*.wesl, shading/dist/*.wesl).
struct TSample {
  f3 dir;
  float prob;
};
CG_HD float diffuse_pdf(f3 d, f3 next, const Surface& s) {
  f3 macro = oriented(d, s.sn);
  return fmx(dot(next, macro), 0.0f) / kPi;
}
CG_HD float diffuse_dist(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 macro = oriented(d, s.sn);
  float np = fmx(dot(next, macro), 0.0f);
  return m.diffuse * np / kPi;
}
CG_HD TSample diffuse_sample(f3 d, const Surface& s, uint32_t& rng) {
  f3 macro = oriented(d, s.sn);
  float r = CG_SQRT(rng_float(rng));
  float angle = 2.0f * kPi * rng_float(rng);
  float sa, ca;
  sincos_decl(angle, sa, ca);
  f3 local = mk3(r * ca, r * sa, CG_SQRT(fmx(0.0f, 1.0f - r * r)));
  f3 next = to_world(tangent_space(macro), local);
  return TSample{next, fmx(dot(next, macro), 0.0f) / kPi};
}
CG_HD f3 specular_dir(f3 d, const Surface& s) { return reflect(d, oriented(d, s.sn)); }
CG_HD float spec_reflect_pdf(f3 d, f3 next, const Surface& s) {
  f3 df = next - specular_dir(d, s);
  return dot(df, df) <= 1e-6f ? 1.0f : 0.0f;
}
CG_HD float spec_reflect_dist(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 macro = oriented(d, s.sn);
  f3 df = next - reflect(d, macro);
  float distribution = dot(df, df) <= 1e-6f ? 1.0f : 0.0f;
  return fresnel(d, macro, rel_ior(s.sn, d, m)) * distribution;
}
CG_HD f3 spec_trans_dir(f3 d, const Surface& s, const Mat& m) {
  if (m.volume) return refract(d, oriented(d, s.sn), rel_ior(s.sn, d, m).re);
  return d;
}
CG_HD float spec_trans_pdf(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 e = spec_trans_dir(d, s, m);
  f3 df = next - e;
  return (dot(e, e) > 1e-6f && dot(df, df) <= 1e-6f) ? 1.0f : 0.0f;
}
CG_HD float spec_trans_dist(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 macro = oriented(d, s.sn);
  f3 df = next - spec_trans_dir(d, s, m);
  float distribution = dot(df, df) <= 1e-6f ? 1.0f : 0.0f;
  return m.transmission * (1.0f - fresnel(d, macro, rel_ior(s.sn, d, m))) * distribution;
}
// GGX with the reference's alpha (not alpha^2) in D and Lambda
// (pdf/rough_reflection.wesl:120-139).
CG_HD float ggx_d(f3 micro, f3 macro, float alpha) {
  float c = clampf(dot(micro, macro), 0.0f, 1.0f);
  float c2 = c * c;
  float t2 = fmx(0.0f, 1.0f - c2) / fmx(c2, 1e-6f);
  float at = alpha + t2;
  float den = kPi * (c2 * c2) * (at * at);
  return alpha / fmx(den, 1e-6f);
}
// tan(acos(c))^2 of a projection c in [1e-6, 1], as its exact identity
// max(0, 1 - c^2) / c^2 (the composed form is ill-conditioned at grazing c).
CG_HD float tan2_acos(float c) { return fmx(0.0f, 1.0f - c * c) / (c * c); }
CG_HD float ggx_g1(float c, float alpha) {
  float t2 = tan2_acos(c);
  return 1.0f / (1.0f + (CG_SQRT(1.0f + alpha * t2) - 1.0f) / 2.0f);
}
CG_HD float ggx_g2(float c1, float c2, float alpha) {
  float t1 = tan2_acos(c1), t2 = tan2_acos(c2);
  return 1.0f / (1.0f + (CG_SQRT(1.0f + alpha * t1) - 1.0f) / 2.0f + (CG_SQRT(1.0f + alpha * t2) - 1.0f) / 2.0f);
}
CG_HD f3 sample_micro_normal(f3 d, const Surface& s, float alpha, uint32_t& rng) {
  f3 macro = oriented(d, s.sn);
  Frame f = tangent_space(macro);
  f3 local = normalize(to_local(f, -d));
  f3 st = normalize(mk3(alpha * local.x, alpha * local.y, fmx(local.z, 1e-6f)));
  float tl2 = st.x * st.x + st.y * st.y;
  f3 tx = tl2 > 1e-6f ? mk3(-st.y, st.x, 0.0f) / CG_SQRT(fmx(tl2, 1e-6f)) : mk3(1.0f, 0.0f, 0.0f);
  f3 ty = cross(st, tx);
  float radius = CG_SQRT(rng_float(rng));
  float angle = 2.0f * kPi * rng_float(rng);
  float sa, ca;
  sincos_decl(angle, sa, ca);
  float px = radius * ca;
  float py = radius * sa;
  float vis = 0.5f * (1.0f + st.z);
  py = mixf(CG_SQRT(fmx(0.0f, 1.0f - px * px)), py, vis);
  f3 lsm = px * tx + py * ty + CG_SQRT(fmx(0.0f, 1.0f - px * px - py * py)) * st;
  f3 lm = normalize(mk3(alpha * lsm.x, alpha * lsm.y, fmx(lsm.z, 1e-6f)));
  return normalize(to_world(f, lm));
}
CG_HD float reflect_pdf(f3 d, f3 micro, f3 macro, float alpha) {
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float cmi = fmx(dot(-d, micro), 1e-6f);
  float vnp = ggx_d(micro, macro, alpha) * ggx_g1(cmp, alpha) * cmi / cmp;
  return vnp / (4.0f * cmi);
}
CG_HD float rough_reflect_pdf(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 macro = oriented(d, s.sn);
  if (dot(next, macro) <= 0.0f) return 0.0f;
  f3 mu = next - d;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  return reflect_pdf(d, oriented(d, normalize(mu)), macro, m.alpha);
}
CG_HD TSample rough_reflect_sample(f3 d, const Surface& s, const Mat& m, uint32_t& rng) {
  f3 micro = oriented(d, sample_micro_normal(d, s, m.alpha, rng));
  f3 next = reflect(d, micro);
  f3 macro = oriented(d, s.sn);
  if (dot(next, macro) <= 0.0f) return TSample{next, 0.0f};
  return TSample{next, reflect_pdf(d, micro, macro, m.alpha)};
}
CG_HD float rough_reflect_dist(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 macro = oriented(d, s.sn);
  float np = fmx(dot(next, macro), 0.0f);
  if (np <= 0.0f) return 0.0f;
  f3 mu = next - d;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  f3 raw = normalize(mu);
  f3 micro = oriented(d, raw);
  IOR rel = rel_ior(raw, d, m);
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float nmp = fmx(dot(next, macro), 1e-6f);
  float rd = ggx_d(micro, macro, m.alpha) * ggx_g2(cmp, nmp, m.alpha) / (4.0f * cmp * nmp);
  return fresnel(d, micro, rel) * rd * np;
}
CG_HD float rel_value(IOR i) { return fmx(i.re, 1e-6f); }
CG_HD float transmit_pdf(f3 d, f3 micro, f3 macro, f3 next, float tior, float alpha) {
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float cmi = fmx(dot(-d, micro), 1e-6f);
  float vnp = ggx_d(micro, macro, alpha) * ggx_g1(cmp, alpha) * cmi / cmp;
  float nmi = dot(next, micro);
  float den = nmi + cmi / tior;
  return vnp * cmi / fmx(den * den, 1e-6f);
}
CG_HD float rough_trans_pdf(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 macro = oriented(d, s.sn);
  if (dot(next, macro) >= 0.0f) return 0.0f;
  float mrv = rel_value(rel_ior(macro, d, m));
  f3 mu = mrv * d - next;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  f3 micro = oriented(d, normalize(mu));
  float tior = 1.0f / rel_value(rel_ior(micro, d, m));
  return transmit_pdf(d, micro, macro, next, tior, m.alpha);
}
CG_HD TSample rough_trans_sample(f3 d, const Surface& s, const Mat& m, uint32_t& rng) {
  f3 macro = oriented(d, s.sn);
  f3 micro = oriented(d, sample_micro_normal(d, s, m.alpha, rng));
  f3 next = refract(d, micro, rel_value(rel_ior(micro, d, m)));
  if (dot(next, next) <= 1e-6f) return TSample{next, 0.0f};
  if (dot(next, macro) >= 0.0f) return TSample{next, 0.0f};
  return TSample{next, rough_trans_pdf(d, next, s, m)};
}
CG_HD float rough_trans_dist(f3 d, f3 next, const Surface& s, const Mat& m) {
  f3 macro = oriented(d, s.sn);
  if (dot(next, macro) >= 0.0f) return 0.0f;
  float np = fmx(dot(-next, macro), 0.0f);
  float mrv = rel_value(rel_ior(macro, d, m));
  f3 mu = mrv * d - next;
  if (dot(mu, mu) <= 1e-6f) return 0.0f;
  f3 raw = normalize(mu);
  f3 micro = oriented(d, raw);
  IOR rel = rel_ior(raw, d, m);
  float tior = 1.0f / rel_value(rel);
  float cmp = fmx(dot(-d, macro), 1e-6f);
  float nmp = fmx(dot(-next, macro), 1e-6f);
  float cmi = fmx(dot(-d, micro), 1e-6f);
  float nmi = dot(next, micro);
  float den = nmi + cmi / tior;
  float td = m.transmission * ggx_d(micro, macro, m.alpha) * ggx_g2(cmp, nmp, m.alpha) *
             CG_FABS(cmi * nmi) / fmx(cmp * nmp * den * den, 1e-6f);
  return (1.0f - fresnel(d, micro, rel)) * td * np;
}

// ---------------------------------------------------------------------------
// Emitters (shading/pdf/emission.wesl:28-191).
CG_HD float emit_hit_dist(f3 pos, f3 dir, const EmitRec& e) {
  f3 p0 = ld3(e.v0);
  f3 e1 = ld3(e.v1) - p0, e2 = ld3(e.v2) - p0;
  f3 pv = cross(dir, e2);
  float det = dot(e1, pv);
  if (CG_FABS(det) <= 1e-3f) return -1.0f;
  float inv = 1.0f / det;
  f3 tv = pos - p0;
  float u = dot(tv, pv) * inv;
  if (u < 0.0f || u > 1.0f) return -1.0f;
  f3 qv = cross(tv, e1);
  float v = dot(dir, qv) * inv;
  if (v < 0.0f || u + v > 1.0f) return -1.0f;
  float hd = dot(e2, qv) * inv;
  return hd > 1e-3f ? hd : -1.0f;
}
// Solid-angle pdf of the emitters the ray from `pos` along `next` passes
// through, summed in emitter order (emission.wesl:120-149).
CG_HD float light_pdf_linear(const DevScene& S, f3 rd, f3 pos) {
  float pdf = 0.0f;
  for (uint32_t i = 0; i < S.emit_count; ++i) {
    const EmitRec& e = S.emit[i];
    float hd = emit_hit_dist(pos, rd, e);
    if (hd <= 0.0f) continue;
    f3 ln = mk3(e.ln_x, e.ln_y, e.ln_z);
    float lp = CG_FABS(dot(ln, -rd));
    pdf += e.area_pdf * hd * hd / fmx(lp, 1e-3f);
  }
  return pdf;
}
// The emitter-BVH walk of light_pdf: among the emitters the ray passes
// through (padded boxes, no distance cut-off) with index above `after`, the
// one of smallest index (CG_MISS_ID: none) and its distance, and how many
// there are.
CG_HD uint32_t emit_next_hit(const DevScene& S, f3 rd, f3 inv, f3 pos, uint32_t after, float& hd_out, uint32_t& count,
                             const TStack& stk) {
  uint32_t best = CG_MISS_ID;
  count = 0u;
  uint32_t ref = S.emit_root;
  int sp = 0;
  for (;;) {
    if ((ref & CG_REF_LEAF) == 0u) {
      // the 64-byte node as four 16-byte loads issued together (see trav_step)
      const uint4* np = reinterpret_cast<const uint4*>(S.nodes + ref);
      uint4 a = np[0], b = np[1], c = np[2], d = np[3];
      asm volatile("" : "+v"(a.x), "+v"(a.y), "+v"(a.z), "+v"(a.w), "+v"(b.x), "+v"(b.y), "+v"(b.z), "+v"(b.w),
                   "+v"(c.x), "+v"(c.y), "+v"(c.z), "+v"(c.w), "+v"(d.x), "+v"(d.y));
      // PairNode: lo_l (a.xyz), hi_l (a.w, b.xy), lo_r (b.zw, c.x), hi_r (c.yzw), ref_l, ref_r (d.xy)
      const float lo_l[3] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)};
      const float hi_l[3] = {__uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y)};
      const float lo_r[3] = {__uint_as_float(b.z), __uint_as_float(b.w), __uint_as_float(c.x)};
      const float hi_r[3] = {__uint_as_float(c.y), __uint_as_float(c.z), __uint_as_float(c.w)};
      const bool hl = box_enter(pos, inv, lo_l, hi_l) < __builtin_inff();
      const bool hr = box_enter(pos, inv, lo_r, hi_r) < __builtin_inff();
      if (hl && hr) {
        stk.push(sp, d.y, 0.0f);
        ref = d.x;
        continue;
      }
      if (hl || hr) {
        ref = hl ? d.x : d.y;
        continue;
      }
    } else {
      const uint32_t slot = ref & CG_REF_SLOT_MASK;
      const uint32_t cnt = ((ref >> CG_REF_COUNT_SHIFT) & 3u) + 1u;
      for (uint32_t i = 0; i < cnt; ++i) {
        const uint32_t e = S.emit_order[slot + i];
        if (after != CG_MISS_ID && e <= after) continue;
        const float hd = emit_hit_dist(pos, rd, S.emit[e]);
        if (hd <= 0.0f) continue;
        ++count;
        if (e < best) {
          best = e;
          hd_out = hd;
        }
      }
    }
    if (sp == 0) break;
    float unused;
    stk.pop(sp, ref, unused);
  }
  return best;
}
// The same sum, with the emitters the ray passes through found by the BVH
// over the emitters and added in ascending emitter index, so the result
// equals the linear loop's bit for bit: one walk finds how many emitters the
// ray passes and the first of them; a ray through several walks again for
// each next one (rare: emitters are small), holding no list per lane.  A ray
// through more than kWalks emitters takes the linear loop (the same sum).
CG_HD float light_pdf(const DevScene& S, f3 next, f3 pos, const TStack& stk) {
  if (S.alias_count == 0u) return 0.0f;
  const f3 rd = normalize(next);
  if (S.emit_root == CG_REF_NONE) return light_pdf_linear(S, rd, pos);
  constexpr uint32_t kWalks = 16u;
  const f3 inv = dir_inv(rd);
  float pdf = 0.0f;
  uint32_t after = CG_MISS_ID, count = 0u;
  float hd = 0.0f;
  uint32_t e = emit_next_hit(S, rd, inv, pos, after, hd, count, stk);
  if (count > kWalks) return light_pdf_linear(S, rd, pos);
  for (uint32_t k = 0; k < count; ++k) {
    if (k > 0u) {
      uint32_t unused;
      e = emit_next_hit(S, rd, inv, pos, after, hd, unused, stk);
    }
    const EmitRec& er = S.emit[e];
    f3 ln = mk3(er.ln_x, er.ln_y, er.ln_z);
    float lp = CG_FABS(dot(ln, -rd));
    pdf += er.area_pdf * hd * hd / fmx(lp, 1e-3f);
    after = e;
  }
  return pdf;
}
struct ESample {
  f3 dir;
  float prob, radiance, dist;
  uint32_t prim;  // the emitter the alias table picked
};
// sampleEmissivePrim (pdf/emission.wesl:181-191) for two uniforms: bin =
// u32(u1 * n), which is n when u1 == 1.0; the table read and the emitter it
// names are clamped to the arrays as WebGPU robustness clamps them.
CG_HD uint32_t alias_pick(const DevScene& S, float u1, float u2) {
  const uint32_t n = S.alias_count;
  const uint32_t bin = f2u(u1 * static_cast<float>(n));
  const float2 ent = S.alias[umin(bin, n - 1u)];
  return umin(u2 < ent.y ? bin : __builtin_bit_cast(uint32_t, ent.x), S.emit_count - 1u);
}
CG_HD ESample emission_sample(const DevScene& S, const Surface& s, const SmitsBin& b, uint32_t& rng) {
  const float u1 = rng_float(rng);
  const float u2 = rng_float(rng);
  const uint32_t pick = alias_pick(S, u1, u2);
  const EmitRec& e = S.emit[pick];
  float tx = rng_float(rng);
  float ty = rng_float(rng);
  float sq = CG_SQRT(tx);
  float bx = 1.0f - sq, by = sq * (1.0f - ty), bz = sq * ty;
  f3 sp = (bx * ld3(e.v0) + by * ld3(e.v1)) + bz * ld3(e.v2);
  f3 off = sp - s.pos;
  float dsq = fmx(dot(off, off), 1e-3f);
  f3 next = off / CG_SQRT(dsq);
  f3 ln = mk3(e.ln_x, e.ln_y, e.ln_z);
  float lp = CG_FABS(dot(ln, -next));
  float sapdf = e.area_pdf * dsq / fmx(lp, 1e-3f);
  return ESample{next, sapdf, rgb_radiance(ld3(e.rgb), b), CG_SQRT(dsq), pick};
}

// ---------------------------------------------------------------------------
// Lobe counts and balance-heuristic MIS (shading/sample.wesl:110-501).
struct Counts6 {
  uint32_t d, sr, rr, st, rt, l;
};
struct TraceParams {
  int D;            // lod MAX_DEPTH == TRACE_DEPTH (clamped to [1, 64])
  uint32_t T, E, DIF, ROUGH;
};
CG_HD uint32_t scaled_count(float h, float scale, uint32_t mx) {
  if (h <= 1e-6f) return 0u;
  return umin(umax(1u, f2u(h * scale)), mx);
}
CG_HD Counts6 lobe_counts(const Mat& m, bool can_light, int depth, const TraceParams& P) {
  float lo = m.ior.re - 1.0f, hi = m.ior.re + 1.0f, isq = m.ior.im * m.ior.im;
  float refl = clampf((lo * lo + isq) / fmx(hi * hi + isq, 1e-6f), 0.0f, 1.0f);
  float hd = m.diffuse, trans = m.transmission;
  float hsr = m.alpha == 0.0f ? refl : 0.0f;
  float hrr = m.alpha > 0.0f ? refl : 0.0f;
  float hst = (trans > 1e-6f && m.alpha == 0.0f) ? trans : 0.0f;
  float hrt = (trans > 1e-6f && m.alpha > 0.0f) ? trans : 0.0f;
  float hl = can_light ? 1.0f : 0.0f;
  float heur = static_cast<float>(P.D - depth) / static_cast<float>(P.D);
  uint32_t t_avail = f2u(heur * static_cast<float>(umin(P.T, 4u)));
  uint32_t e_avail = f2u(heur * static_cast<float>(umin(P.E, 1u)));
  Counts6 c{0u, 0u, 0u, 0u, 0u, 0u};
  float sum = hd + hsr + hrr + hst + hrt;
  if (!(t_avail == 0u || sum <= 1e-6f)) {
    float scale = static_cast<float>(t_avail) / sum;
    uint32_t rem = t_avail;
    c.d = umin(scaled_count(hd, scale, umin(P.DIF, 1u)), rem);
    rem -= c.d;
    c.sr = umin(scaled_count(hsr, scale, 1u), rem);
    rem -= c.sr;
    c.rr = umin(scaled_count(hrr, scale, umin(P.ROUGH, 1u)), rem);
    rem -= c.rr;
    c.st = umin(scaled_count(hst, scale, 1u), rem);
    rem -= c.st;
    c.rt = umin(scaled_count(hrt, scale, umin(P.ROUGH, 1u)), rem);
  }
  c.l = (hl > 1e-6f && e_avail > 0u) ? umax(1u, e_avail) : 0u;
  return c;
}
enum { kNone = 0, kDiffuse, kSpecR, kRoughR, kSpecT, kRoughT };
// The light pdf multiplies a zero count everywhere below the first bounce;
// it is finite for finite emitters, so it is only evaluated when counted.
CG_HD float balance_den(const DevScene& S, f3 d, f3 next, const Surface& s, const Mat& m, const Counts6& c,
                        int model, float model_pdf, float light, const TStack& stk) {
  float pd = model == kDiffuse ? model_pdf : diffuse_pdf(d, next, s);
  float psr = model == kSpecR ? model_pdf : spec_reflect_pdf(d, next, s);
  float prr = model == kRoughR ? model_pdf : rough_reflect_pdf(d, next, s, m);
  float pst = model == kSpecT ? model_pdf : spec_trans_pdf(d, next, s, m);
  float prt = model == kRoughT ? model_pdf : rough_trans_pdf(d, next, s, m);
  float pl = light >= 0.0f ? light : (c.l != 0u ? light_pdf(S, next, s.pos, stk) : 0.0f);
  return fmx(static_cast<float>(c.d) * pd + static_cast<float>(c.sr) * psr + static_cast<float>(c.rr) * prr +
                 static_cast<float>(c.st) * pst + static_cast<float>(c.rt) * prt + static_cast<float>(c.l) * pl,
             1e-6f);
}

// One shading frame of the depth-first transport tree (shading.wesl:24-53).
struct Frame5 {
  f3 pos, gn;
  f3 tdir[4];
  float tcoeff[4];
  f3 edir;
  float erad, edist, emission, acc;
  uint32_t tcount;
  uint32_t ecount;  // 0: no light sample; else 1 + the emitter it picked
  uint32_t iter;
};
CG_HD void forward_pass(const DevScene& S, f3 d, float l, const SmitsBin& b, const Surface& s, int depth,
                        const TraceParams& P, uint32_t& rng, Frame5& x, const TStack& stk) {
  const MatRec& pm = S.mats[umin(s.material, S.mat_count - 1u)];
  Mat m = decode(S, pm, l, b, s.uv0, s.uv1);
  Counts6 c = lobe_counts(m, S.alias_count != 0u, depth, P);
  x.tcount = 0u;
  x.ecount = 0u;
  x.erad = 0.0f;
  x.edist = 0.0f;
  x.edir = mk3(0.0f, 0.0f, 0.0f);
  for (uint32_t i = 0; i < c.d; ++i) {
    TSample t = diffuse_sample(d, s, rng);
    x.tdir[x.tcount] = t.dir;
    x.tcoeff[x.tcount] = diffuse_dist(d, t.dir, s, m) / balance_den(S, d, t.dir, s, m, c, kDiffuse, t.prob, -1.0f, stk);
    x.tcount++;
  }
  for (uint32_t i = 0; i < c.sr; ++i) {
    f3 nd = specular_dir(d, s);
    float prob = spec_reflect_pdf(d, nd, s);
    x.tdir[x.tcount] = nd;
    x.tcoeff[x.tcount] = spec_reflect_dist(d, nd, s, m) / balance_den(S, d, nd, s, m, c, kSpecR, prob, -1.0f, stk);
    x.tcount++;
  }
  for (uint32_t i = 0; i < c.rr; ++i) {
    TSample t = rough_reflect_sample(d, s, m, rng);
    x.tdir[x.tcount] = t.dir;
    x.tcoeff[x.tcount] = rough_reflect_dist(d, t.dir, s, m) / balance_den(S, d, t.dir, s, m, c, kRoughR, t.prob, -1.0f, stk);
    x.tcount++;
  }
  for (uint32_t i = 0; i < c.st; ++i) {
    f3 nd = spec_trans_dir(d, s, m);
    float prob = spec_trans_pdf(d, nd, s, m);
    x.tdir[x.tcount] = nd;
    x.tcoeff[x.tcount] = spec_trans_dist(d, nd, s, m) / balance_den(S, d, nd, s, m, c, kSpecT, prob, -1.0f, stk);
    x.tcount++;
  }
  for (uint32_t i = 0; i < c.rt; ++i) {
    TSample t = rough_trans_sample(d, s, m, rng);
    x.tdir[x.tcount] = t.dir;
    x.tcoeff[x.tcount] = rough_trans_dist(d, t.dir, s, m) / balance_den(S, d, t.dir, s, m, c, kRoughT, t.prob, -1.0f, stk);
    x.tcount++;
  }
  for (uint32_t i = 0; i < c.l; ++i) {
    ESample e = emission_sample(S, s, b, rng);
    if (e.prob <= 1e-6f) continue;
    float proj = CG_FABS(dot(s.sn, e.dir));
    float coeff = (diffuse_dist(d, e.dir, s, m) + rough_reflect_dist(d, e.dir, s, m)) + rough_trans_dist(d, e.dir, s, m);
    x.edir = e.dir;
    x.erad = e.radiance * coeff * proj / balance_den(S, d, e.dir, s, m, c, kNone, -1.0f, e.prob, stk);
    x.edist = e.dist;
    x.ecount = 1u + e.prim;  // one light sample at most (MAX_EMISSION_SAMPLE_COUNT)
  }
  x.emission = m.emission;
  x.pos = s.pos;
  x.gn = s.gn;
  x.acc = 0.0f;
  x.iter = 0u;
}
// shading/spawn_surface_ray.wesl:6-15
CG_HD f3 spawn_origin(f3 pos, f3 gn, f3 next) {
  f3 off = dot(next, gn) > 0.0f ? gn : -gn;
  return pos + off * 1e-3f;
}
// environment.wesl:7-25
CG_HD float env_radiance(const DevScene& S, f3 dir, const SmitsBin& b, uint32_t& texel) {
  f3 d = normalize(dir);
  float u = atan2f(d.y, d.x) / (2.0f * kPi) + 0.5f;
  float v = acosf(clampf(d.z, -1.0f, 1.0f)) / kPi;
  float wu = u - floorf(u);
  float cv = clampf(1.0f - v, 0.0f, 0.999999f);
  uint32_t x = umin(S.env_w - 1u, f2u(wu * static_cast<float>(S.env_w)));
  uint32_t y = umin(S.env_h - 1u, f2u(cv * static_cast<float>(S.env_h)));
  texel = y * S.env_w + x;
  float4 t = S.env[texel];
  return rgb_radiance(mk3(t.x, t.y, t.z), b);
}
CG_HD uint32_t hmix(uint32_t h, uint32_t w) { return (h ^ w) * 16777619u; }

// physics/srgb.wesl:5-27
CG_HD float encode_channel(float lin) {
  float v = fmx(lin, 0.0f);
  // pow(v, 1/2.4) with the f32 exponent, as the binary32 rounding of the
  // double-precision power: the same bits on the host and the device
  float e = v <= 0.0031308f ? 12.92f * v
                            : 1.055f * static_cast<float>(pow(static_cast<double>(v),
                                                              static_cast<double>(1.0f / 2.4f))) - 0.055f;
  return clampf(e, 0.0f, 1.0f);
}
CG_HD uint32_t xyz_to_rgba8(f3 xyz) {
  f3 c0 = mk3(2.36461385f, -0.51516621f, 0.0052037f);
  f3 c1 = mk3(-0.89654057f, 1.4264081f, -0.01440816f);
  f3 c2 = mk3(-0.46807328f, 0.0887581f, 1.00920446f);
  f3 lin = (c0 * xyz.x + c1 * xyz.y) + c2 * xyz.z;
  uint32_t r = static_cast<uint32_t>(floorf(encode_channel(lin.x) * 255.0f + 0.5f));
  uint32_t g = static_cast<uint32_t>(floorf(encode_channel(lin.y) * 255.0f + 0.5f));
  uint32_t bb = static_cast<uint32_t>(floorf(encode_channel(lin.z) * 255.0f + 0.5f));
  return r | (g << 8) | (bb << 16) | (255u << 24);
}

// compute.wesl:64-79
CG_HD void camera_ray(const float* cpos, const float* cdir, const float* vp, uint32_t W, uint32_t H, uint32_t px,
                      uint32_t py, f3& o, f3& d) {
  float cx = (static_cast<float>(px) / static_cast<float>(W)) * 2.0f - 1.0f;
  float cy = (static_cast<float>(py) / static_cast<float>(H)) * 2.0f - 1.0f;
  f3 cd = ld3(cdir);
  f3 dx = normalize(cross(cd, mk3(0.0f, 0.0f, 1.0f))) * vp[0];
  f3 dy = normalize(cross(cd, dx)) * vp[1];
  o = ld3(cpos);
  d = normalize(cd + cx * dx + cy * dy);
}

}  // namespace cg

#endif  // CG_CORE_H_
