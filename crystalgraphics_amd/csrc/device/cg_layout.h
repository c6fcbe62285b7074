// cg_layout.h — MI355X device layouts of the scene, built from the reference
// SceneData (cg_scene.h) at upload time.
//
// Every record is a whole number of 16-byte vectors and either 64 B or a
// multiple of 16 B, so one traversal step is a handful of dwordx4 loads from
// one or two 128-byte lines:
//   Node8Q   128 B  one 8-wide BVH node (one line): quantised boxes of up
//                   to eight children, their refs
//   PairNode  64 B  both child boxes of one interior node of the binary
//                   emitter BVH + child refs
//   TriRec    48 B  v0, e1 = v1 - v0, e2 = v2 - v0, source triangle id (and,
//                   in the merged world BLAS, the triangle's instance)
//   InstRec   80 B  full column-major inverse transform, BLAS root ref, material
//   VAttr     32 B  normal + both uv sets (read once, for the final hit)
//   MatRec    64 B  material factors and texture layers
//   EmitRec   64 B  emissive triangle, its unit normal, RGB and area pdf
#ifndef CG_LAYOUT_H_
#define CG_LAYOUT_H_

#include <stdint.h>

// Child / root references (32 bits):
//   00 | index       interior node (30-bit index: Node8Q in the scene BVH,
//                    PairNode in the emitter BVH)
//   01 | index       TLAS leaf: instance index (30-bit)
//   1 cc | slot      BLAS leaf: cc + 1 triangles (1-4) from slot (29-bit)
//   kRefNone         empty (BLAS with no triangles, or empty TLAS)
//   kRefMarker       traversal-stack marker: return from a BLAS to the TLAS
#define CG_REF_LEAF 0x80000000u
#define CG_REF_INST 0x40000000u
#define CG_REF_INDEX_MASK 0x3FFFFFFFu
#define CG_REF_SLOT_MASK 0x1FFFFFFFu
#define CG_REF_COUNT_SHIFT 29
#define CG_REF_NONE 0x3FFFFFFFu
#define CG_REF_MARKER 0x7FFFFFFFu
#define CG_LEAF_MAX 4u
#define CG_MAKE_LEAF(slot, count) (CG_REF_LEAF | (((count) - 1u) << CG_REF_COUNT_SHIFT) | (slot))

// Up to eight children: exact boxes (build time only).
#define CG_WIDE 8
struct Node8 {
  float lo[3][CG_WIDE], hi[3][CG_WIDE];  // [axis][child]; an empty slot has ref CG_REF_NONE
  uint32_t ref[CG_WIDE];
};

// The node the kernels traverse, 128 B (one line): child boxes as 8-bit
// offsets on a per-axis power-of-two grid from the node's lower corner,
// rounded outward (decoded lo <= true lo, decoded hi >= true hi, checked at
// build time with the decoder's own float operations).
struct __attribute__((aligned(16))) Node8Q {
  float origin[3];
  uint32_t exps;                // bits 0-7, 8-15, 16-23: biased exponents of the x, y, z grid steps
  uint8_t qlo[3][CG_WIDE];      // [axis][child]
  uint8_t qhi[3][CG_WIDE];
  uint32_t ref[CG_WIDE];
  uint32_t _pad[8];
};

struct __attribute__((aligned(16))) PairNode {
  float lo_l[3];   // left child box
  float hi_l[3];
  float lo_r[3];   // right child box
  float hi_r[3];
  uint32_t ref_l;
  uint32_t ref_r;
  uint32_t _pad[2];
};

struct __attribute__((aligned(16))) TriRec {
  float v0[3];
  float e1[3];
  float e2[3];
  uint32_t tri_id;
  uint32_t inst;  // instance of a triangle merged into the world BLAS, else CG_MISS_ID
  uint32_t _pad;
};

struct __attribute__((aligned(16))) InstRec {
  float inv[16];    // column-major
  uint32_t root;    // BLAS root ref
  uint32_t material;
  uint32_t _pad[2];
};

struct __attribute__((aligned(16))) VAttr {
  float nrm[3];
  float uv0[2];
  float uv1[2];
  float _pad;
};

// Material factors as glTF gives them; the kernels multiply in the texture
// samples and clamp (material.wesl:181-262).  Texture layers are 16-bit, the
// value CG_TEX_NONE meaning no texture.
#define CG_TEX_NONE 0xFFFFu
#define CG_MAT_VOLUME 1u          // thickness_factor > 0
#define CG_MAT_UV1_BASE 2u        // tex_coord == 1, per textured slot
#define CG_MAT_UV1_MR 4u
#define CG_MAT_UV1_TRANS 8u
#define CG_MAT_UV1_EMIT 16u
struct __attribute__((aligned(16))) MatRec {
  float base[3];       // base_color_factor rgb
  float metallic;      // metallic_factor
  float roughness;     // roughness_factor
  float transmission;  // transmission_factor
  float ior;
  float dispersion;
  float emission[3];   // emissive_factor
  float strength;      // emissive_strength
  uint32_t flags;      // CG_MAT_*
  uint32_t tex_base_mr;   // base colour layer | metallic-roughness layer << 16
  uint32_t tex_trans_emit;  // transmission layer | emissive layer << 16
  uint32_t _pad;
};

struct __attribute__((aligned(16))) EmitRec {
  float v0[3];
  float ln_x;
  float v1[3];
  float ln_y;
  float v2[3];
  float ln_z;
  float rgb[3];
  float area_pdf;      // max(weight / area, 1e-3)
};

// Wavelength sampling table (one per library): the reference's 16-step
// bisection (shader/compute/physics/wavelength.wesl:20-45) unrolled into a
// heap of its 65535 midpoint CDF values and the 65536 possible outcomes.
#define CG_WL_LEVELS 16
#define CG_WL_LEAVES 65536
struct __attribute__((aligned(16))) WlLeaf {
  float lambda;
  float prob;      // max(xyzMatchSum(lambda) / total, 1e-6)
  float x, y, z;   // xyzMatch(lambda)
  float _pad[3];
};

static_assert(sizeof(Node8Q) == 128, "Node8Q");
static_assert(sizeof(PairNode) == 64, "PairNode");
static_assert(sizeof(TriRec) == 48, "TriRec");
static_assert(sizeof(InstRec) == 80, "InstRec");
static_assert(sizeof(VAttr) == 32, "VAttr");
static_assert(sizeof(MatRec) == 64, "MatRec");
static_assert(sizeof(EmitRec) == 64, "EmitRec");
static_assert(sizeof(WlLeaf) == 32, "WlLeaf");

#endif  // CG_LAYOUT_H_
