/*
 * cg_scene.h — data contract of the spectral path-tracing core.
 *
 * Plain C records, byte-for-byte the layouts of the reference's host-side
 * scene description and of the device buffers its compute pass binds
 * (bind group 2, uniform, history).  These are the inputs the C-ABI in
 * cg_device.h consumes; the MI355X device layouts (64-byte pair nodes,
 * 48-byte triangle records, ...) are private to the library and built from
 * these at upload time.
 *
 * Reference layouts mirrored here:
 *   cg_vertex        src/vertex.h:12-17                     48 B
 *   cg_tex_info      src/material.h:23-28                   16 B
 *   cg_texture       src/texture.h:12-22 after DecodeImageRgba8 (scene input)
 *   cg_material      src/material.h:30-87                  432 B
 *   cg_primitive     src/scene_impl.h:24-30 + AbsTrans()    (scene input)
 *   cg_tlas_node     src/pathtracing/tlas.h:17-30           32 B
 *   cg_instance      src/pathtracing/instance.h:9-17        80 B
 *   cg_blas_node     src/pathtracing/blas.h:15-28           32 B
 *   cg_index         src/pathtracing/blas.h:30-32           16 B
 *   cg_emissive_prim src/pathtracing/emissive_buffer.h:27-40 80 B
 *   cg_alias_entry   src/pathtracing/alias_table.h:14-17     8 B
 *   cg_camera        include/CrystalGraphics/camera.h:9-13  32 B
 *   cg_history_pixel src/wgpu/shader/compute.wesl:18-21     16 B
 */
#ifndef CG_SCENE_H_
#define CG_SCENE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Vertex: position, normal, two uv sets (src/vertex.h:12-17). */
typedef struct cg_vertex {
  float pos[3];
  float _pad0;
  float nrm[3];
  float _pad1;
  float uv0[2];
  float uv1[2];
} cg_vertex;

/* glTF texture reference (src/material.h:23-28). index < 0: no texture. */
typedef struct cg_tex_info {
  int32_t index;
  uint32_t tex_coord;
  float scale;
  float strength;
} cg_tex_info;

/* A decoded material texture (the reference's TextureData after
 * DecodeImageRgba8, src/texture.h:12-22, src/image_decode.h): RGBA8 texels,
 * rows top to bottom, and the glTF sampler's wrap modes.  Only the emissive
 * buffer's centroid lookup honours wrap_s / wrap_t
 * (src/pathtracing/emissive_buffer.h:170-204); the device sampler always
 * repeats (src/wgpu/resource.cpp:589-606). */
typedef struct cg_texture {
  const uint8_t* rgba8;  /* width * height * 4 bytes */
  uint32_t width, height;
  int32_t wrap_s, wrap_t;
} cg_texture;

#define CG_WRAP_REPEAT 10497
#define CG_WRAP_CLAMP_TO_EDGE 33071
#define CG_WRAP_MIRRORED_REPEAT 33648

enum {
  CG_TEX_BASE_COLOR = 0,
  CG_TEX_METALLIC_ROUGHNESS = 1,
  CG_TEX_NORMAL = 2,
  CG_TEX_OCCLUSION = 3,
  CG_TEX_EMISSIVE = 4,
  CG_TEX_TRANSMISSION = 15,
  CG_TEX_COUNT = 17
};

/* glTF PBR material, packed as the reference uploads it (src/material.h:30-87,
 * shader/compute/shading/material.wesl:119-166). */
typedef struct cg_material {
  float base_color_factor[4];
  float emissive_factor[3];
  float emissive_strength;
  float metallic_factor;
  float roughness_factor;
  float alpha_cutoff;
  uint32_t alpha_mode;
  float ior;
  float dispersion;
  float transmission_factor;
  float thickness_factor;
  float attenuation_color[3];
  float attenuation_distance;
  float anisotropy_strength;
  float anisotropy_rotation;
  float clearcoat_factor;
  float clearcoat_roughness_factor;
  float iridescence_factor;
  float iridescence_ior;
  float iridescence_thickness_minimum;
  float iridescence_thickness_maximum;
  float sheen_color_factor[3];
  float sheen_roughness_factor;
  float specular_color_factor[3];
  float specular_factor;
  uint32_t flags;
  uint32_t _padding[3];
  cg_tex_info tex[CG_TEX_COUNT];
} cg_material;

/* One drawable primitive object in scene-graph iteration order (the
 * reference's ObjView<Primitive> order, src/scene_impl.h:24-30), with its
 * absolute transform.  `world` is column-major (glm convention).  `inv_world`
 * is its inverse; the reference computes it with glm::inverse
 * (src/pathtracing/tlas.h:61-65), which is not available here, so it is part
 * of the scene input and computed once by the scene's author. */
typedef struct cg_primitive {
  uint32_t vertex_offset;
  uint32_t vertex_count;
  uint32_t index_offset;
  uint32_t index_count;
  uint32_t material_idx;
  uint32_t _pad[3];
  float world[16];
  float inv_world[16];
} cg_primitive;

/* TLAS node (src/pathtracing/tlas.h:17-30): child == 0 marks a leaf. */
typedef struct cg_tlas_node {
  float lb[3];
  uint32_t child;
  float ub[3];
  uint32_t instance_idx;
} cg_tlas_node;

/* TLAS instance (src/pathtracing/instance.h:9-17); inv_trans column-major. */
typedef struct cg_instance {
  float inv_trans[16];
  uint32_t blas_root_idx;
  uint32_t material_idx;
  uint32_t _pad[2];
} cg_instance;

/* BLAS node (src/pathtracing/blas.h:15-28): triangle_count == 0 marks an
 * interior node whose children are child and child + 1; a leaf holds
 * triangle_count (1 or 2) BVH-ordered triangles starting at child. */
typedef struct cg_blas_node {
  float lb[3];
  uint32_t child;
  float ub[3];
  uint32_t triangle_count;
} cg_blas_node;

/* BVH-ordered triangle vertex indices (src/pathtracing/blas.h:30-32). */
typedef struct cg_index {
  uint32_t i0, i1, i2;
  uint32_t _pad;
} cg_index;

/* Emissive triangle in world space (src/pathtracing/emissive_buffer.h:27-40). */
typedef struct cg_emissive_prim {
  float v[3][4];
  float emission[3];
  float area;
  float weight;
  float _pad[3];
} cg_emissive_prim;

/* Vose alias-table entry (src/pathtracing/alias_table.h:14-17). */
typedef struct cg_alias_entry {
  uint32_t alias;
  float prob;
} cg_alias_entry;

/* Camera (include/CrystalGraphics/camera.h:9-13). */
typedef struct cg_camera {
  float position[3];
  float _pad0;
  float direction[3];
  float _pad1;
  float viewport[2];
  float _pad2[2];
} cg_camera;

/* Per-pixel progressive history (compute.wesl:18-21): running-mean CIE XYZ
 * and the pixel's carried RNG state. */
typedef struct cg_history_pixel {
  float xyz[3];
  uint32_t rand_seed;
} cg_history_pixel;

/* Path-trace configuration (include/CrystalGraphics/environment.h:11-18);
 * the window fields are kept for API fidelity and ignored (headless). */
typedef struct cg_conf {
  uint32_t window_width, window_height;
  uint32_t render_width, render_height;
  uint32_t sample_count;
  uint32_t trace_depth;
  uint32_t max_transport_sample_count;
  uint32_t max_diffuse_sample_count;
  uint32_t max_rough_sample_count;
  uint32_t max_emission_sample_count;
} cg_conf;

/* A ray for the batch closest-hit entry point: 32 B. */
typedef struct cg_ray {
  float org[3];
  float _pad0;
  float dir[3];
  float _pad1;
} cg_ray;

/* Closest-hit result: 16 B.  Miss: t = 1e6 and inst = tri = 0xFFFFFFFF.
 * `inst` is the instance (primitive object) index, `tri` the triangle's index
 * within its primitive's source index buffer. */
typedef struct cg_hit {
  float t;
  uint32_t inst;
  uint32_t tri;
  uint32_t _pad;
} cg_hit;

#define CG_MISS_ID 0xFFFFFFFFu
#define CG_MISS_T 1e6f

/* The flattened scene the reference's SceneData hands its device backend
 * (src/pathtracing/scene_data.h:23-54, uploaded by src/wgpu/resource.cpp:
 * 338-414), plus `tri_ids`: for each BVH-ordered triangle slot, its index in
 * its primitive's source index buffer (the reference keeps no triangle id).
 *
 * The material texture array is what WriteMaterialTextures uploads
 * (src/wgpu/resource.cpp:471-546): tex_layers images (layer = glTF texture
 * index), each resized nearest to the largest width and height, RGBA8 texels
 * row-major, top row first.  tex_layers == 0 stands for the reference's 1x1
 * white layer: every texture lookup then reads 1.0. */
typedef struct cg_scene_data {
  const cg_tlas_node* tlas_nodes;   uint32_t tlas_node_count;
  const cg_instance* instances;     uint32_t instance_count;
  const cg_blas_node* blas_nodes;   uint32_t blas_node_count;
  const cg_index* indices;          uint32_t index_count;   /* triangles */
  const uint32_t* tri_ids;          /* index_count entries */
  const cg_vertex* vertices;        uint32_t vertex_count;
  const cg_material* materials;     uint32_t material_count;
  const cg_emissive_prim* emissive; uint32_t emissive_count;
  const cg_alias_entry* alias;      uint32_t alias_count;
  const float* env_rgba32f;         uint32_t env_width, env_height;
  const uint8_t* tex_rgba8;         uint32_t tex_width, tex_height, tex_layers;
} cg_scene_data;

#ifdef __cplusplus
}
#endif

#endif /* CG_SCENE_H_ */
