// scene_impl.h — host-side scene and SceneData of the crystal::graphics layer.
//
// Scene::Impl mirrors the reference's Scene::Impl (src/scene_impl.h:32-93):
// global vertex and index arrays, materials (default material at index 0),
// primitive objects in scene-graph order with their absolute transforms,
// decoded material textures, and an optional environment texture.  SceneData mirrors
// src/pathtracing/scene_data.h:23-54: the flattened buffers the device
// backend consumes.
#ifndef CG_HOST_SCENE_IMPL_H_
#define CG_HOST_SCENE_IMPL_H_

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/CrystalGraphics/scene.h"
#include "../../../include/cg_scene.h"

namespace crystal::graphics {

// A decoded material texture (src/texture.h:12-22 after DecodeImageRgba8).
struct HostTexture {
  std::uint32_t width = 0, height = 0;
  std::int32_t wrap_s = CG_WRAP_REPEAT, wrap_t = CG_WRAP_REPEAT;
  std::vector<std::uint8_t> rgba8;  // width * height * 4
};

struct Scene::Impl {
  std::uint64_t tag = 0;
  std::vector<cg_vertex> vertices;
  std::vector<std::uint32_t> indices;
  std::vector<cg_material> materials;
  std::vector<cg_primitive> prims;
  std::vector<HostTexture> textures;  // glTF texture index -> image
  mutable std::vector<cg_texture> texture_views;  // C views of `textures` (cgh_scene_describe)
  std::vector<float> env_rgba32f;  // empty: fallback sky
  std::uint32_t env_w = 0, env_h = 0;
  // Suggested view for procedural scenes.
  cg_camera camera{};
  cg_conf conf{};
};

// The reference's default material (src/scene.cpp:659-706).
cg_material DefaultMaterial();

// Procedural scenes; `name` is one of "c0", "c1", "c2", "quad", "spheres".
// Returns an empty string on success, else the error.
std::string GenerateScene(const std::string& name, std::uint64_t seed, const double* params, int nparams,
                          Scene::Impl& out);
// glTF 2.0 scene file (.gltf or .glb); returns an empty string on success.
std::string LoadGltf(const std::string& path, Scene::Impl& out);
// Ray batch of the traversal microbenchmark over a generated "c2" scene:
// half pinhole-camera rays on a grid, half surface-origin cosine rays.
void GenerateRays(const Scene::Impl& scene, std::uint64_t seed, std::uint32_t n, cg_ray* out);

struct SceneData {
  std::vector<cg_tlas_node> tlas;
  std::vector<cg_instance> instances;
  std::vector<cg_blas_node> blas;
  std::vector<cg_index> indices;
  std::vector<std::uint32_t> tri_ids;
  std::vector<cg_emissive_prim> emissive;
  std::vector<cg_alias_entry> alias;
  std::vector<float> env;
  std::uint32_t env_w = 0, env_h = 0;
  // Material texture array (src/wgpu/resource.cpp:471-546): every texture
  // resized nearest to the largest width and height; empty when the scene has
  // no textures.
  std::vector<std::uint8_t> tex;
  std::uint32_t tex_w = 0, tex_h = 0, tex_layers = 0;
  const Scene::Impl* scene = nullptr;  // vertices and materials are referenced
  cg_scene_data View() const;
};

// Build the SceneData (BLAS per primitive object, optionally shared between
// objects drawing the same index range; TLAS; emissive buffer; alias table;
// environment).  threads > 1 builds large BLAS trees in parallel with the
// same node numbering as the serial build.
SceneData BuildSceneData(const Scene::Impl& scene, bool share_blas, int threads);

// Nearest resize of an RGBA8 image (src/wgpu/resource.cpp:51-75).
std::vector<std::uint8_t> ResizeRgba8Nearest(const HostTexture& t, std::uint32_t w, std::uint32_t h);

// Vose alias table in double (src/pathtracing/alias_table.h:24-70).
std::vector<cg_alias_entry> BuildAliasTable(const std::vector<double>& weights);

}  // namespace crystal::graphics

#endif
