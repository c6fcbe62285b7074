// scene_impl.h — host-side scene and SceneData of the crystal::graphics layer.
//
// Scene::Impl mirrors the reference's Scene::Impl (src/scene_impl.h:32-93):
// global vertex and index arrays, materials (default material at index 0),
// primitive objects in scene-graph order with their absolute transforms, and
// an optional environment texture.  SceneData mirrors
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

struct Scene::Impl {
  std::uint64_t tag = 0;
  std::vector<cg_vertex> vertices;
  std::vector<std::uint32_t> indices;
  std::vector<cg_material> materials;
  std::vector<cg_primitive> prims;
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
  const Scene::Impl* scene = nullptr;  // vertices and materials are referenced
  cg_scene_data View() const;
};

// Build the SceneData (BLAS per primitive object, optionally shared between
// objects drawing the same index range; TLAS; emissive buffer; alias table;
// environment).  threads > 1 builds large BLAS trees in parallel with the
// same node numbering as the serial build.
SceneData BuildSceneData(const Scene::Impl& scene, bool share_blas, int threads);

// Vose alias table in double (src/pathtracing/alias_table.h:24-70).
std::vector<cg_alias_entry> BuildAliasTable(const std::vector<double>& weights);

}  // namespace crystal::graphics

#endif
