// Scene handle (include/CrystalGraphics/scene.h:12-37 in the reference):
// move-only, pimpl.  LoadScene reads glTF 2.0 files (.gltf, .glb) as the
// reference does, and the procedural scenes of this library
// ("procedural:<name>?seed=<n>").
#ifndef CRYSTALGRAPHICS_SCENE_H_
#define CRYSTALGRAPHICS_SCENE_H_

#include <filesystem>
#include <memory>

#include "error.h"

namespace crystal::graphics {

class Scene {
 public:
  const std::filesystem::path& FilePath() const { return filepath_; }
  ~Scene();
  Scene(Scene&& other) noexcept;

  struct Impl;
  const Impl& impl() const { return *impl_; }

 private:
  std::filesystem::path filepath_;
  std::unique_ptr<Impl> impl_;

  Scene(const std::filesystem::path& filepath, std::unique_ptr<Impl> impl);
  friend expected<Scene, Error> LoadScene(std::filesystem::path file);
  friend Scene MakeScene(const std::filesystem::path&, std::unique_ptr<Impl>);
};

expected<Scene, Error> LoadScene(std::filesystem::path file);

}  // namespace crystal::graphics

#endif
