// Rendering entry points (include/CrystalGraphics/api.h:13 in the reference),
// plus the headless ViewToBuffer: one View, then the running-mean CIE XYZ
// image (render_width * render_height * 3 floats) copied out.
#ifndef CRYSTALGRAPHICS_API_H_
#define CRYSTALGRAPHICS_API_H_

#include <cstddef>

#include "camera.h"
#include "error.h"
#include "scene.h"

namespace crystal::graphics {

expected<void, Error> View(const Scene& scene, const Camera& camera = {});
expected<void, Error> ViewToBuffer(const Scene& scene, const Camera& camera, float* xyz, std::size_t count);

}  // namespace crystal::graphics

#endif
