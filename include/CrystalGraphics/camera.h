// Camera (include/CrystalGraphics/camera.h:9-13 in the reference): same
// fields, defaults and 16/16/8-byte alignment.
#ifndef CRYSTALGRAPHICS_CAMERA_H_
#define CRYSTALGRAPHICS_CAMERA_H_

#include "public.h"

namespace crystal::graphics {

struct Camera {
  alignas(16) vec3 position{0, 0, 0};
  alignas(16) vec3 direction{1, 0, 0};
  alignas(8) vec2 viewport{0.5, 0.5};
};

}  // namespace crystal::graphics

#endif
