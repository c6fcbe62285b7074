// Shared small types (include/CrystalGraphics/public.h:11 in the reference),
// plus the glm-compatible vectors the API uses (glm is not a dependency here).
#ifndef CRYSTALGRAPHICS_PUBLIC_H_
#define CRYSTALGRAPHICS_PUBLIC_H_

#include <cstdint>

namespace crystal::graphics {

using size32_t = uint32_t;

struct vec2 {
  float x = 0.0f, y = 0.0f;
};
struct vec3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
};

}  // namespace crystal::graphics

#endif
