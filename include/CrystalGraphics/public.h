// Shared small types (include/CrystalGraphics/public.h:11 in the reference),
// plus the glm-compatible vectors the API uses (glm is not a dependency here).
#ifndef CRYSTALGRAPHICS_PUBLIC_H_
#define CRYSTALGRAPHICS_PUBLIC_H_

#include <cstdint>
#include <type_traits>

namespace crystal::graphics {

using size32_t = uint32_t;

// Plain float vectors with the layout of glm::vec2 / glm::vec3.  They convert
// from and to any vector type with x, y (and z) members that is constructible
// from two (three) floats, so a caller that assigns glm vectors into a Camera,
// or reads them back into glm vectors, compiles unchanged without glm here.
struct vec2 {
  float x = 0.0f, y = 0.0f;
  constexpr vec2() = default;
  constexpr vec2(float x_, float y_) : x(x_), y(y_) {}
  template <class V, class = decltype(static_cast<float>(std::declval<const V&>().x) +
                                      static_cast<float>(std::declval<const V&>().y)),
            class = std::enable_if_t<!std::is_same_v<std::decay_t<V>, vec2>>>
  constexpr vec2(const V& v) : x(static_cast<float>(v.x)), y(static_cast<float>(v.y)) {}
  template <class V, class = std::enable_if_t<!std::is_same_v<V, vec2> && std::is_constructible_v<V, float, float> &&
                                              !std::is_arithmetic_v<V>>>
  constexpr operator V() const {
    return V(x, y);
  }
};
struct vec3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
  constexpr vec3() = default;
  constexpr vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
  template <class V, class = decltype(static_cast<float>(std::declval<const V&>().x) +
                                      static_cast<float>(std::declval<const V&>().y) +
                                      static_cast<float>(std::declval<const V&>().z)),
            class = std::enable_if_t<!std::is_same_v<std::decay_t<V>, vec3>>>
  constexpr vec3(const V& v) : x(static_cast<float>(v.x)), y(static_cast<float>(v.y)), z(static_cast<float>(v.z)) {}
  template <class V, class = std::enable_if_t<!std::is_same_v<V, vec3> &&
                                              std::is_constructible_v<V, float, float, float> &&
                                              !std::is_arithmetic_v<V>>>
  constexpr operator V() const {
    return V(x, y, z);
  }
};
static_assert(sizeof(vec2) == 8 && sizeof(vec3) == 12, "glm::vec2 / glm::vec3 layouts");

}  // namespace crystal::graphics

#endif
