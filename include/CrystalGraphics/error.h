// Error type of the crystal::graphics API (include/CrystalGraphics/error.h:8
// in the reference) and a minimal std::expected stand-in for standard
// libraries without <expected>.
#ifndef CRYSTALGRAPHICS_ERROR_H_
#define CRYSTALGRAPHICS_ERROR_H_

#include <string>
#include <utility>
#include <variant>

#if defined(__has_include)
#if __has_include(<expected>) && __cplusplus > 202002L
#include <expected>
#endif
#endif

namespace crystal::graphics {

using Error = std::string;

#if defined(__cpp_lib_expected)
template <class T, class E>
using expected = std::expected<T, E>;
template <class E>
using unexpected = std::unexpected<E>;
#else
template <class E>
class unexpected {
 public:
  explicit unexpected(E e) : e_(std::move(e)) {}
  const E& error() const& { return e_; }
  E&& error() && { return std::move(e_); }

 private:
  E e_;
};
template <class E>
unexpected(E) -> unexpected<E>;

template <class T, class E>
class expected {
 public:
  expected() : v_(std::in_place_index<0>) {}
  expected(T v) : v_(std::in_place_index<0>, std::move(v)) {}
  expected(unexpected<E> u) : v_(std::in_place_index<1>, std::move(u).error()) {}
  bool has_value() const { return v_.index() == 0; }
  explicit operator bool() const { return has_value(); }
  T& value() & { return std::get<0>(v_); }
  const T& value() const& { return std::get<0>(v_); }
  T& operator*() & { return std::get<0>(v_); }
  const T& operator*() const& { return std::get<0>(v_); }
  T* operator->() { return &std::get<0>(v_); }
  const T* operator->() const { return &std::get<0>(v_); }
  const E& error() const& { return std::get<1>(v_); }

 private:
  std::variant<T, E> v_;
};
template <class E>
class expected<void, E> {
 public:
  expected() = default;
  expected(unexpected<E> u) : ok_(false), e_(std::move(u).error()) {}
  bool has_value() const { return ok_; }
  explicit operator bool() const { return ok_; }
  const E& error() const& { return e_; }

 private:
  bool ok_ = true;
  E e_{};
};
#endif

}  // namespace crystal::graphics

#endif
