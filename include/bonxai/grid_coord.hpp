// grid_coord.hpp -- coordinate and point types of the GPU-backed Bonxai drop-in.
//
// Same names, members and conversions as the reference's
// bonxai_core/include/bonxai/grid_coord.hpp:
//   Point3D                        grid_coord.hpp:36-61
//   CoordT                         grid_coord.hpp:63-80
//   PosToCoord / CoordToPos        grid_coord.hpp:80-92
//   ConvertPoint (x() / .x / [i])  grid_coord.hpp:134-162
//   std::hash<CoordT>              grid_coord.hpp:209-218
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

namespace Bonxai {

template <typename PointOut, typename PointIn>
PointOut ConvertPoint(const PointIn& v);

/// A point in world coordinates (metres).
struct Point3D {
  double x;
  double y;
  double z;

  Point3D() = default;
  Point3D(const Point3D&) = default;
  Point3D(Point3D&&) = default;
  Point3D& operator=(const Point3D&) = default;
  Point3D& operator=(Point3D&&) = default;

  Point3D(double px, double py, double pz) : x(px), y(py), z(pz) {}

  /// Any xyz-like type: x()/y()/z() methods, x/y/z members or v[0..2].
  template <typename T>
  Point3D(const T& v) {
    *this = ConvertPoint<Point3D>(v);
  }

  template <typename T>
  Point3D& operator=(const T& v) {
    *this = ConvertPoint<Point3D>(v);
    return *this;
  }

  /// x, y, z by index 0, 1, 2.
  [[nodiscard]] double& operator[](size_t index) {
    if (index == 0) return x;
    if (index == 1) return y;
    if (index == 2) return z;
    throw std::runtime_error("out of bound index");
  }
};

/// Integer voxel coordinate.
struct CoordT {
  int32_t x;
  int32_t y;
  int32_t z;

  [[nodiscard]] int32_t& operator[](size_t index) {
    if (index == 0) return x;
    if (index == 1) return y;
    if (index == 2) return z;
    throw std::runtime_error("out of bound index");
  }
  [[nodiscard]] bool operator==(const CoordT& o) const { return x == o.x && y == o.y && z == o.z; }
  [[nodiscard]] bool operator!=(const CoordT& o) const { return !(*this == o); }
  [[nodiscard]] CoordT operator+(const CoordT& o) const { return {x + o.x, y + o.y, z + o.z}; }
  [[nodiscard]] CoordT operator-(const CoordT& o) const { return {x - o.x, y - o.y, z - o.z}; }
  CoordT& operator+=(const CoordT& o) {
    x += o.x;
    y += o.y;
    z += o.z;
    return *this;
  }
  CoordT& operator-=(const CoordT& o) {
    x -= o.x;
    y -= o.y;
    z -= o.z;
    return *this;
  }
};

/// floor(p * inv_resolution) per axis, in double (grid_coord.hpp:80-85).
[[nodiscard]] inline CoordT PosToCoord(const Point3D& point, double inv_resolution) {
  return {static_cast<int32_t>(std::floor(point.x * inv_resolution)),
          static_cast<int32_t>(std::floor(point.y * inv_resolution)),
          static_cast<int32_t>(std::floor(point.z * inv_resolution))};
}

/// Minimum corner of a voxel (grid_coord.hpp:87-92).
[[nodiscard]] inline Point3D CoordToPos(const CoordT& coord, double resolution) {
  return {static_cast<double>(coord.x) * resolution, static_cast<double>(coord.y) * resolution,
          static_cast<double>(coord.z) * resolution};
}

namespace point_traits {

template <class T, class = void>
struct has_method_x : std::false_type {};
template <class T>
struct has_method_x<T, std::void_t<decltype(std::declval<const T&>().x())>> : std::true_type {};

template <class T, class = void>
struct has_member_x : std::false_type {};
template <class T>
struct has_member_x<T, std::void_t<decltype(std::declval<const T&>().x)>>
    : std::integral_constant<bool, !has_method_x<T>::value> {};

template <class T>
struct is_xyz_sequence : std::false_type {};
template <class T, class A>
struct is_xyz_sequence<std::vector<T, A>> : std::true_type {};
template <class T>
struct is_xyz_sequence<std::array<T, 3>> : std::true_type {};

template <class T>
inline constexpr bool is_point_v =
    has_method_x<T>::value || has_member_x<T>::value || is_xyz_sequence<T>::value;

/// component i of an xyz-like value, by whichever access the type offers
template <class T>
inline auto get(const T& v, int i) {
  if constexpr (has_method_x<T>::value) {
    return i == 0 ? v.x() : (i == 1 ? v.y() : v.z());
  } else if constexpr (has_member_x<T>::value) {
    return i == 0 ? v.x : (i == 1 ? v.y : v.z);
  } else {
    return v[i];
  }
}

/// the scalar type of an xyz-like type
template <class T>
using scalar_t = std::decay_t<decltype(get(std::declval<const T&>(), 0))>;

}  // namespace point_traits

/// Convert between xyz-like types (grid_coord.hpp:134-162): the source is read
/// with x()/y()/z(), .x/.y/.z or [0..2], and the target is brace-initialised
/// from the three values.
template <typename PointOut, typename PointIn>
inline PointOut ConvertPoint(const PointIn& v) {
  static_assert(std::is_same_v<PointIn, PointOut> || point_traits::is_point_v<PointIn>,
                "Can't convert from the specified type");
  static_assert(std::is_same_v<PointIn, PointOut> || point_traits::is_point_v<PointOut>,
                "Can't convert to the specified type");
  if constexpr (std::is_same_v<PointIn, PointOut>) {
    return v;
  } else {
    return {point_traits::get(v, 0), point_traits::get(v, 1), point_traits::get(v, 2)};
  }
}

}  // namespace Bonxai

namespace std {
/// The OpenVDB-style 20-bit coordinate hash of the reference (grid_coord.hpp:209-218).
template <>
struct hash<Bonxai::CoordT> {
  std::size_t operator()(const Bonxai::CoordT& p) const {
    const int64_t h = (static_cast<int64_t>(p.x) * 73856093) ^
                      (static_cast<int64_t>(p.y) * 19349669) ^
                      (static_cast<int64_t>(p.z) * 83492791);
    return static_cast<std::size_t>(h & ((1 << 20) - 1));
  }
};
}  // namespace std
