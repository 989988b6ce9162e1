// caviar2/vector3d.hpp -- 3-vector value type of the caviar2 API.
//
// Same interface and arithmetic as the reference's Vector3d<T>
// (inc/caviar2/vector3d.hpp:24-135): component-wise +, -, scalar * and /,
// and operator* between two vectors is the dot product evaluated as
// (x*x' + y*y') + z*z', the order the FP64 pair predicate depends on.
#pragma once
#include <cmath>
#include <iostream>

namespace caviar2 {

template <typename T>
struct Vector3d {
  T x, y, z;
  Vector3d() : x(0), y(0), z(0) {}
  Vector3d(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
};

template <typename T>
inline Vector3d<T> operator+(const Vector3d<T> &a) { return a; }
template <typename T>
inline Vector3d<T> operator-(const Vector3d<T> &a) { return Vector3d<T>{-a.x, -a.y, -a.z}; }
template <typename T>
inline Vector3d<T> operator+(const Vector3d<T> &a, const Vector3d<T> &b) {
  return Vector3d<T>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename T>
inline Vector3d<T> operator-(const Vector3d<T> &a, const Vector3d<T> &b) {
  return Vector3d<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T>
inline T operator*(const Vector3d<T> &a, const Vector3d<T> &b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T>
inline Vector3d<T> operator*(const Vector3d<T> &a, const double &s) {
  return Vector3d<T>{a.x * s, a.y * s, a.z * s};
}
template <typename T>
inline Vector3d<T> operator*(const double &s, const Vector3d<T> &a) {
  return Vector3d<T>{a.x * s, a.y * s, a.z * s};
}
template <typename T>
inline Vector3d<T> operator/(const Vector3d<T> &a, const double &s) {
  return Vector3d<T>{a.x / s, a.y / s, a.z / s};
}
template <typename T1, typename T2>
inline Vector3d<T1> operator/(const Vector3d<T1> &a, const Vector3d<T2> &b) {
  return Vector3d<T1>{a.x / b.x, a.y / b.y, a.z / b.z};
}
template <typename T>
inline bool operator==(const Vector3d<T> &a, const Vector3d<T> &b) {
  return a.x == b.x && a.y == b.y && a.z == b.z;
}
template <typename T>
inline Vector3d<T> &operator+=(Vector3d<T> &a, const Vector3d<T> &b) {
  a.x += b.x; a.y += b.y; a.z += b.z;
  return a;
}
template <typename T>
inline Vector3d<T> &operator-=(Vector3d<T> &a, const Vector3d<T> &b) {
  a.x -= b.x; a.y -= b.y; a.z -= b.z;
  return a;
}
template <typename T>
inline Vector3d<T> &operator*=(Vector3d<T> &a, const double &s) {
  a.x *= s; a.y *= s; a.z *= s;
  return a;
}
template <typename T>
inline Vector3d<T> &operator/=(Vector3d<T> &a, const double &s) {
  a.x /= s; a.y /= s; a.z /= s;
  return a;
}
template <typename T>
inline std::ostream &operator<<(std::ostream &o, const Vector3d<T> &a) {
  return o << a.x << ' ' << a.y << ' ' << a.z;
}
template <typename T>
constexpr Vector3d<T> cross_product(const Vector3d<T> &a, const Vector3d<T> &b) {
  return Vector3d<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T>
constexpr T dot_product(const Vector3d<T> &a, const Vector3d<T> &b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T>
inline T norm(const Vector3d<T> &a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

}  // namespace caviar2
