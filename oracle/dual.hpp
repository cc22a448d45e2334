// dual.hpp — forward-mode (dual-number) arithmetic for the CPU oracle.  TEST
// INFRASTRUCTURE ONLY (see mpm_oracle.hpp).
//
// Dual carries an FP64 value and kTangents FP64 directional derivatives.
// Instantiating the oracle's substep with R = Dual runs the FP64 twin (the
// reference's algorithm in FP64, every branch decided on the value) and
// propagates the derivatives of every quantity along kTangents input
// directions at once: the exact derivative of the FP64 substep map, with no
// step size.  It checks the reverse-mode kernels where finite differences are
// ill-posed (a perturbation moves particles across the yield surface).
//
// DualT<float> (Dual32) is the same arithmetic on FP32 values and tangents:
// its values follow the FP32 oracle's, rounding for rounding (the FP32
// oracle also computes with every operation rounded as written), and its
// tangents are the derivatives of that FP32 map taken in FP32.  How far
// they stray from Dual's shows what differentiation in FP32 can reach on a
// given trajectory.
//
// Rules: +, -, *, / and the functions below differentiate as usual; floor,
// comparisons and the integer conversions act on the value (their derivative
// is zero or undefined, and the oracle only uses them to pick cells and
// branches); |x| takes the derivative of the branch x >= 0 at 0; sqrt has
// derivative 0 at 0 (the oracle's sqrt(0) arguments are zero-length vectors
// guarded by a branch).
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

constexpr int kTangents = 4;

template <class V>
struct DualT {
  V v;
  V d[kTangents];
  DualT() : v(0) {
    for (int k = 0; k < kTangents; k++) d[k] = V(0);
  }
  DualT(double a) : v(V(a)) {  // NOLINT: implicit, as R(0.5) and mixed arithmetic need
    for (int k = 0; k < kTangents; k++) d[k] = V(0);
  }
  explicit operator double() const { return double(v); }
  explicit operator float() const { return float(v); }
  explicit operator int() const { return int(v); }
  explicit operator uint32_t() const { return uint32_t(v); }
  explicit operator int64_t() const { return int64_t(v); }
};
using Dual = DualT<double>;
using Dual32 = DualT<float>;

template <class V>
inline DualT<V> operator+(const DualT<V>& a, const DualT<V>& b) {
  DualT<V> r;
  r.v = a.v + b.v;
  for (int k = 0; k < kTangents; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <class V>
inline DualT<V> operator-(const DualT<V>& a, const DualT<V>& b) {
  DualT<V> r;
  r.v = a.v - b.v;
  for (int k = 0; k < kTangents; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <class V>
inline DualT<V> operator-(const DualT<V>& a) {
  DualT<V> r;
  r.v = -a.v;
  for (int k = 0; k < kTangents; k++) r.d[k] = -a.d[k];
  return r;
}
template <class V>
inline DualT<V> operator*(const DualT<V>& a, const DualT<V>& b) {
  DualT<V> r;
  r.v = a.v * b.v;
  for (int k = 0; k < kTangents; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <class V>
inline DualT<V> operator/(const DualT<V>& a, const DualT<V>& b) {
  const V q = a.v / b.v;
  DualT<V> r;
  r.v = q;
  for (int k = 0; k < kTangents; k++) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
  return r;
}
template <class V> inline DualT<V>& operator+=(DualT<V>& a, const DualT<V>& b) { return a = a + b; }
template <class V> inline DualT<V>& operator-=(DualT<V>& a, const DualT<V>& b) { return a = a - b; }
template <class V> inline DualT<V>& operator*=(DualT<V>& a, const DualT<V>& b) { return a = a * b; }
template <class V> inline DualT<V>& operator/=(DualT<V>& a, const DualT<V>& b) { return a = a / b; }
template <class V> inline bool operator<(const DualT<V>& a, const DualT<V>& b) { return a.v < b.v; }
template <class V> inline bool operator>(const DualT<V>& a, const DualT<V>& b) { return a.v > b.v; }
template <class V> inline bool operator<=(const DualT<V>& a, const DualT<V>& b) { return a.v <= b.v; }
template <class V> inline bool operator>=(const DualT<V>& a, const DualT<V>& b) { return a.v >= b.v; }
template <class V> inline bool operator==(const DualT<V>& a, const DualT<V>& b) { return a.v == b.v; }
template <class V> inline bool operator!=(const DualT<V>& a, const DualT<V>& b) { return a.v != b.v; }

// f(a) with f'(a) = fp
template <class V>
inline DualT<V> chain(const DualT<V>& a, V f, V fp) {
  DualT<V> r;
  r.v = f;
  for (int k = 0; k < kTangents; k++) r.d[k] = fp * a.d[k];
  return r;
}
template <class V> inline DualT<V> o_fabs(const DualT<V>& a) { return a.v < V(0) ? -a : a; }
template <class V>
inline DualT<V> o_sqrt(const DualT<V>& a) {
  const V s = std::sqrt(a.v);
  return chain(a, s, s > V(0) ? V(0.5) / s : V(0));
}
template <class V> inline DualT<V> o_log(const DualT<V>& a) { return chain(a, V(std::log(a.v)), V(1) / a.v); }
template <class V>
inline DualT<V> o_exp(const DualT<V>& a) {
  const V e = std::exp(a.v);
  return chain(a, e, e);
}
template <class V> inline DualT<V> o_floor(const DualT<V>& a) { return DualT<V>(double(std::floor(a.v))); }
template <class V>
inline DualT<V> o_pow(const DualT<V>& a, const DualT<V>& b) {
  // the oracle raises to constant exponents only (J^(1/3)); b's tangent is dropped
  const V pw = std::pow(a.v, b.v);
  return chain(a, pw, a.v != V(0) ? b.v * pw / a.v : V(0));
}
// the noise hash's sine: the FP32 oracle's correctly rounded float sine for
// Dual32, the FP64 sine for Dual
inline Dual hash_sin(const Dual& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
inline Dual32 hash_sin(const Dual32& a) {
  return chain(a, float(std::sin(double(a.v))), float(std::cos(double(a.v))));
}

}  // namespace oracle
