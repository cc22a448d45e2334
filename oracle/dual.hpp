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

struct Dual {
  double v;
  double d[kTangents];
  Dual() : v(0.0) {
    for (int k = 0; k < kTangents; k++) d[k] = 0.0;
  }
  Dual(double a) : v(a) {  // NOLINT: implicit, as R(0.5) and mixed arithmetic need
    for (int k = 0; k < kTangents; k++) d[k] = 0.0;
  }
  explicit operator double() const { return v; }
  explicit operator float() const { return float(v); }
  explicit operator int() const { return int(v); }
  explicit operator uint32_t() const { return uint32_t(v); }
  explicit operator int64_t() const { return int64_t(v); }
};

inline Dual operator+(const Dual& a, const Dual& b) {
  Dual r(a.v + b.v);
  for (int k = 0; k < kTangents; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
  Dual r(a.v - b.v);
  for (int k = 0; k < kTangents; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
inline Dual operator-(const Dual& a) {
  Dual r(-a.v);
  for (int k = 0; k < kTangents; k++) r.d[k] = -a.d[k];
  return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
  Dual r(a.v * b.v);
  for (int k = 0; k < kTangents; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  Dual r(q);
  for (int k = 0; k < kTangents; k++) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
  return r;
}
inline Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
inline Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }
inline Dual& operator*=(Dual& a, const Dual& b) { return a = a * b; }
inline Dual& operator/=(Dual& a, const Dual& b) { return a = a / b; }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

// f(a) with f'(a) = fp
inline Dual chain(const Dual& a, double f, double fp) {
  Dual r(f);
  for (int k = 0; k < kTangents; k++) r.d[k] = fp * a.d[k];
  return r;
}
inline Dual o_fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
inline Dual o_sqrt(const Dual& a) {
  const double s = std::sqrt(a.v);
  return chain(a, s, s > 0.0 ? 0.5 / s : 0.0);
}
inline Dual o_log(const Dual& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
inline Dual o_exp(const Dual& a) {
  const double e = std::exp(a.v);
  return chain(a, e, e);
}
inline Dual o_floor(const Dual& a) { return Dual(std::floor(a.v)); }
inline Dual o_pow(const Dual& a, const Dual& b) {
  // the oracle raises to constant exponents only (J^(1/3)); b's tangent is dropped
  const double p = std::pow(a.v, b.v);
  return chain(a, p, a.v != 0.0 ? b.v * p / a.v : 0.0);
}
inline Dual hash_sin(const Dual& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }

}  // namespace oracle
