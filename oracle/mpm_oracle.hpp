// mpm_oracle.hpp — CPU oracle for the MPM substep.  TEST INFRASTRUCTURE ONLY.
//
// This is the checker, not the product: only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it.  The HIP library never links it.
//
// It restates, function by function, the semantics of the reference's GLSL
// substep so that the GPU kernels can be compared with it on identical inputs.
// The template parameter R selects the arithmetic: R = float gives the
// reference-semantics FP32 oracle, R = double the FP64 twin used to state the
// error of both FP32 codes.  Every function cites the reference lines it follows.
//
// Parity status of each part (what pins it; see the tests named in brackets):
//   morton/cell keys ............ brute-force bit interleave       [test_oracle_kats]
//   binning (stable, END offsets) numpy stable argsort             [test_oracle_kats]
//   bspline2/bspline2g .......... KATs 1-1, 1-2 of tests/test_mpm.cpp + closed form
//   P2G ......................... KAT 2-1, APIC affine reproduction, brute force
//   grid update / walls / forces  KATs 3-1, 3-2; collider algebra
//   G2P ......................... KATs 4-1, 6-2, 7-1, 7-2, 7-3, 7-4, 8-1
//   SVD / Hencky / return maps .. KATs 5-1, 5-2; numpy LAPACK SVD; yield invariants
//   Degenerate SVD branches (sign(0) stall, sigma <= 1e-8 unit vectors) and the
//   Jacobian reset: results worked by hand from the GLSL text on inputs that
//   reach them (test_oracle_kats, and through G2P on the GPU)
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "dual.hpp"

namespace oracle {

// Math through overloadable names: float and double take the std functions
// (the float overloads, so the FP32 oracle rounds in FP32); dual.hpp adds the
// forward-mode versions.
template <class R> inline R o_fabs(R x) { return std::fabs(x); }
template <class R> inline R o_sqrt(R x) { return std::sqrt(x); }
template <class R> inline R o_log(R x) { return std::log(x); }
template <class R> inline R o_exp(R x) { return std::exp(x); }
template <class R> inline R o_floor(R x) { return std::floor(x); }
template <class R> inline R o_pow(R x, R y) { return std::pow(x, y); }

// ---------------------------------------------------------------------------
// Small column-major 3x3 algebra with GLSL conventions (m[col][row]).
// ---------------------------------------------------------------------------
template <class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R length(V3<R> a) { return o_sqrt(dot(a, a)); }

template <class R> struct M3 {
  V3<R> c[3];  // columns
  M3() {}
  static M3 diag(R a, R b, R d) {
    M3 m;
    m.c[0] = {a, 0, 0};
    m.c[1] = {0, b, 0};
    m.c[2] = {0, 0, d};
    return m;
  }
  static M3 ident() { return diag(1, 1, 1); }
  R& at(int col, int row) { return c[col][row]; }
  R at(int col, int row) const { return c[col][row]; }
};
template <class R> inline M3<R> operator+(const M3<R>& a, const M3<R>& b) {
  M3<R> m;
  for (int i = 0; i < 3; i++) m.c[i] = a.c[i] + b.c[i];
  return m;
}
template <class R> inline M3<R> operator-(const M3<R>& a, const M3<R>& b) {
  M3<R> m;
  for (int i = 0; i < 3; i++) m.c[i] = a.c[i] - b.c[i];
  return m;
}
template <class R> inline M3<R> operator*(R s, const M3<R>& a) {
  M3<R> m;
  for (int i = 0; i < 3; i++) m.c[i] = s * a.c[i];
  return m;
}
// (A*B)[col][row] = sum_k A[k][row] * B[col][k]
template <class R> inline M3<R> operator*(const M3<R>& A, const M3<R>& B) {
  M3<R> m;
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) {
      R s = A.at(0, row) * B.at(col, 0);
      s = s + A.at(1, row) * B.at(col, 1);
      s = s + A.at(2, row) * B.at(col, 2);
      m.at(col, row) = s;
    }
  return m;
}
// (M*v)[row] = sum_col M[col][row] * v[col]
template <class R> inline V3<R> operator*(const M3<R>& M, V3<R> v) {
  V3<R> r;
  for (int row = 0; row < 3; row++) r[row] = M.at(0, row) * v.x + M.at(1, row) * v.y + M.at(2, row) * v.z;
  return r;
}
template <class R> inline M3<R> transpose(const M3<R>& a) {
  M3<R> m;
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) m.at(col, row) = a.at(row, col);
  return m;
}
// outerProduct(a, b)[j][i] = a[i] * b[j]
template <class R> inline M3<R> outer(V3<R> a, V3<R> b) {
  M3<R> m;
  for (int j = 0; j < 3; j++) m.c[j] = a * b[j];
  return m;
}
// cofactor expansion along the first column (glm / GLSL determinant)
template <class R> inline R determinant(const M3<R>& m) {
  return m.at(0, 0) * (m.at(1, 1) * m.at(2, 2) - m.at(2, 1) * m.at(1, 2)) -
         m.at(1, 0) * (m.at(0, 1) * m.at(2, 2) - m.at(2, 1) * m.at(0, 2)) +
         m.at(2, 0) * (m.at(0, 1) * m.at(1, 2) - m.at(1, 1) * m.at(0, 2));
}
template <class R> inline R glsl_sign(R x) { return x > R(0) ? R(1) : (x < R(0) ? R(-1) : R(0)); }
template <class R> inline R glsl_clamp(R x, R lo, R hi) { return std::min(std::max(x, lo), hi); }

// ---------------------------------------------------------------------------
// Morton keys — shaders/mpm_common.glsl:161-195
// ---------------------------------------------------------------------------
inline uint32_t morton_expand(uint32_t v) {
  v = (v | (v << 16u)) & 0x030000FFu;
  v = (v | (v << 8u)) & 0x0300F00Fu;
  v = (v | (v << 4u)) & 0x030C30C3u;
  v = (v | (v << 2u)) & 0x09249249u;
  return v;
}
inline uint32_t morton_compact(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2u)) & 0x030C30C3u;
  v = (v | (v >> 4u)) & 0x0300F00Fu;
  v = (v | (v >> 8u)) & 0x030000FFu;
  v = (v | (v >> 16u)) & 0x000003FFu;
  return v;
}
inline uint32_t morton_encode(int x, int y, int z) {
  return morton_expand(uint32_t(x)) | (morton_expand(uint32_t(y)) << 1u) | (morton_expand(uint32_t(z)) << 2u);
}
inline void morton_decode(uint32_t code, int& x, int& y, int& z) {
  x = int(morton_compact(code));
  y = int(morton_compact(code >> 1u));
  z = int(morton_compact(code >> 2u));
}

// ---------------------------------------------------------------------------
// Parameters of one substep: the fields of MPMSimPC the MPM shaders read
// (src/core/MPMSimPC.h:14-73), plus the buffers they index.
// ---------------------------------------------------------------------------
struct Material {  // src/core/MaterialParams.h:16-38
  float mu, lambda, rho0;
  uint32_t model;
  float M_friction, q_cohesion, q_max, bulkK, fluidGamma, viscosity, hardening, xi, pad[4];
};
struct Collider {  // src/core/Collider.h:19-29
  uint32_t type;
  float px, py, pz, nx, ny, nz, radius, vx, vy, vz, restitution, friction, pad[3];
};
struct Force {  // src/core/Force.h:27-32
  uint16_t type, affect_mask;
  float data[7];
};

struct Params {
  float dt;
  float cell_size;
  uint32_t grid_res[3];
  float world_min[3];
  float world_max[3];
  float flip_ratio;
  uint32_t cube_res;  // Morton cube edge (src/core/Domain.h:25-30)
  const Material* materials;
  uint32_t material_count;
  const Collider* colliders;
  uint32_t collider_count;
  const Force* forces;
  uint32_t force_count;
  const float* sdf;  // dense Morton SDF (NC floats) or null
  bool corrected_svd;  // see svd3
  // dual form only: tangents of each material's (mu, lambda, M_friction,
  // q_cohesion), [material][parameter][tangent], or null
  const double* mat_tangent = nullptr;
};

// Material parameter `which` (0 mu, 1 lambda, 2 M_friction, 3 q_cohesion) of
// material `id` (already clamped) as R; the dual form attaches its tangents.
template <class R> inline R mat_param(const Params&, uint32_t, int, float value) { return R(value); }
template <class V>
inline DualT<V> mat_param_dual(const Params& p, uint32_t id, int which, float value) {
  DualT<V> r{double(value)};
  if (p.mat_tangent)
    for (int k = 0; k < kTangents; k++) r.d[k] = V(p.mat_tangent[(size_t(id) * 4 + size_t(which)) * kTangents + k]);
  return r;
}
template <> inline Dual mat_param<Dual>(const Params& p, uint32_t id, int which, float value) {
  return mat_param_dual<double>(p, id, which, value);
}
template <> inline Dual32 mat_param<Dual32>(const Params& p, uint32_t id, int which, float value) {
  return mat_param_dual<float>(p, id, which, value);
}

inline uint32_t total_cells(const Params& p) { return p.cube_res * p.cube_res * p.cube_res; }

// ---------------------------------------------------------------------------
// Quadratic B-spline and derivative — shaders/mpm_common.glsl:198-210
// ---------------------------------------------------------------------------
template <class R> inline R bspline2(R d) {
  R ad = o_fabs(d);
  if (ad < R(0.5)) return R(0.75) - ad * ad;
  if (ad < R(1.5)) {
    R t = R(1.5) - ad;
    return R(0.5) * t * t;
  }
  return R(0);
}
template <class R> inline R bspline2g(R d) {
  R ad = o_fabs(d);
  if (ad < R(0.5)) return R(-2) * d;
  if (ad < R(1.5)) return -glsl_sign(d) * (R(1.5) - ad);
  return R(0);
}

// ---------------------------------------------------------------------------
// Cell key of a position — shaders/mpm_common.glsl:190-195
// (clamp of the IEEE quotient to [0, gridRes-1], then truncation)
// ---------------------------------------------------------------------------
template <class R> inline uint32_t cell_id_from_pos(const Params& p, R x, R y, R z) {
  R h = R(p.cell_size);
  R lx = glsl_clamp((x - R(p.world_min[0])) / h, R(0), R(p.grid_res[0]) - R(1));
  R ly = glsl_clamp((y - R(p.world_min[1])) / h, R(0), R(p.grid_res[1]) - R(1));
  R lz = glsl_clamp((z - R(p.world_min[2])) / h, R(0), R(p.grid_res[2]) - R(1));
  // NaN compares false in both clamp bounds and would keep NaN; std::min/max
  // return the first argument then, so force it to zero explicitly.
  if (!(lx == lx)) lx = 0;
  if (!(ly == ly)) ly = 0;
  if (!(lz == lz)) lz = 0;
  return morton_expand(uint32_t(lx)) | (morton_expand(uint32_t(ly)) << 1u) | (morton_expand(uint32_t(lz)) << 2u);
}

// ---------------------------------------------------------------------------
// Binning: the stable counting sort with END offsets that
// tests/helpers/MPMHarness.cpp:132-178 builds on the CPU, which is the intended
// semantics of mpm_hash_count/hash_scan_*/mpm_hash_sort
// (shaders/mpm_hash_count.comp:12-13, mpm_hash_sort.comp:12-14).
// ---------------------------------------------------------------------------
template <class R>
void bin(const Params& p, const R* P4, uint32_t n, uint32_t* keys, uint32_t* sorted, uint32_t* cell_count,
         uint32_t* cell_end) {
  const uint32_t NC = total_cells(p);
  std::fill(cell_count, cell_count + NC, 0u);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t k = cell_id_from_pos<R>(p, P4[4 * i + 0], P4[4 * i + 1], P4[4 * i + 2]);
    keys[i] = k;
    cell_count[k]++;
  }
  uint32_t acc = 0;
  for (uint32_t c = 0; c < NC; c++) {
    acc += cell_count[c];
    cell_end[c] = acc;
  }
  std::vector<uint32_t> wp(NC);
  for (uint32_t c = 0; c < NC; c++) wp[c] = cell_end[c] - cell_count[c];
  for (uint32_t i = 0; i < n; i++) sorted[wp[keys[i]]++] = i;
}

// ---------------------------------------------------------------------------
// Jacobi eigen-solver and SVD — shaders/mpm_common.glsl:213-281
// ---------------------------------------------------------------------------
template <class R> inline void jacobi_eigen3(M3<R> A, V3<R>& D, M3<R>& V) {
  V = M3<R>::ident();
  for (int iter = 0; iter < 20; iter++) {
    R a01 = A.at(0, 1), a02 = A.at(0, 2), a12 = A.at(1, 2);
    if (std::max(o_fabs(a01), std::max(o_fabs(a02), o_fabs(a12))) < R(1e-8)) break;
    int p0, p1;
    R apq;
    if (o_fabs(a01) >= std::max(o_fabs(a02), o_fabs(a12))) {
      p0 = 0; p1 = 1; apq = a01;
    } else if (o_fabs(a02) >= o_fabs(a12)) {
      p0 = 0; p1 = 2; apq = a02;
    } else {
      p0 = 1; p1 = 2; apq = a12;
    }
    R app = A.at(p0, p0), aqq = A.at(p1, p1);
    R tau = (aqq - app) / (R(2) * apq);
    R t = glsl_sign(tau) / (o_fabs(tau) + o_sqrt(R(1) + tau * tau));
    R c = R(1) / o_sqrt(R(1) + t * t);  // inversesqrt
    R s = t * c;
    A.at(p0, p0) = app - t * apq;
    A.at(p1, p1) = aqq + t * apq;
    A.at(p0, p1) = R(0);
    A.at(p1, p0) = R(0);
    for (int r = 0; r < 3; r++) {
      if (r == p0 || r == p1) continue;
      R arp = A.at(r, p0), arq = A.at(r, p1);
      A.at(r, p0) = c * arp - s * arq;
      A.at(p0, r) = A.at(r, p0);
      A.at(r, p1) = s * arp + c * arq;
      A.at(p1, r) = A.at(r, p1);
    }
    for (int r = 0; r < 3; r++) {
      R vrp = V.at(r, p0), vrq = V.at(r, p1);
      V.at(r, p0) = c * vrp - s * vrq;
      V.at(r, p1) = s * vrp + c * vrq;
    }
  }
  D = {A.at(0, 0), A.at(1, 1), A.at(2, 2)};
}

// corrected = false: the GLSL result as stored (singular vectors in the rows of
// U and V).  corrected = true (an extension, not reference semantics): V is
// made a rotation before U is formed and both are transposed, so that
// F = U diag(sigma) V^T with U, V in SO(3) and U diag U^T, U diag V^T are the
// true spectral forms.
template <class R> inline void svd3(const M3<R>& F, M3<R>& U, V3<R>& sigma, M3<R>& V, bool corrected = false) {
  M3<R> FtF = transpose(F) * F;
  V3<R> sig2;
  jacobi_eigen3(FtF, sig2, V);
  sigma = {o_sqrt(std::max(sig2.x, R(0))), o_sqrt(std::max(sig2.y, R(0))), o_sqrt(std::max(sig2.z, R(0)))};
  for (int i = 0; i < 2; i++)
    for (int j = i + 1; j < 3; j++) {
      if (sigma[j] > sigma[i]) {
        std::swap(sigma[i], sigma[j]);
        for (int r = 0; r < 3; r++) std::swap(V.at(r, i), V.at(r, j));
      }
    }
  if (corrected && determinant(V) < R(0)) {  // keep V a rotation: flip the third singular vector
    V.at(0, 2) = -V.at(0, 2);
    V.at(1, 2) = -V.at(1, 2);
    V.at(2, 2) = -V.at(2, 2);
  }
  U = M3<R>();
  for (int i = 0; i < 3; i++) {
    V3<R> vi(V.at(0, i), V.at(1, i), V.at(2, i));
    if (sigma[i] > R(1e-8)) {
      V3<R> Fv = F * vi;
      U.at(0, i) = Fv[0] / sigma[i];
      U.at(1, i) = Fv[1] / sigma[i];
      U.at(2, i) = Fv[2] / sigma[i];
    } else {
      U.at(i, i) = R(1);
    }
  }
  if (determinant(U) < R(0)) {
    U.at(0, 2) = -U.at(0, 2);
    U.at(1, 2) = -U.at(1, 2);
    U.at(2, 2) = -U.at(2, 2);
    sigma[2] = -sigma[2];
  }
  if (corrected) {
    U = transpose(U);
    V = transpose(V);
  }
}

// Hencky Kirchhoff stress — shaders/mpm_common.glsl:284-293
template <class R> inline M3<R> hencky_stress(const M3<R>& F, R mu, R lam, bool corrected = false) {
  M3<R> U, V;
  V3<R> s;
  svd3(F, U, s, V, corrected);
  s = {std::max(o_fabs(s.x), R(1e-6)), std::max(o_fabs(s.y), R(1e-6)), std::max(o_fabs(s.z), R(1e-6))};
  V3<R> eps(o_log(s.x), o_log(s.y), o_log(s.z));
  R tr = eps.x + eps.y + eps.z;
  V3<R> kp = V3<R>(lam * tr, lam * tr, lam * tr) + (R(2) * mu) * eps;
  return U * M3<R>::diag(kp.x, kp.y, kp.z) * transpose(U);
}

// Von Mises projection — shaders/mpm_g2p.comp:19-38
template <class R> inline M3<R> project_von_mises(M3<R> F, R mu, R lam, R q_max, bool corrected = false) {
  (void)lam;
  M3<R> U, V;
  V3<R> s;
  svd3(F, U, s, V, corrected);
  s = {std::max(o_fabs(s.x), R(1e-6)), std::max(o_fabs(s.y), R(1e-6)), std::max(o_fabs(s.z), R(1e-6))};
  V3<R> eps(o_log(s.x), o_log(s.y), o_log(s.z));
  R tr = eps.x + eps.y + eps.z;
  V3<R> dev = eps - (tr / R(3)) * V3<R>(1, 1, 1);
  R nd = length(dev);
  R e_mu = R(2) * mu;
  R delta = nd - q_max / e_mu;
  if (delta > R(0) && nd > R(1e-8)) {
    dev = dev - (delta / nd) * dev;
    eps = dev + (tr / R(3)) * V3<R>(1, 1, 1);
    s = {o_exp(eps.x), o_exp(eps.y), o_exp(eps.z)};
    F = U * M3<R>::diag(s.x, s.y, s.z) * transpose(V);
  }
  return F;
}

// Drucker-Prager projection — shaders/mpm_g2p.comp:41-68
template <class R> inline M3<R> project_drucker_prager(const M3<R>& F, R mu, R lam, R M, R q_c,
                                                      bool corrected = false) {
  M3<R> U, V;
  V3<R> s;
  svd3(F, U, s, V, corrected);
  s = {std::max(o_fabs(s.x), R(1e-6)), std::max(o_fabs(s.y), R(1e-6)), std::max(o_fabs(s.z), R(1e-6))};
  V3<R> eps(o_log(s.x), o_log(s.y), o_log(s.z));
  R tr = eps.x + eps.y + eps.z;
  V3<R> dev = eps - (tr / R(3)) * V3<R>(1, 1, 1);
  R nd = length(dev);
  R K = lam + R(2) * mu / R(3);
  R pr = -K * tr;
  R q_yield = M * pr + q_c;
  if (q_yield <= R(1e-8)) {
    R p_proj = -q_c / std::max(M, R(1e-8));
    tr = -p_proj / K;
    eps = (tr / R(3)) * V3<R>(1, 1, 1);
  } else {
    R delta = nd - q_yield / (R(2) * mu);
    if (delta > R(0) && nd > R(1e-8)) {
      dev = dev - (delta / nd) * dev;
      eps = dev + (tr / R(3)) * V3<R>(1, 1, 1);
    }
  }
  s = {o_exp(eps.x), o_exp(eps.y), o_exp(eps.z)};
  return U * M3<R>::diag(s.x, s.y, s.z) * transpose(V);
}

// ---------------------------------------------------------------------------
// Simulation state: the eight float4 arrays (src/engine/MPMEngine.cpp:46-53)
// and the dense grid (gridMom 2*NC vec4, gridMass NC).
// ---------------------------------------------------------------------------
template <class R> struct State {
  uint32_t n = 0;
  std::vector<R> f[8];  // P, v, F0, F1, F2, B0, B1, B2 (4 per particle)
  std::vector<uint32_t> matid;  // v.w bits, carried as an integer
  std::vector<uint32_t> keys, sorted, cell_count, cell_end;
  std::vector<R> grid_mom;   // 2*NC*4
  std::vector<R> grid_mass;  // NC

  void resize(uint32_t n_, uint32_t NC) {
    n = n_;
    for (auto& a : f) a.assign(size_t(n) * 4, R(0));
    matid.assign(n, 0u);
    keys.assign(n, 0u);
    sorted.assign(n, 0u);
    cell_count.assign(NC, 0u);
    cell_end.assign(NC, 0u);
    grid_mom.assign(size_t(NC) * 8, R(0));
    grid_mass.assign(NC, R(0));
  }
  M3<R> read_mat(int f0, uint32_t p) const {
    M3<R> m;
    for (int c = 0; c < 3; c++) m.c[c] = {f[f0 + c][4 * p + 0], f[f0 + c][4 * p + 1], f[f0 + c][4 * p + 2]};
    return m;
  }
  // readStressW — shaders/mpm_common.glsl:104-113
  M3<R> read_tau(uint32_t p) const {
    R txx = f[2][4 * p + 3], tyy = f[3][4 * p + 3], tzz = f[4][4 * p + 3];
    R txy = f[5][4 * p + 3], txz = f[6][4 * p + 3], tyz = f[7][4 * p + 3];
    M3<R> m;
    m.c[0] = {txx, txy, txz};
    m.c[1] = {txy, tyy, tyz};
    m.c[2] = {txz, tyz, tzz};
    return m;
  }
};

template <class R> inline const Material& material_of(const Params& p, uint32_t id) {
  // shaders/mpm_g2p.comp:128-130: clamp to materialCount-1
  if (id > p.material_count - 1u) id = p.material_count - 1u;
  return p.materials[id];
}

// ---------------------------------------------------------------------------
// P2G node-centric gather — shaders/mpm_p2g.comp:17-87 (all NC nodes; nodes
// whose 27-cell neighbourhood is empty stay at zero, as the reference writes).
// ---------------------------------------------------------------------------
template <class R> void p2g(const Params& p, State<R>& s) {
  const uint32_t NC = total_cells(p);
  const R h = R(p.cell_size);
  const R dt = R(p.dt);
  std::fill(s.grid_mom.begin(), s.grid_mom.end(), R(0));  // mpm_zero_grid.comp:16-18
  std::fill(s.grid_mass.begin(), s.grid_mass.end(), R(0));
  // nodes reachable from an occupied cell
  std::vector<uint8_t> touched(NC, 0);
  const int gx = int(p.grid_res[0]), gy = int(p.grid_res[1]), gz = int(p.grid_res[2]);
  const int cr = int(p.cube_res);
  for (uint32_t c = 0; c < NC; c++) {
    if (!s.cell_count[c]) continue;
    int cx, cy, cz;
    morton_decode(c, cx, cy, cz);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int nx = cx + dx, ny = cy + dy, nz = cz + dz;
          if (nx < 0 || ny < 0 || nz < 0 || nx >= cr || ny >= cr || nz >= cr) continue;
          touched[morton_encode(nx, ny, nz)] = 1;
        }
  }
  const bool apic = p.flip_ratio < -0.5f;
  const R D_inv = R(4) / (h * h);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t gid = 0; gid < int64_t(NC); gid++) {
    if (!touched[gid]) continue;
    int ix, iy, iz;
    morton_decode(uint32_t(gid), ix, iy, iz);
    V3<R> xi((R(ix) + R(0.5)) * h + R(p.world_min[0]), (R(iy) + R(0.5)) * h + R(p.world_min[1]),
             (R(iz) + R(0.5)) * h + R(p.world_min[2]));
    R tm = 0;
    V3<R> tmom;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int cx = ix + dx, cy = iy + dy, cz = iz + dz;
          if (cx < 0 || cy < 0 || cz < 0 || cx >= gx || cy >= gy || cz >= gz) continue;
          uint32_t cid = morton_encode(cx, cy, cz);
          uint32_t cnt = s.cell_count[cid];
          uint32_t start = s.cell_end[cid] - cnt;
          for (uint32_t k = start; k < start + cnt; k++) {
            uint32_t q = s.sorted[k];
            V3<R> xp(s.f[0][4 * q + 0], s.f[0][4 * q + 1], s.f[0][4 * q + 2]);
            R Vp = s.f[0][4 * q + 3];
            V3<R> vp(s.f[1][4 * q + 0], s.f[1][4 * q + 1], s.f[1][4 * q + 2]);
            R rho0 = R(material_of<R>(p, s.matid[q]).rho0);
            R mp = rho0 * Vp;
            V3<R> d = (xp - xi) / h;
            R wx = bspline2(d.x), wy = bspline2(d.y), wz = bspline2(d.z);
            R w = wx * wy * wz;
            if (w < R(1e-12)) continue;
            V3<R> dw = V3<R>(bspline2g(d.x) * wy * wz, wx * bspline2g(d.y) * wz, wx * wy * bspline2g(d.z)) / h;
            M3<R> Bp;
            if (apic) Bp = s.read_mat(5, q);
            V3<R> v_eff = vp + D_inv * (Bp * (xi - xp));
            M3<R> tau = s.read_tau(q);
            tm = tm + w * mp;
            tmom = tmom + (w * mp) * v_eff;
            tmom = tmom - (dt * Vp) * (tau * dw);
          }
        }
    s.grid_mom[4 * gid + 0] = tmom.x;
    s.grid_mom[4 * gid + 1] = tmom.y;
    s.grid_mom[4 * gid + 2] = tmom.z;
    s.grid_mom[4 * gid + 3] = 0;
    s.grid_mass[gid] = tm;
  }
}

// ---------------------------------------------------------------------------
// Analytic collider SDFs and BC — shaders/collider_common.glsl:22-85
// ---------------------------------------------------------------------------
template <class R> inline V3<R> normalize3(V3<R> v) { return v / length(v); }
template <class R> inline R plane_sdf(V3<R> x, V3<R> c, V3<R> n, V3<R>& on) {
  on = normalize3(n);
  return dot(x - c, on);
}
template <class R> inline R sphere_sdf(V3<R> x, V3<R> c, R r, V3<R>& on) {
  V3<R> d = x - c;
  R dist = length(d);
  on = dist > R(1e-8) ? d / dist : V3<R>(0, 1, 0);
  return dist - r;
}
template <class R> inline R box_sdf(V3<R> x, V3<R> c, V3<R> he, V3<R>& on) {
  V3<R> rel = x - c;
  V3<R> q(o_fabs(rel.x) - he.x, o_fabs(rel.y) - he.y, o_fabs(rel.z) - he.z);
  V3<R> qp(std::max(q.x, R(0)), std::max(q.y, R(0)), std::max(q.z, R(0)));
  R sdf = length(qp) + std::min(std::max(q.x, std::max(q.y, q.z)), R(0));
  if (sdf < R(0)) {
    V3<R> diff(he.x - o_fabs(rel.x), he.y - o_fabs(rel.y), he.z - o_fabs(rel.z));
    if (diff.x < diff.y && diff.x < diff.z)
      on = V3<R>(glsl_sign(rel.x + R(1e-8)), 0, 0);
    else if (diff.y < diff.z)
      on = V3<R>(0, glsl_sign(rel.y + R(1e-8)), 0);
    else
      on = V3<R>(0, 0, glsl_sign(rel.z + R(1e-8)));
  } else {
    R ql = length(qp);
    V3<R> sg(glsl_sign(rel.x + R(1e-8)), glsl_sign(rel.y + R(1e-8)), glsl_sign(rel.z + R(1e-8)));
    on = ql > R(1e-8) ? normalize3(V3<R>(qp.x * sg.x, qp.y * sg.y, qp.z * sg.z)) : normalize3(sg);
  }
  return sdf;
}
template <class R> inline R capsule_sdf(V3<R> x, V3<R> c, V3<R> axis, R r, V3<R>& on) {
  R al = length(axis);
  V3<R> ad = al > R(1e-8) ? axis / al : V3<R>(0, 1, 0);
  R t = glsl_clamp(dot(x - c, ad), R(0), al);
  V3<R> cl = c + t * ad;
  V3<R> d = x - cl;
  R dist = length(d);
  on = dist > R(1e-8) ? d / dist : V3<R>(0, 1, 0);
  return dist - r;
}
template <class R> inline void apply_collider_bc(V3<R>& v, V3<R> n, V3<R> vc, R rest, R fric) {
  V3<R> vr = v - vc;
  R vn = dot(vr, n);
  if (vn < R(0)) {
    V3<R> vnv = vn * n;
    V3<R> vt = vr - vnv;
    R vtl = length(vt);
    if (vtl > R(1e-8)) vt = std::max(R(0), R(1) - fric * o_fabs(vn) / vtl) * vt;
    v = vc + vt - rest * vnv;
  }
}

// ---------------------------------------------------------------------------
// Procedural noise of the turbulence and noise forces — shaders/force_noise.glsl
// (hash 7-12, value noise 14-37, fBm 40-50, curl 54-82), in precision R, with
// GLSL's fract(x) = x - floor(x) and mix(a, b, t) = a + (b - a) t.
// The hash's sin in FP32 is the correctly rounded sine: sin in FP64 rounded
// once (the HIP kernels do the same; GLSL leaves sin's accuracy to the driver).
inline float hash_sin(float a) { return float(std::sin(double(a))); }
inline double hash_sin(double a) { return std::sin(a); }
template <class R> R gl_fract(R x) { return x - o_floor(x); }
template <class R> R gl_mix(R a, R b, R t) { return a + (b - a) * t; }
template <class R> V3<R> noise_hash33(const V3<R>& p) {
  const R a = p.x * R(127.1f) + p.y * R(311.7f) + p.z * R(74.7f);
  const R b = p.x * R(269.5f) + p.y * R(183.3f) + p.z * R(246.1f);
  const R c = p.x * R(113.5f) + p.y * R(271.9f) + p.z * R(124.6f);
  const R k = R(43758.5453123f);
  return V3<R>(R(-1) + R(2) * gl_fract(hash_sin(a) * k), R(-1) + R(2) * gl_fract(hash_sin(b) * k),
               R(-1) + R(2) * gl_fract(hash_sin(c) * k));
}
template <class R> R noise_value3(const V3<R>& p) {
  const V3<R> i(o_floor(p.x), o_floor(p.y), o_floor(p.z));
  const V3<R> f(gl_fract(p.x), gl_fract(p.y), gl_fract(p.z));
  const V3<R> u(f.x * f.x * (R(3) - R(2) * f.x), f.y * f.y * (R(3) - R(2) * f.y), f.z * f.z * (R(3) - R(2) * f.z));
  R n[8];
  for (int c = 0; c < 8; c++) {
    const V3<R> o(R(c & 1), R((c >> 1) & 1), R((c >> 2) & 1));
    const V3<R> g = noise_hash33(V3<R>(i.x + o.x, i.y + o.y, i.z + o.z));
    n[c] = g.x * (f.x - o.x) + g.y * (f.y - o.y) + g.z * (f.z - o.z);
  }
  const R x00 = gl_mix(n[0], n[1], u.x), x10 = gl_mix(n[2], n[3], u.x);
  const R x01 = gl_mix(n[4], n[5], u.x), x11 = gl_mix(n[6], n[7], u.x);
  return gl_mix(gl_mix(x00, x10, u.y), gl_mix(x01, x11, u.y), u.z);
}
template <class R> R noise_fbm3(const V3<R>& p, uint32_t octaves) {
  R sum = 0, amp = R(0.5), freq = 1;
  const uint32_t n = octaves > 1u ? octaves : 1u;
  for (uint32_t i = 0; i < n; i++) {
    sum += amp * noise_value3(V3<R>(p.x * freq, p.y * freq, p.z * freq));
    freq *= R(2);
    amp *= R(0.5);
  }
  return sum;
}
template <class R> V3<R> noise_curl3(const V3<R>& p, uint32_t oct) {
  const R e = R(0.1f), ox = R(19.1f), oy = R(47.3f), oz = R(71.7f);
  auto F = [&](R x, R y, R z) { return noise_fbm3(V3<R>(x, y, z), oct); };
  const R px1 = F(p.x + ox, p.y + e, p.z), px2 = F(p.x + ox, p.y - e, p.z);
  const R py3 = F(p.x + e, p.y + oy, p.z), py4 = F(p.x - e, p.y + oy, p.z);
  const R py1 = F(p.x, p.y + oy, p.z + e), py2 = F(p.x, p.y + oy, p.z - e);
  const R pz3 = F(p.x, p.y + e, p.z + oz), pz4 = F(p.x, p.y - e, p.z + oz);
  const R px3 = F(p.x + ox, p.y, p.z + e), px4 = F(p.x + ox, p.y, p.z - e);
  const R pz1 = F(p.x + e, p.y, p.z + oz), pz2 = F(p.x - e, p.y, p.z + oz);
  const R s = R(2) * e;
  return V3<R>((pz3 - pz4) / s - (py1 - py2) / s, (px3 - px4) / s - (pz1 - pz2) / s, (py3 - py4) / s - (px1 - px2) / s);
}

// Grid update — shaders/mpm_grid_update.comp:17-101, with the force block the
// reference generates (src/core/ForceShaderCompiler.cpp:72-93, Force.h:136-256).
// ---------------------------------------------------------------------------
template <class R> void grid_update(const Params& p, State<R>& s) {
  const uint32_t NC = total_cells(p);
  const R h = R(p.cell_size);
  const R dt = R(p.dt);
#pragma omp parallel for schedule(static, 4096)
  for (int64_t gid = 0; gid < int64_t(NC); gid++) {
    R mass = s.grid_mass[gid];
    if (mass < R(1e-10)) continue;
    V3<R> mom(s.grid_mom[4 * gid + 0], s.grid_mom[4 * gid + 1], s.grid_mom[4 * gid + 2]);
    V3<R> v = mom / mass;
    {
      R sp = length(v);
      R vmax = R(15);
      if (sp > vmax) v = v * (vmax / sp);
    }
    size_t o = 4 * (size_t(NC) + size_t(gid));
    s.grid_mom[o + 0] = v.x;
    s.grid_mom[o + 1] = v.y;
    s.grid_mom[o + 2] = v.z;
    s.grid_mom[o + 3] = 0;
    int ix, iy, iz;
    morton_decode(uint32_t(gid), ix, iy, iz);
    V3<R> fp((R(ix) + R(0.5)) * h + R(p.world_min[0]), (R(iy) + R(0.5)) * h + R(p.world_min[1]),
             (R(iz) + R(0.5)) * h + R(p.world_min[2]));
    {
      V3<R> fv = v;
      const uint32_t type_flag = 0u;
      for (uint32_t fi = 0; fi < p.force_count; fi++) {
        const Force& f = p.forces[fi];
        uint32_t mask = f.affect_mask;
        if (mask != 0u && ((mask >> type_flag) & 1u) == 0u) continue;
        if (f.type == 0u || f.type == 1u) {  // gravity, constant wind: v += dir*strength*dt
          V3<R> dir(R(f.data[0]), R(f.data[1]), R(f.data[2]));
          fv = fv + (dir * R(f.data[3])) * dt;
        } else if (f.type == 2u) {  // turbulence: v += curl(fBm)(pos freq + seed) strength dt
          uint32_t oct;
          std::memcpy(&oct, &f.data[2], 4);
          const R fr = R(f.data[1]), sd = R(f.data[3]);
          const V3<R> c = noise_curl3(V3<R>(fp.x * fr + sd, fp.y * fr + sd, fp.z * fr + sd), oct);
          fv = fv + (c * R(f.data[0])) * dt;
        } else if (f.type == 3u) {  // noise: v += dir (fBm(pos freq + seed) strength dt)
          uint32_t oct;
          std::memcpy(&oct, &f.data[5], 4);
          const R fr = R(f.data[4]), sd = R(f.data[6]);
          const R n = noise_fbm3(V3<R>(fp.x * fr + sd, fp.y * fr + sd, fp.z * fr + sd), oct);
          V3<R> dir(R(f.data[0]), R(f.data[1]), R(f.data[2]));
          fv = fv + dir * ((n * R(f.data[3])) * dt);
        }
      }
      v = fv;
    }
    int Nx = int(p.grid_res[0]) - 1, Ny = int(p.grid_res[1]) - 1, Nz = int(p.grid_res[2]) - 1;
    if (ix <= 1 && v.x < R(0)) v.x = 0;
    if (ix >= Nx - 1 && v.x > R(0)) v.x = 0;
    if (iy <= 1 && v.y < R(0)) v.y = 0;
    if (iy >= Ny - 1 && v.y > R(0)) v.y = 0;
    if (iz <= 1 && v.z < R(0)) v.z = 0;
    if (iz >= Nz - 1 && v.z > R(0)) v.z = 0;
    for (uint32_t ci = 0; ci < p.collider_count; ci++) {
      const Collider& c = p.colliders[ci];
      V3<R> cpos(c.px, c.py, c.pz), cn(c.nx, c.ny, c.nz), cv(c.vx, c.vy, c.vz), on;
      R sdf;
      if (c.type == 1u)
        sdf = sphere_sdf(fp, cpos, R(c.radius), on);
      else if (c.type == 2u)
        sdf = box_sdf(fp, cpos, cn, on);
      else if (c.type == 3u)
        sdf = capsule_sdf(fp, cpos, cn, R(c.radius), on);
      else
        sdf = plane_sdf(fp, cpos, cn, on);
      if (sdf <= R(0)) apply_collider_bc(v, on, cv, R(c.restitution), R(c.friction));
    }
    s.grid_mom[4 * gid + 0] = v.x;
    s.grid_mom[4 * gid + 1] = v.y;
    s.grid_mom[4 * gid + 2] = v.z;
    s.grid_mom[4 * gid + 3] = 0;
  }
}

// ---------------------------------------------------------------------------
// Dense-SDF node BC — shaders/mpm_nanovdb_bc.comp:23-54
// ---------------------------------------------------------------------------
template <class R> inline R sample_sdf(const Params& p, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= int(p.grid_res[0]) || y >= int(p.grid_res[1]) || z >= int(p.grid_res[2]))
    return R(1e6);
  return R(p.sdf[morton_encode(x, y, z)]);
}
template <class R> void sdf_bc(const Params& p, State<R>& s) {
  if (!p.sdf) return;
  const uint32_t NC = total_cells(p);
  const R h = R(p.cell_size);
#pragma omp parallel for schedule(static, 4096)
  for (int64_t gid = 0; gid < int64_t(NC); gid++) {
    if (s.grid_mass[gid] < R(1e-10)) continue;
    int x, y, z;
    morton_decode(uint32_t(gid), x, y, z);
    R sd = sample_sdf<R>(p, x, y, z);
    if (sd < R(0)) {
      for (int k = 0; k < 4; k++) s.grid_mom[4 * gid + k] = 0;
    } else if (sd < h * R(1.5)) {
      V3<R> v(s.grid_mom[4 * gid + 0], s.grid_mom[4 * gid + 1], s.grid_mom[4 * gid + 2]);
      R sx = sample_sdf<R>(p, x + 1, y, z) - sample_sdf<R>(p, x - 1, y, z);
      R sy = sample_sdf<R>(p, x, y + 1, z) - sample_sdf<R>(p, x, y - 1, z);
      R sz = sample_sdf<R>(p, x, y, z + 1) - sample_sdf<R>(p, x, y, z - 1);
      V3<R> gn(sx, sy, sz);
      R gl = length(gn);
      if (gl > R(1e-6)) {
        V3<R> n = gn / gl;
        R vn = dot(v, n);
        if (vn < R(0)) v = v - vn * n;
      }
      s.grid_mom[4 * gid + 0] = v.x;
      s.grid_mom[4 * gid + 1] = v.y;
      s.grid_mom[4 * gid + 2] = v.z;
      s.grid_mom[4 * gid + 3] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// G2P — shaders/mpm_g2p.comp:70-272
// ---------------------------------------------------------------------------
template <class R> void g2p(const Params& p, State<R>& s) {
  const uint32_t NC = total_cells(p);
  const R h = R(p.cell_size);
  const R dt = R(p.dt);
  const V3<R> wmin(R(p.world_min[0]), R(p.world_min[1]), R(p.world_min[2]));
  const V3<R> wmax(R(p.world_max[0]), R(p.world_max[1]), R(p.world_max[2]));
  const int gres[3] = {int(p.grid_res[0]), int(p.grid_res[1]), int(p.grid_res[2])};
  const bool flip = p.flip_ratio > 0.01f;
#pragma omp parallel for schedule(static, 256)
  for (int64_t pi = 0; pi < int64_t(s.n); pi++) {
    const uint32_t q = uint32_t(pi);
    V3<R> xp(s.f[0][4 * q + 0], s.f[0][4 * q + 1], s.f[0][4 * q + 2]);
    R Vp = s.f[0][4 * q + 3];
    V3<R> rel = (xp - wmin) / h;
    int base[3] = {int(o_floor(rel.x)) - 1, int(o_floor(rel.y)) - 1, int(o_floor(rel.z)) - 1};
    V3<R> vpic, vpic_old;
    M3<R> vgrad, Bn;
    for (int iz = 0; iz < 3; iz++)
      for (int iy = 0; iy < 3; iy++)
        for (int ix = 0; ix < 3; ix++) {
          int n[3] = {base[0] + ix, base[1] + iy, base[2] + iz};
          if (n[0] < 0 || n[1] < 0 || n[2] < 0 || n[0] >= gres[0] || n[1] >= gres[1] || n[2] >= gres[2]) continue;
          uint32_t nid = morton_encode(n[0], n[1], n[2]);
          R mass = s.grid_mass[nid];
          if (mass < R(1e-10)) continue;
          V3<R> xi = V3<R>((R(n[0]) + R(0.5)) * h, (R(n[1]) + R(0.5)) * h, (R(n[2]) + R(0.5)) * h) + wmin;
          V3<R> d = (xp - xi) / h;
          R wx = bspline2(d.x), wy = bspline2(d.y), wz = bspline2(d.z);
          R w = wx * wy * wz;
          if (w < R(1e-12)) continue;
          V3<R> dw = V3<R>(bspline2g(d.x) * wy * wz, wx * bspline2g(d.y) * wz, wx * wy * bspline2g(d.z)) / h;
          V3<R> vi(s.grid_mom[4 * nid + 0], s.grid_mom[4 * nid + 1], s.grid_mom[4 * nid + 2]);
          vpic = vpic + w * vi;
          vgrad = vgrad + outer(vi, dw);
          Bn = Bn + w * outer(vi, xi - xp);
          if (flip) {
            size_t o = 4 * (size_t(nid) + NC);
            vpic_old = vpic_old + w * V3<R>(s.grid_mom[o + 0], s.grid_mom[o + 1], s.grid_mom[o + 2]);
          }
        }
    V3<R> vold(s.f[1][4 * q + 0], s.f[1][4 * q + 1], s.f[1][4 * q + 2]);
    const Material& mat = material_of<R>(p, s.matid[q]);
    const uint32_t mid = std::min(s.matid[q], p.material_count - 1u);
    const bool cs = p.corrected_svd;
    R mu = mat_param<R>(p, mid, 0, mat.mu), lam = mat_param<R>(p, mid, 1, mat.lambda);
    const R Mf = mat_param<R>(p, mid, 2, mat.M_friction), qc = mat_param<R>(p, mid, 3, mat.q_cohesion);
    M3<R> F = s.read_mat(2, q);
    M3<R> Fn = (M3<R>::ident() + dt * vgrad) * F;
    M3<R> tau;
    if (mat.model == 4u) {  // FLUID
      R J = determinant(Fn);
      J = std::max(J, R(1e-4));
      Fn = o_pow(J, R(1) / R(3)) * M3<R>::ident();
      tau = (R(mat.bulkK) * (J - R(1))) * M3<R>::ident();
    } else if (mat.model == 1u) {  // VON_MISES
      Fn = project_von_mises(Fn, mu, lam, R(mat.q_max), cs);
      tau = hencky_stress(Fn, mu, lam, cs);
    } else if (mat.model == 2u || mat.model == 3u) {  // DRUCKER_PRAGER / GRANULAR_POWDER
      Fn = project_drucker_prager(Fn, mu, lam, Mf, qc, cs);
      tau = hencky_stress(Fn, mu, lam, cs);
    } else if (mat.model == 5u) {  // VISCOPLASTIC_MUD
      Fn = project_drucker_prager(Fn, mu, lam, Mf, qc, cs);
      tau = hencky_stress(Fn, mu, lam, cs);
      M3<R> D = R(0.5) * (vgrad + transpose(vgrad));
      R trD = D.at(0, 0) + D.at(1, 1) + D.at(2, 2);
      M3<R> devD = D - (trD / R(3)) * M3<R>::ident();
      tau = tau + (R(2) * R(mat.viscosity)) * devD;
    } else {
      tau = hencky_stress(Fn, mu, lam, cs);
    }
    bool jreset = false;
    {
      R Jd = o_fabs(determinant(Fn));
      if (Jd < R(0.15) || Jd > R(10)) {
        Fn = M3<R>::ident();
        tau = M3<R>();
        jreset = true;
      }
    }
    if (p.flip_ratio > -0.5f) Bn = M3<R>();
    V3<R> vn;
    if (flip) {
      V3<R> vf = vold + (vpic - vpic_old);
      R a = R(p.flip_ratio);
      vn = vpic * (R(1) - a) + vf * a;  // GLSL mix(x, y, a) = x*(1-a) + y*a
    } else {
      vn = vpic;
    }
    if (jreset) vn = vn * R(0.1);
    V3<R> xn = xp + dt * vn;
    if (p.sdf) {
      int pj[3];
      V3<R> r2 = (xn - wmin) / h;
      pj[0] = int(o_floor(r2.x));
      pj[1] = int(o_floor(r2.y));
      pj[2] = int(o_floor(r2.z));
      for (int a = 0; a < 3; a++) pj[a] = std::min(std::max(pj[a], 0), gres[a] - 1);
      auto at = [&](int x, int y, int z) {
        x = std::min(std::max(x, 0), gres[0] - 1);
        y = std::min(std::max(y, 0), gres[1] - 1);
        z = std::min(std::max(z, 0), gres[2] - 1);
        return R(p.sdf[morton_encode(x, y, z)]);
      };
      R psdf = at(pj[0], pj[1], pj[2]);
      if (psdf < R(0)) {
        V3<R> gn(at(pj[0] + 1, pj[1], pj[2]) - at(pj[0] - 1, pj[1], pj[2]),
                 at(pj[0], pj[1] + 1, pj[2]) - at(pj[0], pj[1] - 1, pj[2]),
                 at(pj[0], pj[1], pj[2] + 1) - at(pj[0], pj[1], pj[2] - 1));
        R gl = length(gn);
        if (gl > R(1e-6)) {
          V3<R> nn = gn / gl;
          xn = xn - psdf * nn;
          R vdn = dot(vn, nn);
          if (vdn < R(0)) vn = vn - vdn * nn;
        } else {
          xn = xp;
          vn = V3<R>();
        }
        Fn = M3<R>::ident();
        tau = M3<R>();
      }
    }
    R margin = h * R(1.5);
    V3<R> lo = wmin + V3<R>(margin, margin, margin);
    V3<R> hi = wmax - V3<R>(margin, margin, margin);
    for (int a = 0; a < 3; a++) {
      if (xn[a] < lo[a]) { xn[a] = lo[a]; vn[a] = std::max(vn[a], R(0)); }
      if (xn[a] > hi[a]) { xn[a] = hi[a]; vn[a] = std::min(vn[a], R(0)); }
    }
    s.f[0][4 * q + 0] = xn.x; s.f[0][4 * q + 1] = xn.y; s.f[0][4 * q + 2] = xn.z; s.f[0][4 * q + 3] = Vp;
    s.f[1][4 * q + 0] = vn.x; s.f[1][4 * q + 1] = vn.y; s.f[1][4 * q + 2] = vn.z;
    for (int c = 0; c < 3; c++) {
      for (int r = 0; r < 3; r++) {
        s.f[2 + c][4 * q + r] = Fn.at(c, r);
        s.f[5 + c][4 * q + r] = Bn.at(c, r);
      }
    }
    s.f[2][4 * q + 3] = tau.at(0, 0);
    s.f[3][4 * q + 3] = tau.at(1, 1);
    s.f[4][4 * q + 3] = tau.at(2, 2);
    s.f[5][4 * q + 3] = tau.at(0, 1);
    s.f[6][4 * q + 3] = tau.at(0, 2);
    s.f[7][4 * q + 3] = tau.at(1, 2);
  }
}

// One substep in the reference's dispatch order (src/engine/MPMEngine.cpp:415-473).
template <class R> void substep(const Params& p, State<R>& s) {
  bin<R>(p, s.f[0].data(), s.n, s.keys.data(), s.sorted.data(), s.cell_count.data(), s.cell_end.data());
  p2g<R>(p, s);
  grid_update<R>(p, s);
  sdf_bc<R>(p, s);
  g2p<R>(p, s);
}

}  // namespace oracle
