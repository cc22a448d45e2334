// nb_device_math.hpp — gfx950 device math for the MPM substep.
//
// Register-resident 3x3 algebra (column-major, m[col][row] as in GLSL), Morton
// keys, quadratic B-splines, the Jacobi/F^T F SVD, Hencky stress and the two
// plasticity return maps.  The algorithms follow the reference's GLSL
// (shaders/mpm_common.glsl:161-302, shaders/mpm_g2p.comp:19-68) step for step,
// because the particle state they produce is compared with it; the code is
// written for wave64 VALU execution: every loop has a compile-time trip count
// so the compiler keeps the matrices in VGPRs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nb {

struct m3 {
  float m[3][3];  // m[col][row]
};

__device__ __forceinline__ m3 m3_zero() {
  m3 r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = 0.f;
  return r;
}
__device__ __forceinline__ m3 m3_ident() {
  m3 r = m3_zero();
  r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.f;
  return r;
}
__device__ __forceinline__ m3 m3_diag(float a, float b, float c) {
  m3 r = m3_zero();
  r.m[0][0] = a;
  r.m[1][1] = b;
  r.m[2][2] = c;
  return r;
}
// (A*B)[col][row] = sum_k A[k][row] * B[col][k]
__device__ __forceinline__ m3 m3_mul(const m3& A, const m3& B) {
  m3 r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = A.m[0][k] * B.m[c][0] + A.m[1][k] * B.m[c][1] + A.m[2][k] * B.m[c][2];
  return r;
}
__device__ __forceinline__ m3 m3_transpose(const m3& A) {
  m3 r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = A.m[k][c];
  return r;
}
__device__ __forceinline__ float m3_det(const m3& m) {
  return m.m[0][0] * (m.m[1][1] * m.m[2][2] - m.m[2][1] * m.m[1][2]) -
         m.m[1][0] * (m.m[0][1] * m.m[2][2] - m.m[2][1] * m.m[0][2]) +
         m.m[2][0] * (m.m[0][1] * m.m[1][2] - m.m[1][1] * m.m[0][2]);
}
// U * diag(d) * V^T
__device__ __forceinline__ m3 m3_udvt(const m3& U, float d0, float d1, float d2, const m3& V) {
  m3 r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = U.m[0][k] * d0 * V.m[0][c] + U.m[1][k] * d1 * V.m[1][c] + U.m[2][k] * d2 * V.m[2][c];
  return r;
}

__device__ __forceinline__ float glsl_sign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

// ---- Morton keys (shaders/mpm_common.glsl:161-187) ----
__host__ __device__ __forceinline__ uint32_t morton_expand(uint32_t v) {
  v = (v | (v << 16u)) & 0x030000FFu;
  v = (v | (v << 8u)) & 0x0300F00Fu;
  v = (v | (v << 4u)) & 0x030C30C3u;
  v = (v | (v << 2u)) & 0x09249249u;
  return v;
}
__host__ __device__ __forceinline__ uint32_t morton_compact(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2u)) & 0x030C30C3u;
  v = (v | (v >> 4u)) & 0x0300F00Fu;
  v = (v | (v >> 8u)) & 0x030000FFu;
  v = (v | (v >> 16u)) & 0x000003FFu;
  return v;
}
__host__ __device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return morton_expand(x) | (morton_expand(y) << 1u) | (morton_expand(z) << 2u);
}

// ---- hardware reciprocal, square root and transcendentals ----
// One VALU instruction each, accurate to about 1 ulp (log/exp: base-2
// instructions with the base change folded into one multiply).  GLSL only
// bounds division (2.5 ulp), sqrt/inversesqrt, log and exp (3 ulp; exp
// 3 + 2|x| ulp) by ulps, so the reference's shaders run with accuracy of this
// order.  Cell keys and stencil bases keep the IEEE quotient: they must land in
// the same cell as the reference's.
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float hw_log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
__device__ __forceinline__ float hw_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// ---- Jacobi eigen-decomposition of symmetric A (shaders/mpm_common.glsl:213-247) ----
// Same pivot rule, rotation and stopping test as the GLSL.  The matrix is held
// as its six independent entries and the three pivot cases are resolved with
// selects, so every lane runs one instruction stream and nothing is indexed
// dynamically (no scratch).  V keeps the GLSL indexing V[col][row]: the GLSL
// rotates V[r][p0], V[r][p1], i.e. here V.m[r][p0], V.m[r][p1].
__device__ __forceinline__ void jacobi_eigen3(float a00, float a11, float a22, float a01, float a02, float a12,
                                              float D[3], m3& V) {
  V = m3_ident();
  for (int iter = 0; iter < 20; iter++) {
    float b01 = fabsf(a01), b02 = fabsf(a02), b12 = fabsf(a12);
    if (fmaxf(b01, fmaxf(b02, b12)) < 1e-8f) break;
    // pivot case: 0 -> (0,1), 1 -> (0,2), 2 -> (1,2)
    int cs = (b01 >= fmaxf(b02, b12)) ? 0 : ((b02 >= b12) ? 1 : 2);
    float app = (cs == 2) ? a11 : a00;
    float aqq = (cs == 0) ? a11 : a22;
    float apq = (cs == 0) ? a01 : ((cs == 1) ? a02 : a12);
    float arp = (cs == 0) ? a02 : a01;  // A[r][p0]
    float arq = (cs == 2) ? a02 : a12;  // A[r][p1]
    // tau = (aqq - app) / (2 apq); t = sign(tau) / (|tau| + sqrt(1 + tau^2));
    // c = 1 / sqrt(1 + t^2); s = t c
    float tau = (aqq - app) * hw_rcp(2.f * apq);
    float t = glsl_sign(tau) * hw_rcp(fabsf(tau) + hw_sqrt(1.f + tau * tau));
    float c = hw_rsq(1.f + t * t);
    float s = t * c;
    float napp = app - t * apq, naqq = aqq + t * apq;
    float nrp = c * arp - s * arq;
    float nrq = s * arp + c * arq;
    a00 = (cs == 2) ? a00 : napp;
    a11 = (cs == 0) ? naqq : ((cs == 2) ? napp : a11);
    a22 = (cs == 0) ? a22 : naqq;
    a01 = (cs == 0) ? 0.f : nrp;
    a02 = (cs == 0) ? nrp : ((cs == 1) ? 0.f : nrq);
    a12 = (cs == 2) ? 0.f : nrq;
#pragma unroll
    for (int rr = 0; rr < 3; rr++) {
      float vp = (cs == 2) ? V.m[rr][1] : V.m[rr][0];
      float vq = (cs == 0) ? V.m[rr][1] : V.m[rr][2];
      float nvp = c * vp - s * vq;
      float nvq = s * vp + c * vq;
      V.m[rr][0] = (cs == 2) ? V.m[rr][0] : nvp;
      V.m[rr][1] = (cs == 0) ? nvq : ((cs == 2) ? nvp : V.m[rr][1]);
      V.m[rr][2] = (cs == 0) ? V.m[rr][2] : nvq;
    }
  }
  D[0] = a00;
  D[1] = a11;
  D[2] = a22;
}

// ---- SVD via F^T F (shaders/mpm_common.glsl:250-281) ----
// corrected = false keeps the GLSL storage (singular vectors in the rows of
// U, V as the GLSL mat3 holds them); corrected = true makes V a rotation before
// U is formed and transposes both, so F = U diag(sigma) V^T with U, V in SO(3)
// and the downstream U diag U^T / U diag V^T are the true spectral forms.
__device__ __forceinline__ void svd3(const m3& F, m3& U, float sig[3], m3& V, bool corrected) {
  // FtF = transpose(F) * F: FtF[c][k] = dot(F col k, F col c)
  float f00 = F.m[0][0] * F.m[0][0] + F.m[0][1] * F.m[0][1] + F.m[0][2] * F.m[0][2];
  float f11 = F.m[1][0] * F.m[1][0] + F.m[1][1] * F.m[1][1] + F.m[1][2] * F.m[1][2];
  float f22 = F.m[2][0] * F.m[2][0] + F.m[2][1] * F.m[2][1] + F.m[2][2] * F.m[2][2];
  float f01 = F.m[0][0] * F.m[1][0] + F.m[0][1] * F.m[1][1] + F.m[0][2] * F.m[1][2];
  float f02 = F.m[0][0] * F.m[2][0] + F.m[0][1] * F.m[2][1] + F.m[0][2] * F.m[2][2];
  float f12 = F.m[1][0] * F.m[2][0] + F.m[1][1] * F.m[2][1] + F.m[1][2] * F.m[2][2];
  float D[3];
  jacobi_eigen3(f00, f11, f22, f01, f02, f12, D, V);
  sig[0] = hw_sqrt(fmaxf(D[0], 0.f));
  sig[1] = hw_sqrt(fmaxf(D[1], 0.f));
  sig[2] = hw_sqrt(fmaxf(D[2], 0.f));
  // selection sort, descending; swap V columns i<->j (GLSL V[0][i] = column 0
  // row i: it swaps ROWS of our m[col][row] storage at index i, j)
#define NB_SWAP_SV(i, j)                                                        \
  if (sig[j] > sig[i]) {                                                        \
    float ts = sig[i]; sig[i] = sig[j]; sig[j] = ts;                            \
    _Pragma("unroll") for (int cc = 0; cc < 3; cc++) {                          \
      float tv = V.m[cc][i]; V.m[cc][i] = V.m[cc][j]; V.m[cc][j] = tv;          \
    }                                                                           \
  }
  NB_SWAP_SV(0, 1)
  NB_SWAP_SV(0, 2)
  NB_SWAP_SV(1, 2)
#undef NB_SWAP_SV
  if (corrected && m3_det(V) < 0.f) {  // keep V a rotation: flip the third singular vector
    V.m[0][2] = -V.m[0][2];
    V.m[1][2] = -V.m[1][2];
    V.m[2][2] = -V.m[2][2];
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    // GLSL vi = vec3(V[0][i], V[1][i], V[2][i]); U[r][i] = (F*vi)[r]/sigma_i,
    // or the unit vector e_i when sigma_i <= 1e-8 (U starts as zero there)
    const float v0 = V.m[0][i], v1 = V.m[1][i], v2 = V.m[2][i];
    const float fv0 = F.m[0][0] * v0 + F.m[1][0] * v1 + F.m[2][0] * v2;
    const float fv1 = F.m[0][1] * v0 + F.m[1][1] * v1 + F.m[2][1] * v2;
    const float fv2 = F.m[0][2] * v0 + F.m[1][2] * v1 + F.m[2][2] * v2;
    const bool big = sig[i] > 1e-8f;
    const float inv = hw_rcp(big ? sig[i] : 1.f);
    U.m[0][i] = big ? fv0 * inv : (i == 0 ? 1.f : 0.f);
    U.m[1][i] = big ? fv1 * inv : (i == 1 ? 1.f : 0.f);
    U.m[2][i] = big ? fv2 * inv : (i == 2 ? 1.f : 0.f);
  }
  if (m3_det(U) < 0.f) {
    U.m[0][2] = -U.m[0][2];
    U.m[1][2] = -U.m[1][2];
    U.m[2][2] = -U.m[2][2];
    sig[2] = -sig[2];
  }
  if (corrected) {
    U = m3_transpose(U);
    V = m3_transpose(V);
  }
}

// ---- Hencky Kirchhoff stress (shaders/mpm_common.glsl:284-293) ----
// tau = U diag(lambda tr(eps) + 2 mu eps) U^T with U as the GLSL stores it;
// eps_i = log(max(|sigma_i|, 1e-6)).
__device__ __forceinline__ m3 hencky_from_log(const m3& U, float e0, float e1, float e2, float mu, float lam) {
  const float lt = lam * (e0 + e1 + e2), tm = 2.f * mu;
  return m3_udvt(U, lt + tm * e0, lt + tm * e1, lt + tm * e2, U);
}
__device__ __forceinline__ float log_strain(float s) { return hw_log(fmaxf(fabsf(s), 1e-6f)); }
__device__ __forceinline__ m3 hencky_stress(const m3& F, float mu, float lam, bool corrected) {
  m3 U, V;
  float sig[3];
  svd3(F, U, sig, V, corrected);
  return hencky_from_log(U, log_strain(sig[0]), log_strain(sig[1]), log_strain(sig[2]), mu, lam);
}

// ---- plastic return maps (shaders/mpm_g2p.comp:19-68) ----
// One SVD serves both maps.  Von Mises (dp = false) scales the deviatoric log
// strain back to the yield radius q_max / 2mu and keeps F untouched when
// inside it; Drucker-Prager (dp = true) projects onto the cone
// q <= M p + q_c (apex: volumetric strain only) and always recomposes F.
// Returns the projected F; eps[] receives its log strains and U its left
// singular vectors, which give the Hencky stress of the result directly when
// U, V are the true SVD factors (corrected mode).
struct PlasticOut {
  m3 F;
  m3 U;
  float eps[3];
  bool recomposed;  // false: F was returned as given, and U, eps are its own SVD's
};
__device__ __forceinline__ PlasticOut project_plastic(const m3& F, bool dp, float mu, float lam, float q_max, float M,
                                                      float q_c, bool corrected) {
  m3 U, V;
  float sig[3];
  svd3(F, U, sig, V, corrected);
  float e0 = log_strain(sig[0]), e1 = log_strain(sig[1]), e2 = log_strain(sig[2]);
  const float tr = e0 + e1 + e2;
  const float m = tr / 3.f;
  float d0 = e0 - m, d1 = e1 - m, d2 = e2 - m;
  const float nd = hw_sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  bool recompose = dp;
  float q_lim = q_max;
  bool apex = false;
  if (dp) {
    const float K = lam + 2.f * mu / 3.f;
    const float q_yield = M * (-K * tr) + q_c;
    apex = q_yield <= 1e-8f;
    q_lim = q_yield;
    if (apex) {
      const float m2 = (q_c / fmaxf(M, 1e-8f) / K) / 3.f;  // tr2 = -p_proj / K, p_proj = -q_c / M
      e0 = m2; e1 = m2; e2 = m2;
    }
  }
  if (!apex) {
    const float delta = nd - q_lim / (2.f * mu);
    if (delta > 0.f && nd > 1e-8f) {
      const float k = delta * hw_rcp(nd);
      e0 = d0 - k * d0 + m;
      e1 = d1 - k * d1 + m;
      e2 = d2 - k * d2 + m;
      recompose = true;
    }
  }
  PlasticOut o;
  o.U = U;
  o.recomposed = recompose;
  if (recompose) {
    const float x0 = hw_exp(e0), x1 = hw_exp(e1), x2 = hw_exp(e2);
    o.F = m3_udvt(U, x0, x1, x2, V);
    o.eps[0] = log_strain(x0);
    o.eps[1] = log_strain(x1);
    o.eps[2] = log_strain(x2);
  } else {
    o.F = F;
    o.eps[0] = e0;
    o.eps[1] = e1;
    o.eps[2] = e2;
  }
  return o;
}

}  // namespace nb
