// nb_adjoint_math.hpp — reverse-mode derivatives of the particle-local part of
// G2P: the plastic return maps (shaders/mpm_g2p.comp:19-68) and the Hencky
// stress (shaders/mpm_common.glsl:284-293), in the proper polar SVD semantics
// (F = U diag(sigma) V^T with U, V rotations).
//
// Both maps are isotropic:
//   F' = U diag(g(sigma)) V^T        (return map; g = exp(eps*(ln sigma)))
//   tau = U diag(k(sigma')) U^T      (Hencky; k_i = lambda tr eps' + 2 mu eps'_i)
// With dF^ = U^T dF V the SVD differential gives, for i != j,
//   (U^T dU)_ij = (s_j dF^_ij + s_i dF^_ji) / (s_j^2 - s_i^2)
//   (V^T dV)_ij = (s_i dF^_ij + s_j dF^_ji) / (s_j^2 - s_i^2)
// so the adjoints are formed in the singular frame and rotated back.  Every
// ratio of differences is evaluated as a divided difference (log, exp) that
// stays finite when singular values coincide (F = I is the common case).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/nebula_mpm.h"
#include "nb_device_math.hpp"

namespace nb {

// log strain of a singular value, as the forward uses it
__device__ __forceinline__ float eps_of(float s) { return logf(fmaxf(fabsf(s), 1e-6f)); }
__device__ __forceinline__ float deps_of(float s) { return fabsf(s) > 1e-6f ? 1.f / s : 0.f; }

// (eps_of(b) - eps_of(a)) / (b - a)
__device__ __forceinline__ float dd_eps(float a, float b) {
  const float d = b - a;
  if (a * b > 0.f && fabsf(a) > 1e-6f && fabsf(d) < 1e-3f * fabsf(a)) {
    const float r = d / a;
    return (1.f - r * (0.5f - r * (1.f / 3.f - 0.25f * r))) / a;
  }
  if (d == 0.f) return deps_of(a);
  return (eps_of(b) - eps_of(a)) / d;
}

// (e^y - e^x) / (y - x)
__device__ __forceinline__ float dd_exp(float x, float y) {
  const float d = y - x;
  if (fabsf(d) < 1e-3f) return expf(x) * (1.f + d * (0.5f + d * (1.f / 6.f + d * (1.f / 24.f))));
  return (expf(y) - expf(x)) / d;
}

// (A^T B) and (A B^T) in the m[col][row] storage
__device__ __forceinline__ m3 m3_tmul(const m3& A, const m3& B) { return m3_mul(m3_transpose(A), B); }
__device__ __forceinline__ m3 m3_mult(const m3& A, const m3& B) { return m3_mul(A, m3_transpose(B)); }

// The return map in log-strain space with its Jacobian and parameter
// derivatives.  model: 0 elastic (identity, F passes through), 1 Von Mises,
// 2 Drucker-Prager (always recomposes, shaders/mpm_g2p.comp:41-68).
struct ReturnMap {
  float es[3];      // projected log strains
  float J[3][3];    // d es_i / d e_j
  float dmu[3], dlam[3], dM[3], dqc[3];
  float s;          // (es_b - es_a) / (e_b - e_a), the deviatoric scale
  bool recompose;   // F is rebuilt from es (false: F passes through)
};

__device__ __forceinline__ void return_map(const float e[3], int model, float mu, float lam, float q_max, float M,
                                           float q_c, ReturnMap& o) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    o.es[i] = e[i];
    o.dmu[i] = o.dlam[i] = o.dM[i] = o.dqc[i] = 0.f;
#pragma unroll
    for (int j = 0; j < 3; j++) o.J[i][j] = (i == j) ? 1.f : 0.f;
  }
  o.s = 1.f;
  o.recompose = (model == 2);
  if (model == 0) return;
  const float tr = e[0] + e[1] + e[2];
  const float m = tr / 3.f;
  const float d[3] = {e[0] - m, e[1] - m, e[2] - m};
  const float nd = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  float c = 0.f, dcde = 0.f;  // radius of the return and its derivative along tr e / 3 per component
  float dc_mu = 0.f, dc_lam = 0.f, dc_M = 0.f, dc_qc = 0.f;
  bool yield = false;
  if (model == 2) {
    const float K = lam + 2.f * mu / 3.f;
    const float q_y = M * (-K * tr) + q_c;
    if (q_y <= 1e-8f) {
      // apex: pure volumetric strain, independent of e
      const float Ms = fmaxf(M, 1e-8f);
      const float m2 = (q_c / Ms / K) / 3.f;
      const float dm2_dK = -m2 / K;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        o.es[i] = m2;
#pragma unroll
        for (int j = 0; j < 3; j++) o.J[i][j] = 0.f;
        o.dqc[i] = 1.f / (3.f * Ms * K);
        o.dM[i] = M >= 1e-8f ? -m2 / M : 0.f;
        o.dmu[i] = dm2_dK * (2.f / 3.f);
        o.dlam[i] = dm2_dK;
      }
      o.s = 0.f;
      return;
    }
    if (nd - q_y / (2.f * mu) > 0.f && nd > 1e-8f) {
      yield = true;
      c = q_y / (2.f * mu);
      dcde = -M * K / (2.f * mu);
      const float dc_K = -M * tr / (2.f * mu);
      dc_mu = -c / mu + dc_K * (2.f / 3.f);
      dc_lam = dc_K;
      dc_M = -K * tr / (2.f * mu);
      dc_qc = 1.f / (2.f * mu);
    }
  } else if (model == 1) {
    if (nd - q_max / (2.f * mu) > 0.f && nd > 1e-8f) {
      yield = true;
      c = q_max / (2.f * mu);
      dc_mu = -c / mu;
    }
  }
  if (!yield) return;
  o.recompose = true;
  const float dh[3] = {d[0] / nd, d[1] / nd, d[2] / nd};
  const float sc = c / nd;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    o.es[i] = m + c * dh[i];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const float P = (i == j ? 1.f : 0.f) - 1.f / 3.f;
      o.J[i][j] = 1.f / 3.f + sc * (P - dh[i] * dh[j]) + dh[i] * dcde;
    }
    o.dmu[i] = dh[i] * dc_mu;
    o.dlam[i] = dh[i] * dc_lam;
    o.dM[i] = dh[i] * dc_M;
    o.dqc[i] = dh[i] * dc_qc;
  }
  o.s = sc;
}

// The forward of the particle-local map from one SVD of F_tmp: singular
// frame, return map and the singular values of F' (ss; F' = U diag(ss) V^T
// when rm.recompose, F' = F_tmp otherwise).
struct PlasticFwd {
  m3 U, V;
  float sig[3], ss[3];
  ReturnMap rm;
};

__device__ __forceinline__ void plastic_forward(const m3& Ft, int model, float mu, float lam, float q_max, float M,
                                                float q_c, PlasticFwd& f) {
  svd3(Ft, f.U, f.sig, f.V, true);
  float e[3];
#pragma unroll
  for (int i = 0; i < 3; i++) e[i] = eps_of(f.sig[i]);
  return_map(e, model, mu, lam, q_max, M, q_c, f.rm);
#pragma unroll
  for (int i = 0; i < 3; i++) f.ss[i] = f.rm.recompose ? expf(f.rm.es[i]) : f.sig[i];
}

// Adjoint of  F' = return_map(F_tmp), tau = hencky(F').
// gFp = dL/dF', gT = dL/dtau (full symmetric matrix).  Returns dL/dF_tmp and
// adds dL/d(mu, lambda, M, q_c) to g4.
__device__ inline m3 plastic_hencky_adjoint(const PlasticFwd& f, float mu, float lam, const m3& gFp, const m3& gT,
                                            float g4[4]) {
  const m3& U = f.U;
  const m3& V = f.V;
  const float* sig = f.sig;
  const float* ss = f.ss;
  const ReturnMap& rm = f.rm;
  float ep[3];
#pragma unroll
  for (int i = 0; i < 3; i++) ep[i] = eps_of(ss[i]);
  // Hencky: tau = U diag(k) U^T, k_i = lam tr(ep) + 2 mu ep_i
  const m3 Gt = m3_mul(m3_tmul(U, gT), U);  // U^T gT U
  const float trG = Gt.m[0][0] + Gt.m[1][1] + Gt.m[2][2];
  const float Sep = ep[0] + ep[1] + ep[2];
  g4[0] += 2.f * (Gt.m[0][0] * ep[0] + Gt.m[1][1] * ep[1] + Gt.m[2][2] * ep[2]);
  g4[1] += trG * Sep;
  m3 Gh = m3_mul(m3_tmul(U, gFp), V);  // U^T gF' V  (Gh.m[b][a] = entry (a, b))
#pragma unroll
  for (int j = 0; j < 3; j++) Gh.m[j][j] += (lam * trG + 2.f * mu * Gt.m[j][j]) * deps_of(ss[j]);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      if (a == b) continue;
      const float den = ss[a] + ss[b];
      const float R = fabsf(den) > 1e-12f ? Gt.m[b][a] * 2.f * mu * dd_eps(ss[a], ss[b]) / den : 0.f;
      Gh.m[b][a] += 2.f * ss[b] * R;
    }
  if (!rm.recompose) return m3_mult(m3_mul(U, Gh), V);  // U Gh V^T
  // return map: ss_i = exp(es_i(e(sig)))
  m3 G2;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    float gs = 0.f;
#pragma unroll
    for (int i = 0; i < 3; i++) gs += Gh.m[i][i] * ss[i] * rm.J[i][j];
    G2.m[j][j] = gs * deps_of(sig[j]);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float gi = Gh.m[i][i] * ss[i];
    g4[0] += gi * rm.dmu[i];
    g4[1] += gi * rm.dlam[i];
    g4[2] += gi * rm.dM[i];
    g4[3] += gi * rm.dqc[i];
  }
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      if (a == b) continue;
      const float den = sig[a] + sig[b];
      const float S = fabsf(den) > 1e-12f ? (ss[a] + ss[b]) / den : 0.f;
      const float D = dd_exp(rm.es[a], rm.es[b]) * rm.s * dd_eps(sig[a], sig[b]);
      G2.m[b][a] = Gh.m[b][a] * 0.5f * (S + D) + Gh.m[a][b] * 0.5f * (D - S);
    }
  return m3_mult(m3_mul(U, G2), V);
}

}  // namespace nb
