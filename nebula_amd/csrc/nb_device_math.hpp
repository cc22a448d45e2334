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

// column-major 3x3, m[col][row] as in GLSL; float for the substep, double
// for the FP64 SVD fallback of near-degenerate F
template <typename T>
struct mat3 {
  T m[3][3];
};
using m3 = mat3<float>;
using m3d = mat3<double>;

template <typename T>
__device__ __forceinline__ mat3<T> mat3_zero() {
  mat3<T> r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = T(0);
  return r;
}
template <typename T>
__device__ __forceinline__ mat3<T> mat3_ident() {
  mat3<T> r = mat3_zero<T>();
  r.m[0][0] = r.m[1][1] = r.m[2][2] = T(1);
  return r;
}
__device__ __forceinline__ m3 m3_zero() { return mat3_zero<float>(); }
__device__ __forceinline__ m3 m3_ident() { return mat3_ident<float>(); }
__device__ __forceinline__ m3 m3_diag(float a, float b, float c) {
  m3 r = m3_zero();
  r.m[0][0] = a;
  r.m[1][1] = b;
  r.m[2][2] = c;
  return r;
}
// (A*B)[col][row] = sum_k A[k][row] * B[col][k]
template <typename T>
__device__ __forceinline__ mat3<T> m3_mul(const mat3<T>& A, const mat3<T>& B) {
  mat3<T> r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = A.m[0][k] * B.m[c][0] + A.m[1][k] * B.m[c][1] + A.m[2][k] * B.m[c][2];
  return r;
}
template <typename T>
__device__ __forceinline__ mat3<T> m3_transpose(const mat3<T>& A) {
  mat3<T> r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = A.m[k][c];
  return r;
}
template <typename T>
__device__ __forceinline__ T m3_det(const mat3<T>& m) {
  return m.m[0][0] * (m.m[1][1] * m.m[2][2] - m.m[2][1] * m.m[1][2]) -
         m.m[1][0] * (m.m[0][1] * m.m[2][2] - m.m[2][1] * m.m[0][2]) +
         m.m[2][0] * (m.m[0][1] * m.m[1][2] - m.m[1][1] * m.m[0][2]);
}
// U * diag(d) * V^T
template <typename T>
__device__ __forceinline__ mat3<T> m3_udvt(const mat3<T>& U, T d0, T d1, T d2, const mat3<T>& V) {
  mat3<T> r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = U.m[0][k] * d0 * V.m[0][c] + U.m[1][k] * d1 * V.m[1][c] + U.m[2][k] * d2 * V.m[2][c];
  return r;
}
template <typename T>
__device__ __forceinline__ mat3<T> m3_convert(const mat3<float>& A) {
  mat3<T> r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = T(A.m[c][k]);
  return r;
}
__device__ __forceinline__ m3 m3_round(const m3d& A) {
  m3 r;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) r.m[c][k] = float(A.m[c][k]);
  return r;
}

template <typename T>
__device__ __forceinline__ T glsl_sign(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }
__device__ __forceinline__ float tmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double tmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float tabs(float a) { return fabsf(a); }
__device__ __forceinline__ double tabs(double a) { return fabs(a); }

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
// The FP64 fallback: the hardware reciprocal and reciprocal square root
// refined by two Newton steps each, so quotients and roots land within an
// ulp or two of FP64 (the oracle's correctly rounded operations), far below
// the single FP32 rounding the fallback's results take; log and exp are the
// library's.  Arguments are finite, and the reciprocal's nonzero.
__device__ __forceinline__ double hw_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double hw_rsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = fma(y, fma(-h * y, y, 0.5), y);
  return fma(y, fma(-h * y, y, 0.5), y);
}
__device__ __forceinline__ double hw_sqrt(double x) {
  if (!(x > 0.0)) return x == 0.0 ? 0.0 : sqrt(x);
  const double y = hw_rsq(x), s = x * y;
  return fma(0.5 * y, fma(-s, s, x), s);
}
__device__ __forceinline__ double hw_log(double x) { return log(x); }
__device__ __forceinline__ double hw_exp(double x) { return exp(x); }

// a / b: float multiplies by the hardware reciprocal; double multiplies by
// the refined reciprocal and corrects the quotient once
__device__ __forceinline__ float tdiv(float a, float b) { return a * hw_rcp(b); }
__device__ __forceinline__ double tdiv(double a, double b) {
  const double r = hw_rcp(b), q = a * r;
  return fma(r, fma(-b, q, a), q);
}

// ---- Jacobi eigen-decomposition of symmetric A (shaders/mpm_common.glsl:213-247) ----
// Two forms of the same iteration.  jacobi_eigen3_select keeps the matrix in
// its physical layout and resolves the three pivot cases with selects;
// jacobi_eigen3 (below) keeps it relative to the last pivot and is the one the
// kernels use.  Both take the same pivots in the same order and differ only in
// the rounding of FMA groupings.
// Same pivot rule, rotation and stopping test as the GLSL.  The matrix is held
// as its six independent entries and the three pivot cases are resolved with
// selects, so every lane runs one instruction stream and nothing is indexed
// dynamically (no scratch).  V keeps the GLSL indexing V[col][row]: the GLSL
// rotates V[r][p0], V[r][p1], i.e. here V.m[r][p0], V.m[r][p1].
template <typename T>
__device__ __forceinline__ void jacobi_eigen3_select(T a00, T a11, T a22, T a01, T a02, T a12, T D[3], mat3<T>& V) {
  V = mat3_ident<T>();
  for (int iter = 0; iter < 20; iter++) {
    T b01 = tabs(a01), b02 = tabs(a02), b12 = tabs(a12);
    if (tmax(b01, tmax(b02, b12)) < T(1e-8)) break;
    // pivot case: 0 -> (0,1), 1 -> (0,2), 2 -> (1,2)
    int cs = (b01 >= tmax(b02, b12)) ? 0 : ((b02 >= b12) ? 1 : 2);
    T app = (cs == 2) ? a11 : a00;
    T aqq = (cs == 0) ? a11 : a22;
    T apq = (cs == 0) ? a01 : ((cs == 1) ? a02 : a12);
    T arp = (cs == 0) ? a02 : a01;  // A[r][p0]
    T arq = (cs == 2) ? a02 : a12;  // A[r][p1]
    // tau = (aqq - app) / (2 apq); t = sign(tau) / (|tau| + sqrt(1 + tau^2));
    // c = 1 / sqrt(1 + t^2); s = t c
    T tau = tdiv(aqq - app, T(2) * apq);
    T t = tdiv(glsl_sign(tau), tabs(tau) + hw_sqrt(T(1) + tau * tau));
    T c = hw_rsq(T(1) + t * t);
    T s = t * c;
    T napp = app - t * apq, naqq = aqq + t * apq;
    T nrp = c * arp - s * arq;
    T nrq = s * arp + c * arq;
    a00 = (cs == 2) ? a00 : napp;
    a11 = (cs == 0) ? naqq : ((cs == 2) ? napp : a11);
    a22 = (cs == 0) ? a22 : naqq;
    a01 = (cs == 0) ? T(0) : nrp;
    a02 = (cs == 0) ? nrp : ((cs == 1) ? T(0) : nrq);
    a12 = (cs == 2) ? T(0) : nrq;
#pragma unroll
    for (int rr = 0; rr < 3; rr++) {
      T vp = (cs == 2) ? V.m[rr][1] : V.m[rr][0];
      T vq = (cs == 0) ? V.m[rr][1] : V.m[rr][2];
      T nvp = c * vp - s * vq;
      T nvq = s * vp + c * vq;
      V.m[rr][0] = (cs == 2) ? V.m[rr][0] : nvp;
      V.m[rr][1] = (cs == 0) ? nvq : ((cs == 2) ? nvp : V.m[rr][1]);
      V.m[rr][2] = (cs == 0) ? V.m[rr][2] : nvq;
    }
  }
  D[0] = a00;
  D[1] = a11;
  D[2] = a22;
}

// Role form: after a rotation of the pair (p, q) its off-diagonal entry is
// exactly zero, so the next pivot is one of the two pairs that contain the
// third index r.  The matrix is kept as (d_R, d_P, d_Q, o_RP, o_RQ) with
// o_PQ = 0, the columns of V as (V_R, V_P, V_Q) and the physical index of
// each role; an iteration picks X in {P, Q} (the larger |o_RX|; on a tie the
// pair with the lower physical pivot case, i.e. the smaller physical index of
// X, as the GLSL's ordered comparisons do), rotates the pair (R, X) and
// relabels: (R, P, Q) <- (Y, R, X) with Y the other of P, Q.  In the rotated
// pair A[Y][X] was zero, so the third row's update is o_YR <- c o_YR and
// o_YX <- s o_YR.  The rotation is the GLSL's written for the pair (R, X): the
// Jacobi rotation is symmetric in its two indices (swapping them negates tau,
// t and s and gives the same new entries), so the orientation needs no
// physical order.  Per iteration this needs 12 selects instead of 50, and the
// first rotation (from A itself) and the return to physical order are done
// once.
template <typename T>
__device__ __forceinline__ T jacobi_t(T tau) {
  // t = sign(tau) / (|tau| + sqrt(1 + tau^2)), with GLSL's sign(0) = 0
  const T r = tdiv(T(1), tabs(tau) + hw_sqrt(T(1) + tau * tau));
  return tau == T(0) ? T(0) : copysign(r, tau);
}
template <typename T>
__device__ __forceinline__ void jacobi_eigen3(T a00, T a11, T a22, T a01, T a02, T a12, T D[3], mat3<T>& V) {
  const T b01 = tabs(a01), b02 = tabs(a02), b12 = tabs(a12);
  if (tmax(b01, tmax(b02, b12)) < T(1e-8)) {
    V = mat3_ident<T>();
    D[0] = a00;
    D[1] = a11;
    D[2] = a22;
    return;
  }
  // first rotation, on A as given: pivot case 0 -> (0,1), 1 -> (0,2), 2 -> (1,2)
  const int cs = (b01 >= tmax(b02, b12)) ? 0 : ((b02 >= b12) ? 1 : 2);
  int pP = (cs == 2) ? 1 : 0;        // p0
  int pQ = (cs == 0) ? 1 : 2;        // p1
  int pR = 3 - pP - pQ;              // the third index
  T dR = (cs == 0) ? a22 : ((cs == 1) ? a11 : a00);
  T dP = (cs == 2) ? a11 : a00;
  T dQ = (cs == 0) ? a11 : a22;
  const T apq = (cs == 0) ? a01 : ((cs == 1) ? a02 : a12);
  const T arp = (cs == 0) ? a02 : a01;  // A[r][p0]
  const T arq = (cs == 2) ? a02 : a12;  // A[r][p1]
  T oRP, oRQ;
  T vR[3], vP[3], vQ[3];
  {
    const T tau = tdiv(dQ - dP, T(2) * apq);
    const T t = jacobi_t(tau);
    const T c = hw_rsq(T(1) + t * t);
    const T s = t * c;
    const T np = dP - t * apq, nq = dQ + t * apq;
    dP = np;
    dQ = nq;
    oRP = c * arp - s * arq;
    oRQ = s * arp + c * arq;
#pragma unroll
    for (int rr = 0; rr < 3; rr++) {  // V = I rotated in the pair (p0, p1)
      vR[rr] = (rr == pR) ? T(1) : T(0);
      vP[rr] = (rr == pP) ? c : ((rr == pQ) ? -s : T(0));
      vQ[rr] = (rr == pP) ? s : ((rr == pQ) ? c : T(0));
    }
  }
  for (int iter = 1; iter < 20; iter++) {
    const T bRP = tabs(oRP), bRQ = tabs(oRQ);
    if (bRP < T(1e-8) && bRQ < T(1e-8)) break;
    const bool xp = bRP > bRQ || (bRP == bRQ && pP < pQ);  // X = P, else X = Q
    const T oRX = xp ? oRP : oRQ;
    const T oYR = xp ? oRQ : oRP;
    const T dX = xp ? dP : dQ;
    const T dY = xp ? dQ : dP;
    const int pX = xp ? pP : pQ;
    const int pY = xp ? pQ : pP;
    const T tau = tdiv(dX - dR, T(2) * oRX);
    const T t = jacobi_t(tau);
    const T c = hw_rsq(T(1) + t * t);
    const T s = t * c;
    const T nR = dR - t * oRX, nX = dX + t * oRX;
    // relabel (R, P, Q) <- (Y, R, X)
    oRP = c * oYR;
    oRQ = s * oYR;
    dP = nR;
    dQ = nX;
    dR = dY;
#pragma unroll
    for (int rr = 0; rr < 3; rr++) {
      const T vx = xp ? vP[rr] : vQ[rr];
      const T vy = xp ? vQ[rr] : vP[rr];
      const T vr = vR[rr];
      vP[rr] = c * vr - s * vx;
      vQ[rr] = s * vr + c * vx;
      vR[rr] = vy;
    }
    pQ = pX;
    pP = pR;
    pR = pY;
  }
  // back to physical order
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const bool r = pR == i, pp = pP == i;
    D[i] = r ? dR : (pp ? dP : dQ);
#pragma unroll
    for (int rr = 0; rr < 3; rr++) V.m[rr][i] = r ? vR[rr] : (pp ? vP[rr] : vQ[rr]);
  }
}

// ---- SVD via F^T F (shaders/mpm_common.glsl:250-281) ----
// corrected = false keeps the GLSL storage (singular vectors in the rows of
// U, V as the GLSL mat3 holds them); corrected = true makes V a rotation before
// U is formed and transposes both, so F = U diag(sigma) V^T with U, V in SO(3)
// and the downstream U diag U^T / U diag V^T are the true spectral forms.
template <typename T>
__device__ __forceinline__ void svd3(const mat3<T>& F, mat3<T>& U, T sig[3], mat3<T>& V, bool corrected) {
  // FtF = transpose(F) * F: FtF[c][k] = dot(F col k, F col c)
  T f00 = F.m[0][0] * F.m[0][0] + F.m[0][1] * F.m[0][1] + F.m[0][2] * F.m[0][2];
  T f11 = F.m[1][0] * F.m[1][0] + F.m[1][1] * F.m[1][1] + F.m[1][2] * F.m[1][2];
  T f22 = F.m[2][0] * F.m[2][0] + F.m[2][1] * F.m[2][1] + F.m[2][2] * F.m[2][2];
  T f01 = F.m[0][0] * F.m[1][0] + F.m[0][1] * F.m[1][1] + F.m[0][2] * F.m[1][2];
  T f02 = F.m[0][0] * F.m[2][0] + F.m[0][1] * F.m[2][1] + F.m[0][2] * F.m[2][2];
  T f12 = F.m[1][0] * F.m[2][0] + F.m[1][1] * F.m[2][1] + F.m[1][2] * F.m[2][2];
  T D[3];
  jacobi_eigen3<T>(f00, f11, f22, f01, f02, f12, D, V);
  sig[0] = hw_sqrt(tmax(D[0], T(0)));
  sig[1] = hw_sqrt(tmax(D[1], T(0)));
  sig[2] = hw_sqrt(tmax(D[2], T(0)));
  // selection sort, descending; swap V columns i<->j (GLSL V[0][i] = column 0
  // row i: it swaps ROWS of our m[col][row] storage at index i, j)
#define NB_SWAP_SV(i, j)                                                        \
  if (sig[j] > sig[i]) {                                                        \
    T ts = sig[i]; sig[i] = sig[j]; sig[j] = ts;                                \
    _Pragma("unroll") for (int cc = 0; cc < 3; cc++) {                          \
      T tv = V.m[cc][i]; V.m[cc][i] = V.m[cc][j]; V.m[cc][j] = tv;              \
    }                                                                           \
  }
  NB_SWAP_SV(0, 1)
  NB_SWAP_SV(0, 2)
  NB_SWAP_SV(1, 2)
#undef NB_SWAP_SV
  if (corrected && m3_det(V) < T(0)) {  // keep V a rotation: flip the third singular vector
    V.m[0][2] = -V.m[0][2];
    V.m[1][2] = -V.m[1][2];
    V.m[2][2] = -V.m[2][2];
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    // GLSL vi = vec3(V[0][i], V[1][i], V[2][i]); U[r][i] = (F*vi)[r]/sigma_i,
    // or the unit vector e_i when sigma_i <= 1e-8 (U starts as zero there)
    const T v0 = V.m[0][i], v1 = V.m[1][i], v2 = V.m[2][i];
    const T fv0 = F.m[0][0] * v0 + F.m[1][0] * v1 + F.m[2][0] * v2;
    const T fv1 = F.m[0][1] * v0 + F.m[1][1] * v1 + F.m[2][1] * v2;
    const T fv2 = F.m[0][2] * v0 + F.m[1][2] * v1 + F.m[2][2] * v2;
    const bool big = sig[i] > T(1e-8);
    const T den = big ? sig[i] : T(1);
    U.m[0][i] = big ? tdiv(fv0, den) : (i == 0 ? T(1) : T(0));
    U.m[1][i] = big ? tdiv(fv1, den) : (i == 1 ? T(1) : T(0));
    U.m[2][i] = big ? tdiv(fv2, den) : (i == 2 ? T(1) : T(0));
  }
  if (m3_det(U) < T(0)) {
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

// ---- near-degenerate F: the FP64 fallback's test ----
// On singular values sorted by magnitude (s0 >= s1 >= |s2|) the eigenvalues
// of F^T F are s_i^2.  FP32 rounds F^T F's entries by about 2^-22 s0^2, so
// the singular vectors of a close pair (i, j) are fixed only up to a rotation
// in their plane of angle theta_ij ~ min(1, 2^-22 s0^2 / (s_i^2 - s_j^2)),
// and U = F v / sigma loses about 1e-7 s0 / s2 for a small s2.  Rotating the
// pair (U -> U G, V -> V G) changes
//   the reference convention's X = U^T D V (F' = U^T exp(eps) V and
//   tau = U^T k U, not isotropic functions of F) to G^T X G, by
//   theta [X, K_ij] (K_ij = e_i e_j^T - e_j e_i^T), and
//   the corrected convention's X = U D V^T by theta U [K_ij, D] V^T, whose
//   entries are theta (d_i - d_j).
// A lane takes the FP64 path when s2^2 <= zero_rel s0^2, or when a pair
// with s_i^2 - s_j^2 <= gap_rel s0^2 moves F' or tau by more than
// kFbRoundTol of its scale (max |F'|; max(max |tau|, 2 mu + 3 lambda), the
// stress of a unit strain).  1e-5 is the accuracy the FP32 path has anyway on
// well-separated singular values (its F' and tau land about 2e-6 and 1e-5 of
// their scale from the FP64 map's there): below it the rotation's error
// disappears in that.  Elsewhere rounding does not decide the result:
// an F within rounding of a multiple of the identity (a body at rest or in
// free fall) has all three pairs close, but there F' and tau are within
// rounding of multiples of the identity too, and commute with every K_ij.
constexpr float kFbRound = 2.3841858e-7f;  // 2^-22
constexpr float kFbRoundTol = 1e-5f;

// rotation angles of the pairs (0, 1), (1, 2), (0, 2); 0 for a pair whose
// gap exceeds gap_rel (an equal pair: 1)
__device__ __forceinline__ void fb_pair_angles(const float sig[3], float gap_rel, float th[3]) {
  const float s0 = sig[0] * sig[0], s1 = sig[1] * sig[1], s2 = sig[2] * sig[2];
  const float g[3] = {s0 - s1, s1 - s2, s0 - s2};
#pragma unroll
  for (int k = 0; k < 3; k++)
    th[k] = g[k] <= gap_rel * s0 ? fminf(1.f, kFbRound * s0 * hw_rcp(fmaxf(g[k], 0.f))) : 0.f;
}

// max |(X K_ij - K_ij X)(a, b)| with X(a, b) = X.m[b][a]
template <int I, int J>
__device__ __forceinline__ float comm_pair(const m3& X) {
  float c = 0.f;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const float xk = (b == J ? X.m[I][a] : 0.f) - (b == I ? X.m[J][a] : 0.f);
      const float kx = (a == I ? X.m[b][J] : 0.f) - (a == J ? X.m[b][I] : 0.f);
      c = fmaxf(c, fabsf(xk - kx));
    }
  return c;
}

__device__ __forceinline__ float m3_maxabs(const m3& X) {
  float c = 0.f;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) c = fmaxf(c, fabsf(X.m[a][b]));
  return c;
}

// Whether rounding decides the FP32 result of an SVD-based update (see
// above).  sig: that SVD's singular values; cs: corrected convention;
// with_F: F' = Fp was recomposed from this SVD (projected log strains eps_F);
// with_tau: tau was formed from this SVD's U (projected log strains eps_T).
__device__ __forceinline__ bool fb_rounding_decides(const float sig[3], float gap_rel, float zero_rel, bool cs,
                                                    bool with_F, const m3& Fp, const float eps_F[3], bool with_tau,
                                                    const m3& tau, const float eps_T[3], float mu, float lam) {
  const float s0 = sig[0] * sig[0];
  if (sig[2] * sig[2] <= zero_rel * s0) return true;
  float th[3];
  fb_pair_angles(sig, gap_rel, th);
  if (th[0] == 0.f && th[1] == 0.f && th[2] == 0.f) return false;
  const float tF = kFbRoundTol * m3_maxabs(Fp);
  const float tT = kFbRoundTol * fmaxf(m3_maxabs(tau), 2.f * mu + 3.f * lam);
  constexpr int PI[3] = {0, 1, 0}, PJ[3] = {1, 2, 2};
  bool hit = false;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (th[k] == 0.f) continue;
    float dF, dT;
    if (cs) {
      dF = fabsf(expf(eps_F[PI[k]]) - expf(eps_F[PJ[k]]));
      dT = 2.f * mu * fabsf(eps_T[PI[k]] - eps_T[PJ[k]]);
    } else {
      dF = k == 0 ? comm_pair<0, 1>(Fp) : k == 1 ? comm_pair<1, 2>(Fp) : comm_pair<0, 2>(Fp);
      dT = k == 0 ? comm_pair<0, 1>(tau) : k == 1 ? comm_pair<1, 2>(tau) : comm_pair<0, 2>(tau);
    }
    hit = hit || (with_F && th[k] * dF > tF) || (with_tau && th[k] * dT > tT);
  }
  return hit;
}

// ---- Hencky Kirchhoff stress (shaders/mpm_common.glsl:284-293) ----
// tau = U diag(lambda tr(eps) + 2 mu eps) U^T with U as the GLSL stores it;
// eps_i = log(max(|sigma_i|, 1e-6)).
template <typename T>
__device__ __forceinline__ mat3<T> hencky_from_log(const mat3<T>& U, T e0, T e1, T e2, T mu, T lam) {
  const T lt = lam * (e0 + e1 + e2), tm = T(2) * mu;
  return m3_udvt(U, lt + tm * e0, lt + tm * e1, lt + tm * e2, U);
}
template <typename T>
__device__ __forceinline__ T log_strain(T s) { return hw_log(tmax(tabs(s), T(1e-6))); }
// log_strain(x) of x = exp(e): FP32 takes the log as the reference's shader
// does; FP64 returns e (or the clamp), which that log returns to within an ulp
__device__ __forceinline__ float log_strain_of_exp(float e, float x) { return log_strain(x); }
__device__ __forceinline__ double log_strain_of_exp(double e, double x) { return x > 1e-6 ? e : log(1e-6); }
template <typename T>
__device__ __forceinline__ mat3<T> hencky_stress(const mat3<T>& F, T mu, T lam, bool corrected) {
  mat3<T> U, V;
  T sig[3];
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
// U, V are the true SVD factors (corrected mode).  sig[] keeps the SVD's
// singular values (the fallback's test).
template <typename T>
struct PlasticOutT {
  mat3<T> F;
  mat3<T> U;
  T eps[3];
  T sig[3];
  bool recomposed;  // false: F was returned as given, and U, eps are its own SVD's
};
using PlasticOut = PlasticOutT<float>;
// The return map on a given SVD of F (U, sig, V as svd3 returns them).
template <typename T>
__device__ __forceinline__ PlasticOutT<T> project_plastic_svd(const mat3<T>& F, const mat3<T>& U, const T sig[3],
                                                              const mat3<T>& V, bool dp, T mu, T lam, T q_max, T M,
                                                              T q_c) {
  T e0 = log_strain(sig[0]), e1 = log_strain(sig[1]), e2 = log_strain(sig[2]);
  const T tr = e0 + e1 + e2;
  const T m = tr / T(3);
  T d0 = e0 - m, d1 = e1 - m, d2 = e2 - m;
  const T nd = hw_sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  bool recompose = dp;
  T q_lim = q_max;
  bool apex = false;
  if (dp) {
    const T K = lam + T(2) * mu / T(3);
    const T q_yield = M * (-K * tr) + q_c;
    apex = q_yield <= T(1e-8);
    q_lim = q_yield;
    if (apex) {
      const T m2 = (q_c / tmax(M, T(1e-8)) / K) / T(3);  // tr2 = -p_proj / K, p_proj = -q_c / M
      e0 = m2; e1 = m2; e2 = m2;
    }
  }
  if (!apex) {
    const T delta = nd - q_lim / (T(2) * mu);
    if (delta > T(0) && nd > T(1e-8)) {
      const T k = tdiv(delta, nd);
      e0 = d0 - k * d0 + m;
      e1 = d1 - k * d1 + m;
      e2 = d2 - k * d2 + m;
      recompose = true;
    }
  }
  PlasticOutT<T> o;
  o.U = U;
  o.recomposed = recompose;
  o.sig[0] = sig[0];
  o.sig[1] = sig[1];
  o.sig[2] = sig[2];
  if (recompose) {
    const T x0 = hw_exp(e0), x1 = hw_exp(e1), x2 = hw_exp(e2);
    o.F = m3_udvt(U, x0, x1, x2, V);
    o.eps[0] = log_strain_of_exp(e0, x0);
    o.eps[1] = log_strain_of_exp(e1, x1);
    o.eps[2] = log_strain_of_exp(e2, x2);
  } else {
    o.F = F;
    o.eps[0] = e0;
    o.eps[1] = e1;
    o.eps[2] = e2;
  }
  return o;
}
template <typename T>
__device__ __forceinline__ PlasticOutT<T> project_plastic(const mat3<T>& F, bool dp, T mu, T lam, T q_max, T M, T q_c,
                                                          bool corrected) {
  mat3<T> U, V;
  T sig[3];
  svd3(F, U, sig, V, corrected);
  return project_plastic_svd(F, U, sig, V, dp, mu, lam, q_max, M, q_c);
}

}  // namespace nb
