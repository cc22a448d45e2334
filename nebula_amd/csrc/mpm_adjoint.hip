// mpm_adjoint.hip — reverse mode of the MPM substep (G2P, grid update, P2G),
// for the proper polar SVD semantics, PIC/APIC transfers, elastic, Von Mises
// and Drucker-Prager materials, forces, walls and the dense-SDF boundary.
//
// Per substep s (state x, v, F, B, tau -> x', v', F', B', tau'), given the
// adjoint of the state after the substep, the kernels run in this order on
// the bins of s (rebuilt from its kept keys) and the grid the recorded
// forward kept for s (velocities after the update; momentum / mass and mass
// before it):
//   k_g2p_bwd      per particle: the G2P chain backwards (advection, clamps,
//                  push-out, J reset, return map + Hencky adjoint, F update);
//                  writes d/dF, the position part from G2P's weights, and the
//                  coefficients of dL/dv_node: w A + w M_B d + M_G grad w
//   k_adj_scatter  the coefficients scattered to a grid adjoint, tile by
//                  tile like P2G (register sums, fixed-order LDS merge)
//   k_grid_bwd     per node: the grid update backwards -> dL/dp, dL/dm
//   k_p2g_bwd      per particle: P2G backwards as a 27-node gather
//                  -> d/dv, d/dB, d/dtau and the position part from P2G
//   k_pgrad_chunks / k_pgrad_total  once per backward: the material-parameter
//                  partials k_g2p_bwd's blocks accumulated over the substeps,
//                  summed in a fixed order
// Positions enter every weight, so both transfers contribute the weight
// gradients and Hessians.  Everything is deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nb_adjoint_math.hpp"
#include "nb_device_math.hpp"
#include "nb_engine.hpp"
#include "nb_kernel_util.hpp"

namespace nb {

// Per-axis stencil with second derivatives: weight, d/dx, d2/dx2, x_node - x_p.
struct AxisD {
  float w[3], g[3], hh[3], dx[3];
};

// f = (xp - wmin) / h as cell_coord gives it
__device__ __forceinline__ void axis_stencil_d(float f, float xp, float wmin, float h, int first_node, AxisD& a) {
  const float inv_h = 1.f / h;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float nf = float(first_node + i) + 0.5f;
    const float d = f - nf;
    const float ad = fabsf(d);
    const float t = 1.5f - ad;
    const bool in = ad < 0.5f, mid = ad < 1.5f;
    a.w[i] = in ? 0.75f - ad * ad : (mid ? 0.5f * t * t : 0.f);
    a.g[i] = (in ? -2.f * d : (mid ? (d > 0.f ? -t : t) : 0.f)) * inv_h;
    a.hh[i] = (in ? -2.f : (mid ? 1.f : 0.f)) * inv_h * inv_h;
    a.dx[i] = nf * h + wmin - xp;
  }
}

// weight, gradient and Hessian of node (ix, iy, iz)
struct NodeW {
  float w, dw[3], H[3][3], d[3];
};
__device__ __forceinline__ void node_w(const AxisD& X, const AxisD& Y, const AxisD& Z, int ix, int iy, int iz, NodeW& n) {
  const float wx = X.w[ix], wy = Y.w[iy], wz = Z.w[iz];
  const float gx = X.g[ix], gy = Y.g[iy], gz = Z.g[iz];
  n.w = wx * wy * wz;
  n.dw[0] = gx * wy * wz;
  n.dw[1] = wx * gy * wz;
  n.dw[2] = wx * wy * gz;
  n.H[0][0] = X.hh[ix] * wy * wz;
  n.H[1][1] = wx * Y.hh[iy] * wz;
  n.H[2][2] = wx * wy * Z.hh[iz];
  n.H[0][1] = n.H[1][0] = gx * gy * wz;
  n.H[0][2] = n.H[2][0] = gx * wy * gz;
  n.H[1][2] = n.H[2][1] = wx * gy * gz;
  n.d[0] = X.dx[ix];
  n.d[1] = Y.dx[iy];
  n.d[2] = Z.dx[iz];
}

// The transfers' weight cut as a bit per stencil node (bit (z 3 + y) 3 + x
// set: node dropped) for per-axis weights wx, wy, wz: on a particle with a
// node of 0 < w < limit (the smallest positive per-axis weights bound every
// w from below), the nodes whose weight (wx wy) wz, multiplied in the order
// of the reference and of the forward kernels, is below 1e-12.  Zero for
// every other particle.
__device__ __forceinline__ uint32_t cut_mask(const float wx[3], const float wy[3], const float wz[3], float limit) {
  uint32_t m = 0u;
  if ((min_positive3(wx) * min_positive3(wy)) * min_positive3(wz) < limit) {
#pragma unroll
    for (int kz = 0; kz < 3; kz++)
#pragma unroll
      for (int jy = 0; jy < 3; jy++)
#pragma unroll
        for (int ix = 0; ix < 3; ix++) {
          const float w = (wx[ix] * wy[jy]) * wz[kz];
          if (w < 1e-12f) m |= 1u << ((kz * 3 + jy) * 3 + ix);
        }
  }
  return m;
}

// ---- separable gathers ----
// For a field f on the 27 stencil nodes, the backward kernels need the sums
// of f times 22 products of per-axis factors (w, g = w', hh = w'', wd = w d,
// gd = g d with d = x_node - x_p on that axis):
//   kW               w
//   kWD + a          w d_a
//   kG + a           dw/dx_a
//   gd_idx(b, c)     dw/dx_b d_c     (kGD + a on the diagonal)
//   h_idx(a, b)      d2w/dx_a dx_b   (kHD + a on the diagonal)
// The first seven are returned as sums; each of the last fifteen goes to the
// caller's sink(k, z, q) as a term z * q[3] of its sum, so a caller that
// only needs them contracted with per-particle coefficients keeps no sums.
constexpr int kW = 0, kWD = 1, kG = 4, kGD = 7, kHD = 16, kNSum = 7;
__host__ __device__ constexpr int gd_idx(int b, int c) { return b == c ? kGD + b : 10 + 2 * b + (c < b ? c : c - 1); }
__host__ __device__ constexpr int h_idx(int a, int b) { return a == b ? kHD + a : 18 + a + b; }
// axes (b, c) of gd_idx and (a, b), a <= b, of h_idx
__host__ __device__ constexpr int sep_a(int k) {
  return k < 10 ? k - kGD : (k < kHD ? (k - 10) / 2 : (k < 19 ? k - kHD : (k == 21 ? 1 : 0)));
}
__host__ __device__ constexpr int sep_b(int k) {
  return k < 10 ? k - kGD
                : (k < kHD ? ((k - 10) % 2 < (k - 10) / 2 ? (k - 10) % 2 : (k - 10) % 2 + 1)
                           : (k < 19 ? k - kHD : (k == 19 ? 1 : 2)));
}
static_assert(gd_idx(2, 1) == 15 && h_idx(0, 1) == 19 && h_idx(2, 1) == 21, "separable product table");
static_assert(sep_a(gd_idx(1, 2)) == 1 && sep_b(gd_idx(1, 2)) == 2 && sep_a(gd_idx(2, 0)) == 2 &&
                  sep_b(gd_idx(2, 0)) == 0 && sep_a(h_idx(0, 2)) == 0 && sep_b(h_idx(0, 2)) == 2 &&
                  sep_a(h_idx(1, 2)) == 1 && sep_b(h_idx(2, 1)) == 2 && sep_a(h_idx(1, 1)) == 1,
              "separable axis table");

struct AxisF {
  float w[3], g[3], hh[3], wd[3], gd[3], dx[3];
};
__device__ __forceinline__ void axis_factors(const AxisD& a, AxisF& f) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    f.w[i] = a.w[i];
    f.g[i] = a.g[i];
    f.hh[i] = a.hh[i];
    f.wd[i] = a.w[i] * a.dx[i];
    f.gd[i] = a.g[i] * a.dx[i];
    f.dx[i] = a.dx[i];
  }
}

// Axis-by-axis contraction: x (5 factor kinds per row), then y (the 12
// (x, y) pairs the products use), then z.  load(ix, iy, iz) returns the
// node's field in xyz and, when WITH_M, a scalar field in w whose sums with
// dw/dx_a go to Sm.  The transfers' cut of nodes with weight below 1e-12 is
// applied by load returning zero for such a node (every product above then
// drops it): the cut node's weight is negligible, but its second derivative
// is not (w'' = 1/h^2 on the support's edge), so the derivative of the cut
// map leaves it out.  The z loop stays rolled so that one slab
// of 9 node loads is live at a time; its z factors sit in position 0 and are
// rotated at the end of each iteration.
template <bool WITH_M, class Load, class Sink>
__device__ __forceinline__ void sep_gather(const AxisF& X, const AxisF& Y, const AxisF& Z, Load load,
                                           float S[kNSum][3], float Sm[3], Sink sink) {
#pragma unroll
  for (int k = 0; k < kNSum; k++) S[k][0] = S[k][1] = S[k][2] = 0.f;
  Sm[0] = Sm[1] = Sm[2] = 0.f;
  float zw_[3] = {Z.w[0], Z.w[1], Z.w[2]}, zg_[3] = {Z.g[0], Z.g[1], Z.g[2]};
  float zh_[3] = {Z.hh[0], Z.hh[1], Z.hh[2]}, zwd_[3] = {Z.wd[0], Z.wd[1], Z.wd[2]};
  float zgd_[3] = {Z.gd[0], Z.gd[1], Z.gd[2]};
#pragma unroll 1
  for (int iz = 0; iz < 3; iz++) {
    float Q[12][3], Qm[3];
#pragma unroll
    for (int k = 0; k < 12; k++) Q[k][0] = Q[k][1] = Q[k][2] = 0.f;
    Qm[0] = Qm[1] = Qm[2] = 0.f;
#pragma unroll
    for (int iy = 0; iy < 3; iy++) {
      float R[5][3], Rm[2];  // x kinds: w, g, hh, wd, gd
#pragma unroll
      for (int k = 0; k < 5; k++) R[k][0] = R[k][1] = R[k][2] = 0.f;
      Rm[0] = Rm[1] = 0.f;
#pragma unroll
      for (int ix = 0; ix < 3; ix++) {
        const float4 f4 = load(ix, iy, iz);
        const float f[3] = {f4.x, f4.y, f4.z};
#pragma unroll
        for (int r = 0; r < 3; r++) {
          R[0][r] = fmaf(X.w[ix], f[r], R[0][r]);
          R[1][r] = fmaf(X.g[ix], f[r], R[1][r]);
          R[2][r] = fmaf(X.hh[ix], f[r], R[2][r]);
          R[3][r] = fmaf(X.wd[ix], f[r], R[3][r]);
          R[4][r] = fmaf(X.gd[ix], f[r], R[4][r]);
        }
        if (WITH_M) {
          Rm[0] = fmaf(X.w[ix], f4.w, Rm[0]);
          Rm[1] = fmaf(X.g[ix], f4.w, Rm[1]);
        }
      }
      const float yw = Y.w[iy], yg = Y.g[iy], yh = Y.hh[iy], ywd = Y.wd[iy], ygd = Y.gd[iy];
#pragma unroll
      for (int r = 0; r < 3; r++) {
        Q[0][r] = fmaf(yw, R[0][r], Q[0][r]);    // (w, w)
        Q[1][r] = fmaf(yw, R[3][r], Q[1][r]);    // (wd, w)
        Q[2][r] = fmaf(ywd, R[0][r], Q[2][r]);   // (w, wd)
        Q[3][r] = fmaf(yw, R[1][r], Q[3][r]);    // (g, w)
        Q[4][r] = fmaf(yg, R[0][r], Q[4][r]);    // (w, g)
        Q[5][r] = fmaf(yw, R[4][r], Q[5][r]);    // (gd, w)
        Q[6][r] = fmaf(ygd, R[0][r], Q[6][r]);   // (w, gd)
        Q[7][r] = fmaf(ywd, R[1][r], Q[7][r]);   // (g, wd)
        Q[8][r] = fmaf(yg, R[3][r], Q[8][r]);    // (wd, g)
        Q[9][r] = fmaf(yw, R[2][r], Q[9][r]);    // (hh, w)
        Q[10][r] = fmaf(yh, R[0][r], Q[10][r]);  // (w, hh)
        Q[11][r] = fmaf(yg, R[1][r], Q[11][r]);  // (g, g)
      }
      if (WITH_M) {
        Qm[0] = fmaf(yw, Rm[1], Qm[0]);  // (g, w)
        Qm[1] = fmaf(yg, Rm[0], Qm[1]);  // (w, g)
        Qm[2] = fmaf(yw, Rm[0], Qm[2]);  // (w, w)
      }
    }
    const float zw = zw_[0], zg = zg_[0], zh = zh_[0], zwd = zwd_[0], zgd = zgd_[0];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      S[kW][r] = fmaf(zw, Q[0][r], S[kW][r]);
      S[kWD + 0][r] = fmaf(zw, Q[1][r], S[kWD + 0][r]);
      S[kWD + 1][r] = fmaf(zw, Q[2][r], S[kWD + 1][r]);
      S[kWD + 2][r] = fmaf(zwd, Q[0][r], S[kWD + 2][r]);
      S[kG + 0][r] = fmaf(zw, Q[3][r], S[kG + 0][r]);
      S[kG + 1][r] = fmaf(zw, Q[4][r], S[kG + 1][r]);
      S[kG + 2][r] = fmaf(zg, Q[0][r], S[kG + 2][r]);
    }
    sink(gd_idx(0, 0), zw, Q[5]);
    sink(gd_idx(1, 1), zw, Q[6]);
    sink(gd_idx(2, 2), zgd, Q[0]);
    sink(gd_idx(0, 1), zw, Q[7]);
    sink(gd_idx(0, 2), zwd, Q[3]);
    sink(gd_idx(1, 0), zw, Q[8]);
    sink(gd_idx(1, 2), zwd, Q[4]);
    sink(gd_idx(2, 0), zg, Q[1]);
    sink(gd_idx(2, 1), zg, Q[2]);
    sink(h_idx(0, 0), zw, Q[9]);
    sink(h_idx(1, 1), zw, Q[10]);
    sink(h_idx(2, 2), zh, Q[0]);
    sink(h_idx(0, 1), zw, Q[11]);
    sink(h_idx(0, 2), zg, Q[3]);
    sink(h_idx(1, 2), zg, Q[4]);
    if (WITH_M) {
      Sm[0] = fmaf(zw, Qm[0], Sm[0]);
      Sm[1] = fmaf(zw, Qm[1], Sm[1]);
      Sm[2] = fmaf(zg, Qm[2], Sm[2]);
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
      zw_[i] = zw_[i + 1];
      zg_[i] = zg_[i + 1];
      zh_[i] = zh_[i + 1];
      zwd_[i] = zwd_[i + 1];
      zgd_[i] = zgd_[i + 1];
    }
  }
}

__device__ __forceinline__ int adj_model(uint32_t m) {
  return m == NB_MAT_VON_MISES ? 1 : ((m == NB_MAT_DRUCKER_PRAGER || m == NB_MAT_GRANULAR_POWDER) ? 2 : 0);
}

// record <-> matrices (m[col][row]; F columns packed as in the forward record)
__device__ __forceinline__ m3 rec_F(const float4* r) {
  const float4 t = r[R_T], a = r[R_FA], b = r[R_FB];
  m3 F;
  F.m[0][0] = t.w; F.m[0][1] = a.x; F.m[0][2] = a.y;
  F.m[1][0] = a.z; F.m[1][1] = a.w; F.m[1][2] = b.x;
  F.m[2][0] = b.y; F.m[2][1] = b.z; F.m[2][2] = b.w;
  return F;
}
__device__ __forceinline__ m3 rec_B(const float4* r) {
  const float4 b0 = r[R_B0], b1 = r[R_B1], b2 = r[R_B2];
  m3 B;
  B.m[0][0] = b0.x; B.m[0][1] = b0.y; B.m[0][2] = b0.z;
  B.m[1][0] = b1.x; B.m[1][1] = b1.y; B.m[1][2] = b1.z;
  B.m[2][0] = b2.x; B.m[2][1] = b2.y; B.m[2][2] = b2.z;
  return B;
}
// tau as the full symmetric matrix of its six packed values
__device__ __forceinline__ m3 rec_tau(const float4* r) {
  const float4 t = r[R_T], b0 = r[R_B0], b1 = r[R_B1], b2 = r[R_B2];
  m3 T;
  T.m[0][0] = t.x; T.m[1][1] = t.y; T.m[2][2] = t.z;
  T.m[1][0] = T.m[0][1] = b0.w;
  T.m[2][0] = T.m[0][2] = b1.w;
  T.m[2][1] = T.m[1][2] = b2.w;
  return T;
}
// The adjoint of a packed symmetric value enters both of its matrix slots:
// dL/dT_rc = dL/dT_cr = g_packed / 2 off the diagonal.
__device__ __forceinline__ m3 adj_tau_full(const float4* g) {
  const float4 t = g[R_T], b0 = g[R_B0], b1 = g[R_B1], b2 = g[R_B2];
  m3 T;
  T.m[0][0] = t.x; T.m[1][1] = t.y; T.m[2][2] = t.z;
  T.m[1][0] = T.m[0][1] = 0.5f * b0.w;
  T.m[2][0] = T.m[0][2] = 0.5f * b1.w;
  T.m[2][1] = T.m[1][2] = 0.5f * b2.w;
  return T;
}

constexpr int kCoef = 6;  // float4 per particle: A (3), M_B (9), M_G (9)

// ---------------------------------------------------------------------------
// G2P backward
// ---------------------------------------------------------------------------
template <bool BATCH>
__global__ __launch_bounds__(256) void k_g2p_bwd(SubParams p, const float4* __restrict__ rec,
                                                 const float4* __restrict__ adj_in, float4* __restrict__ adj_out,
                                                 float4* __restrict__ coef, float4* __restrict__ pg_acc,
                                                 const nb_material* __restrict__ mats,
                                                 const uint32_t* __restrict__ sorted,
                                                 const float4* __restrict__ grid_node, const float* __restrict__ sdf) {
  const uint32_t k = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool valid = k < p.n;
  const uint32_t q = valid ? sorted[k] : 0u;
  const float4* r = rec + size_t(q) * kRec;
  const SceneGeo sg = particle_geo<BATCH>(p, q);
  float g4[4] = {0.f, 0.f, 0.f, 0.f};
  uint32_t mid = 0;
  if (valid) {
    const float4 pp = r[R_P];
    const float xs[3] = {pp.x, pp.y, pp.z};
    mid = __float_as_uint(r[R_V].w);
    mid = mid > p.material_count - 1u ? p.material_count - 1u : mid;
    const nb_material mat = mats[mid];
    const int model = adj_model(mat.model);
    int base[3];
    AxisF ax[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const float fq = cell_coord(sg, xs[a], a);
      const float f = fminf(fmaxf(fq, -4.f), float(sg.grid_res[a]) + 4.f);
      base[a] = int(floorf(f)) - 1;
      AxisD d;
      axis_stencil_d(fq, xs[a], sg.wmin[a], sg.h, base[a], d);
      axis_factors(d, ax[a]);
    }
    // G2P's weight cut (shaders/mpm_g2p.comp:93-101) as its forward takes it:
    // on a particle with a node of 0 < w < 1e-12, nodes with (wx wy) wz < 1e-12
    const uint32_t cutm = cut_mask(ax[0].w, ax[1].w, ax[2].w, 1e-12f);
    // grid velocity of node (ix, iy, iz); out-of-grid, massless (mass < 1e-10)
    // and cut nodes read as zero (shaders/mpm_g2p.comp:93-101)
    auto node_v = [&](int ix, int iy, int iz) -> float4 {
      const int nx = base[0] + ix, ny = base[1] + iy, nz = base[2] + iz;
      const uint32_t cx = uint32_t(min(max(nx, 0), sg.grid_res[0] - 1));
      const uint32_t cy = uint32_t(min(max(ny, 0), sg.grid_res[1] - 1));
      const uint32_t cz = uint32_t(min(max(nz, 0), sg.grid_res[2] - 1));
      const float4 g = grid_node[node_id(sg, cx, cy, cz)];
      // bitwise tests: the load stays unconditional (one 16-byte load)
      const bool cut = (cutm >> ((iz * 3 + iy) * 3 + ix)) & 1u;
      const bool live = (uint32_t(nx) < uint32_t(sg.grid_res[0])) & (uint32_t(ny) < uint32_t(sg.grid_res[1])) &
                        (uint32_t(nz) < uint32_t(sg.grid_res[2])) & (g.w >= 1e-10f) & !cut;
      return live ? make_float4(g.x, g.y, g.z, 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    // The incoming adjoint of B is needed during the gather: the position
    // adjoint's gradient-times-offset terms are contracted with it there.
    const bool apic = p.flip_ratio < -0.5f;
    const float4* ga = adj_in + size_t(q) * kRec;
    const m3 MB = apic ? rec_B(ga) : m3_zero();
    // one gather serves the forward recompute (shaders/mpm_g2p.comp:87-118)
    // and the position adjoint of the weights
    float S[kNSum][3], Sm_unused[3];
    float gxB[3] = {0.f, 0.f, 0.f};  // sum_i (v_i^T M_B d_i) grad w_i
    float SH[6][3];                  // sum_i H_i v_i (h_idx order)
#pragma unroll
    for (int k = 0; k < 6; k++) SH[k][0] = SH[k][1] = SH[k][2] = 0.f;
    auto sink = [&](int k, float z, const float qv[3]) {
      if (k < kHD) {
        const int b = sep_a(k), c = sep_b(k);
        gxB[b] = fmaf(z, MB.m[c][0] * qv[0] + MB.m[c][1] * qv[1] + MB.m[c][2] * qv[2], gxB[b]);
      } else {
#pragma unroll
        for (int r = 0; r < 3; r++) SH[k - kHD][r] = fmaf(z, qv[r], SH[k - kHD][r]);
      }
    };
    sep_gather<false>(ax[0], ax[1], ax[2], node_v, S, Sm_unused, sink);
    const float vp[3] = {S[kW][0], S[kW][1], S[kW][2]};
    m3 vg, Bn;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int rr = 0; rr < 3; rr++) {
        vg.m[c][rr] = S[kG + c][rr];
        Bn.m[c][rr] = S[kWD + c][rr];
      }
    const m3 F = rec_F(r);
    m3 L;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int rr = 0; rr < 3; rr++) L.m[c][rr] = (c == rr ? 1.f : 0.f) + p.dt * vg.m[c][rr];
    const m3 Ft = m3_mul(L, F);
    PlasticFwd pf;
    plastic_forward(Ft, model, mat.mu, mat.lambda, mat.q_max, mat.M_friction, mat.q_cohesion, pf);
    // det F' = ss0 ss1 ss2 when recomposed from rotations, det F_tmp otherwise
    const float Jd = pf.rm.recompose ? pf.ss[0] * pf.ss[1] * pf.ss[2] : fabsf(m3_det(Ft));
    const bool jreset = Jd < 0.15f || Jd > 10.f;
    float v1[3] = {vp[0], vp[1], vp[2]};
    if (jreset) { v1[0] *= 0.1f; v1[1] *= 0.1f; v1[2] *= 0.1f; }
    float x2[3] = {xs[0] + p.dt * v1[0], xs[1] + p.dt * v1[1], xs[2] + p.dt * v1[2]};
    float v2[3] = {v1[0], v1[1], v1[2]};
    bool pushed = false, push_proj = false, push_stuck = false;
    float u[3] = {0.f, 0.f, 0.f};
    if (sdf) {
      const int gmx[3] = {sg.grid_res[0] - 1, sg.grid_res[1] - 1, sg.grid_res[2] - 1};
      int ic[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const float f = fminf(fmaxf(cell_coord(sg, x2[a], a), -4.f), float(sg.grid_res[a]) + 4.f);
        ic[a] = min(max(int(floorf(f)), 0), gmx[a]);
      }
      const float psdf = sdf[morton3(ic[0], ic[1], ic[2])];
      if (psdf < 0.f) {
        pushed = true;
        const float gnx = sdf[morton3(min(ic[0] + 1, gmx[0]), ic[1], ic[2])] - sdf[morton3(max(ic[0] - 1, 0), ic[1], ic[2])];
        const float gny = sdf[morton3(ic[0], min(ic[1] + 1, gmx[1]), ic[2])] - sdf[morton3(ic[0], max(ic[1] - 1, 0), ic[2])];
        const float gnz = sdf[morton3(ic[0], ic[1], min(ic[2] + 1, gmx[2]))] - sdf[morton3(ic[0], ic[1], max(ic[2] - 1, 0))];
        const float gl = len3(gnx, gny, gnz);
        if (gl > 1e-6f) {
          u[0] = gnx / gl; u[1] = gny / gl; u[2] = gnz / gl;
#pragma unroll
          for (int a = 0; a < 3; a++) x2[a] -= psdf * u[a];
          const float vn = v2[0] * u[0] + v2[1] * u[1] + v2[2] * u[2];
          if (vn < 0.f) {
            push_proj = true;
#pragma unroll
            for (int a = 0; a < 3; a++) v2[a] -= vn * u[a];
          }
        } else {
          push_stuck = true;
#pragma unroll
          for (int a = 0; a < 3; a++) { x2[a] = xs[a]; v2[a] = 0.f; }
        }
      }
    }
    // clamp flags (shaders/mpm_g2p.comp:255-263)
    bool xcl[3], vcl[3];
    {
      const float mg = sg.h * 1.5f;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const float lo = sg.wmin[a] + mg, hi = sg.wmax[a] - mg;
        xcl[a] = false;
        vcl[a] = false;
        if (x2[a] < lo) { x2[a] = lo; xcl[a] = true; if (v2[a] < 0.f) { vcl[a] = true; v2[a] = 0.f; } }
        if (x2[a] > hi) { x2[a] = hi; xcl[a] = true; if (v2[a] > 0.f) { vcl[a] = true; v2[a] = 0.f; } }
      }
    }

    // ---- backward ----
    const float4 gP = ga[R_P], gV = ga[R_V];
    float gx2[3] = {gP.x, gP.y, gP.z}, gv2[3] = {gV.x, gV.y, gV.z};
    m3 gT = adj_tau_full(ga);
    m3 gF2 = rec_F(ga);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (xcl[a]) gx2[a] = 0.f;
      if (vcl[a]) gv2[a] = 0.f;
    }
    float gx[3] = {0.f, 0.f, 0.f};  // d/dx of the state before the substep
    if (pushed) {
      gF2 = m3_zero();
      gT = m3_zero();
      if (push_stuck) {
#pragma unroll
        for (int a = 0; a < 3; a++) { gx[a] += gx2[a]; gx2[a] = 0.f; gv2[a] = 0.f; }
      } else if (push_proj) {
        const float un = gv2[0] * u[0] + gv2[1] * u[1] + gv2[2] * u[2];
#pragma unroll
        for (int a = 0; a < 3; a++) gv2[a] -= un * u[a];
      }
    }
    // x' = x + dt v'
#pragma unroll
    for (int a = 0; a < 3; a++) { gx[a] += gx2[a]; gv2[a] += p.dt * gx2[a]; }
    float A[3];
    if (jreset) {
      gF2 = m3_zero();
      gT = m3_zero();
#pragma unroll
      for (int a = 0; a < 3; a++) A[a] = 0.1f * gv2[a];
    } else {
#pragma unroll
      for (int a = 0; a < 3; a++) A[a] = gv2[a];
    }
    m3 gFt = m3_zero();
    if (!jreset && !pushed) gFt = plastic_hencky_adjoint(pf, mat.mu, mat.lambda, gF2, gT, g4);
    // F_tmp = L F,  L = I + dt grad v
    m3 MG = m3_mult(gFt, F);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int rr = 0; rr < 3; rr++) MG.m[c][rr] *= p.dt;
    const m3 gF = m3_tmul(L, gFt);
    // position through G2P's weights: with v_i the node velocities,
    //   sum_i (v_i.A + v_i^T M_B d_i) grad w_i - w_i M_B^T v_i + H_i M_G^T v_i
#pragma unroll
    for (int b = 0; b < 3; b++) {
      float acc = gxB[b];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        acc = fmaf(A[a], S[kG + b][a], acc);
        acc = fmaf(-MB.m[b][a], S[kW][a], acc);
#pragma unroll
        for (int c = 0; c < 3; c++) acc = fmaf(MG.m[c][a], SH[h_idx(c, b) - kHD][a], acc);
      }
      gx[b] += acc;
    }
    float4* go = adj_out + size_t(q) * kRec;
    go[R_P] = make_float4(gx[0], gx[1], gx[2], 0.f);
    go[R_T] = make_float4(0.f, 0.f, 0.f, gF.m[0][0]);
    go[R_FA] = make_float4(gF.m[0][1], gF.m[0][2], gF.m[1][0], gF.m[1][1]);
    go[R_FB] = make_float4(gF.m[1][2], gF.m[2][0], gF.m[2][1], gF.m[2][2]);
    float cf[kCoef * 4];
    cf[0] = A[0]; cf[1] = A[1]; cf[2] = A[2];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int rr = 0; rr < 3; rr++) {
        cf[3 + 3 * c + rr] = MB.m[c][rr];
        cf[12 + 3 * c + rr] = MG.m[c][rr];
      }
    cf[21] = cf[22] = cf[23] = 0.f;
    float4* co = coef + size_t(q) * kCoef;
#pragma unroll
    for (int j = 0; j < kCoef; j++) co[j] = make_float4(cf[4 * j], cf[4 * j + 1], cf[4 * j + 2], cf[4 * j + 3]);
  }
  // Material-parameter partials: each wave sums its lanes per material (the
  // groups in order of their first lane), and one thread adds the groups of
  // the block's waves, wave by wave, to the block's row of running sums.
  __shared__ uint32_t s_gmid[4][64];
  __shared__ float4 s_gval[4][64];
  __shared__ uint32_t s_gcnt[4];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t n_grp = 0;
  uint64_t pending = __ballot(valid);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lmid = __shfl(mid, leader);
    const bool mine = valid && mid == lmid;
    float s4[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float x = mine ? g4[j] : 0.f;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
      s4[j] = x;
    }
    if (lane == 0) {
      s_gmid[wv][n_grp] = lmid;
      s_gval[wv][n_grp] = make_float4(s4[0], s4[1], s4[2], s4[3]);
    }
    n_grp++;
    pending &= ~__ballot(mine);
  }
  if (lane == 0) s_gcnt[wv] = n_grp;
  __syncthreads();
  if (threadIdx.x == 0) {
    float4* row = pg_acc + size_t(blockIdx.x) * p.material_count;
    for (uint32_t w = 0; w < 4u; w++)
      for (uint32_t i = 0; i < s_gcnt[w]; i++) {
        const float4 v = s_gval[w][i];
        float4& t = row[s_gmid[w][i]];
        t = make_float4(t.x + v.x, t.y + v.y, t.z + v.z, t.w + v.w);
      }
  }
}

// ---------------------------------------------------------------------------
// scatter of the G2P-backward coefficients to the grid adjoint
// ---------------------------------------------------------------------------
constexpr int kAdjThreads = 256;

template <bool BATCH>
__global__ __launch_bounds__(kAdjThreads) void k_adj_scatter(SubParams p, const float4* __restrict__ rec,
                                                             const float4* __restrict__ coef,
                                                             const uint32_t* __restrict__ cell_count,
                                                             const uint32_t* __restrict__ cell_end,
                                                             const uint32_t* __restrict__ sorted,
                                                             const uint32_t* __restrict__ act_tiles,
                                                             const uint32_t* __restrict__ tile_order,
                                                             const uint32_t* __restrict__ counters,
                                                             float4* __restrict__ partial_adj) {
  __shared__ float s_x[kBoxNodes], s_y[kBoxNodes], s_z[kBoxNodes];
  __shared__ uint16_t s_cell[kTileCells];
  __shared__ uint32_t s_wave[kAdjThreads / 64 + 1];
  const uint32_t na = counters[CNT_ACTIVE];
  for (uint32_t a0 = blockIdx.x; a0 < na; a0 += gridDim.x) {
    // largest tiles first, as in the forward's P2G
    const uint32_t ao = tile_order[a0], a = ao < na ? ao : a0;
    const uint32_t T = act_tiles[a];
    const SceneGeo sg = tile_geo<BATCH>(p, T);
    int tx, ty, tz;
    tile_coords(sg, T, tx, ty, tz);
    const size_t tbase = size_t(T) << kTileBits;
    for (int e = threadIdx.x; e < kBoxNodes; e += kAdjThreads) { s_x[e] = 0.f; s_y[e] = 0.f; s_z[e] = 0.f; }
    const uint32_t ca = 2u * threadIdx.x, cb = ca + 1u;
    const bool oa = cell_count[tbase + ca] != 0u, ob = cell_count[tbase + cb] != 0u;
    uint32_t n_occ;
    uint32_t pos = block_exclusive_scan<kAdjThreads>(uint32_t(oa) + uint32_t(ob), s_wave, n_occ);
    if (oa) s_cell[pos++] = uint16_t(ca);
    if (ob) s_cell[pos] = uint16_t(cb);
    __syncthreads();
    for (uint32_t base = 0; base < n_occ; base += kAdjThreads) {
      const uint32_t j = base + threadIdx.x;
      const bool have = j < n_occ;
      const uint32_t c = have ? uint32_t(s_cell[j]) : 0u;
      const int lx = int(morton_compact(c)), ly = int(morton_compact(c >> 1)), lz = int(morton_compact(c >> 2));
      const int cx = tx * 8 + lx, cy = ty * 8 + ly, cz = tz * 8 + lz;
      uint32_t kbeg = 0, kend = 0;
      if (have) {
        kend = cell_end[tbase + c];
        kbeg = kend - cell_count[tbase + c];
      }
      float ax[27], ay[27], az[27];
#pragma unroll
      for (int o = 0; o < 27; o++) { ax[o] = 0.f; ay[o] = 0.f; az[o] = 0.f; }
      // the next particle's position and coefficients are loaded while this
      // one is scattered
      float4 npp, nco[kCoef];
      auto load = [&](uint32_t kk) {
        const uint32_t q = sorted[kk];
        npp = rec[size_t(q) * kRec + R_P];
        const float4* co = coef + size_t(q) * kCoef;
#pragma unroll
        for (int jj = 0; jj < kCoef; jj++) nco[jj] = co[jj];
      };
      if (kbeg < kend) load(kbeg);
      for (uint32_t kk = kbeg; kk < kend; kk++) {
        const float4 pp = npp;
        float cf[kCoef * 4];
#pragma unroll
        for (int jj = 0; jj < kCoef; jj++) {
          const float4 v4 = nco[jj];
          cf[4 * jj] = v4.x; cf[4 * jj + 1] = v4.y; cf[4 * jj + 2] = v4.z; cf[4 * jj + 3] = v4.w;
        }
        if (kk + 1 < kend) load(kk + 1);
        AxisD X, Y, Z;
        axis_stencil_d(cell_coord(sg, pp.x, 0), pp.x, sg.wmin[0], sg.h, cx - 1, X);
        axis_stencil_d(cell_coord(sg, pp.y, 1), pp.y, sg.wmin[1], sg.h, cy - 1, Y);
        axis_stencil_d(cell_coord(sg, pp.z, 2), pp.z, sg.wmin[2], sg.h, cz - 1, Z);
        // dL/dv_node = w (A + M_B d) + M_G grad w, factored per (y, z) row:
        //   = wx P + (wx dx) Q + gx R  with
        //   P = wy wz (A + dy M_B[1] + dz M_B[2]) + gy wz M_G[1] + wy gz M_G[2],
        //   Q = wy wz M_B[0],  R = wy wz M_G[0]
        // G2P's weight cut as its forward takes it: on a particle with a node
        // of 0 < w < 1e-12, the nodes with (wx wy) wz < 1e-12 get nothing.
        const uint32_t cutm = cut_mask(X.w, Y.w, Z.w, 1e-12f);
        float wdx[3];
#pragma unroll
        for (int ix = 0; ix < 3; ix++) wdx[ix] = X.w[ix] * X.dx[ix];
#pragma unroll
        for (int kz = 0; kz < 3; kz++)
#pragma unroll
          for (int jy = 0; jy < 3; jy++) {
            const float wyz = Y.w[jy] * Z.w[kz], gyz = Y.g[jy] * Z.w[kz], wgz = Y.w[jy] * Z.g[kz];
            float P[3], Q[3], R[3];
#pragma unroll
            for (int rr = 0; rr < 3; rr++) {
              const float t = fmaf(Z.dx[kz], cf[9 + rr], fmaf(Y.dx[jy], cf[6 + rr], cf[rr]));
              P[rr] = fmaf(wgz, cf[18 + rr], fmaf(gyz, cf[15 + rr], wyz * t));
              Q[rr] = wyz * cf[3 + rr];
              R[rr] = wyz * cf[12 + rr];
            }
#pragma unroll
            for (int ix = 0; ix < 3; ix++) {
              const int o = (kz * 3 + jy) * 3 + ix;
              const bool cut = (cutm >> o) & 1u;
              ax[o] = cut ? ax[o] : fmaf(X.g[ix], R[0], fmaf(wdx[ix], Q[0], fmaf(X.w[ix], P[0], ax[o])));
              ay[o] = cut ? ay[o] : fmaf(X.g[ix], R[1], fmaf(wdx[ix], Q[1], fmaf(X.w[ix], P[1], ay[o])));
              az[o] = cut ? az[o] : fmaf(X.g[ix], R[2], fmaf(wdx[ix], Q[2], fmaf(X.w[ix], P[2], az[o])));
            }
          }
      }
#pragma unroll
      for (int kz = 0; kz < 3; kz++)
#pragma unroll
        for (int jy = 0; jy < 3; jy++)
#pragma unroll
          for (int ix = 0; ix < 3; ix++) {
            const int o = (kz * 3 + jy) * 3 + ix;
            const int e = ((lz + kz) * kBox + (ly + jy)) * kBox + (lx + ix);
            __syncthreads();
            if (have) { s_x[e] += ax[o]; s_y[e] += ay[o]; s_z[e] += az[o]; }
          }
      __syncthreads();
    }
    float4* out = partial_adj + size_t(a) * kBoxNodes;
    for (int e = threadIdx.x; e < kBoxNodes; e += kAdjThreads) out[e] = make_float4(s_x[e], s_y[e], s_z[e], 0.f);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// grid update backward
// ---------------------------------------------------------------------------
template <bool BATCH, bool NOISE>
__global__ __launch_bounds__(512) void k_grid_bwd(SubParams p, const float4* __restrict__ grid_pre,
                                                  const float4* __restrict__ partial_adj,
                                                  const int32_t* __restrict__ tile_act,
                                                  const uint32_t* __restrict__ tiles, const uint32_t* __restrict__ count,
                                                  float4* __restrict__ grid_adj, const float* __restrict__ sdf,
                                                  const nb_force* __restrict__ forces) {
  const uint32_t nb_tiles = *count;
  const int lx = int(morton_compact(threadIdx.x)), ly = int(morton_compact(threadIdx.x >> 1)),
            lz = int(morton_compact(threadIdx.x >> 2));
  for (uint32_t b = blockIdx.x; b < nb_tiles; b += gridDim.x) {
    const uint32_t N = tiles[b];
    const SceneGeo sg = tile_geo<BATCH>(p, N);
    const int td = int(sg.tile_dim);
    int tx, ty, tz;
    tile_coords(sg, N, tx, ty, tz);
    const int nx = tx * 8 + lx, ny = ty * 8 + ly, nz = tz * 8 + lz;
    const size_t nid = (size_t(N) << kTileBits) | threadIdx.x;
    float gx = 0.f, gy = 0.f, gz = 0.f;
    bool covered = false;
    for (int dz = -1; dz <= 1; dz++) {
      if ((dz < 0 && lz != 0) || (dz > 0 && lz != 7)) continue;
      const int sz = tz + dz;
      if (sz < 0 || sz >= td) continue;
      for (int dy = -1; dy <= 1; dy++) {
        if ((dy < 0 && ly != 0) || (dy > 0 && ly != 7)) continue;
        const int sy = ty + dy;
        if (sy < 0 || sy >= td) continue;
        for (int dx = -1; dx <= 1; dx++) {
          if ((dx < 0 && lx != 0) || (dx > 0 && lx != 7)) continue;
          const int sx = tx + dx;
          if (sx < 0 || sx >= td) continue;
          const int32_t a = tile_act[tile_id(sg, sx, sy, sz)];
          if (a < 0) continue;
          const int bx = lx - 8 * dx + 1, by = ly - 8 * dy + 1, bz = lz - 8 * dz + 1;
          const size_t e = size_t(a) * kBoxNodes + (bz * kBox + by) * kBox + bx;
          const float4 g = partial_adj[e];
          gx += g.x; gy += g.y; gz += g.z;
          covered = true;
        }
      }
    }
    if (!covered) continue;
    // the recorded forward's node state before the update: (momentum / mass, mass)
    const float4 pre = grid_pre[nid];
    const float mass = pre.w;
    if (mass < 1e-10f) {
      grid_adj[nid] = make_float4(0.f, 0.f, 0.f, 0.f);
      continue;
    }
    // forward recompute (shaders/mpm_grid_update.comp:17-101, mpm_nanovdb_bc.comp:23-54)
    const float u0x = pre.x, u0y = pre.y, u0z = pre.z;
    const float sp = len3(u0x, u0y, u0z);
    const bool capped = sp > 15.f;
    float vx = u0x, vy = u0y, vz = u0z;
    if (capped) { const float kk = 15.f / sp; vx *= kk; vy *= kk; vz *= kk; }
    // forces add constants with respect to the state: they pass the adjoint
    // through, but the flags below depend on the velocity they produce
    apply_forces<NOISE>(vx, vy, vz, (float(nx) + 0.5f) * sg.h + sg.wmin[0], (float(ny) + 0.5f) * sg.h + sg.wmin[1],
                        (float(nz) + 0.5f) * sg.h + sg.wmin[2], forces, p.force_count, p.dt);
    const int Nx = sg.grid_res[0] - 1, Ny = sg.grid_res[1] - 1, Nz = sg.grid_res[2] - 1;
    bool wx = false, wy = false, wz = false;
    if (nx <= 1 && vx < 0.f) { vx = 0.f; wx = true; }
    if (nx >= Nx - 1 && vx > 0.f) { vx = 0.f; wx = true; }
    if (ny <= 1 && vy < 0.f) { vy = 0.f; wy = true; }
    if (ny >= Ny - 1 && vy > 0.f) { vy = 0.f; wy = true; }
    if (nz <= 1 && vz < 0.f) { vz = 0.f; wz = true; }
    if (nz >= Nz - 1 && vz > 0.f) { vz = 0.f; wz = true; }
    bool sdf_zero = false, sdf_proj = false;
    float ux = 0.f, uy = 0.f, uz = 0.f;
    if (sdf) {
      const float sd = sample_sdf(sdf, sg, nx, ny, nz);
      if (sd < 0.f) {
        sdf_zero = true;
      } else if (sd < sg.h * 1.5f) {
        const float ggx = sample_sdf(sdf, sg, nx + 1, ny, nz) - sample_sdf(sdf, sg, nx - 1, ny, nz);
        const float ggy = sample_sdf(sdf, sg, nx, ny + 1, nz) - sample_sdf(sdf, sg, nx, ny - 1, nz);
        const float ggz = sample_sdf(sdf, sg, nx, ny, nz + 1) - sample_sdf(sdf, sg, nx, ny, nz - 1);
        const float gl = len3(ggx, ggy, ggz);
        if (gl > 1e-6f) {
          ux = ggx / gl; uy = ggy / gl; uz = ggz / gl;
          if (vx * ux + vy * uy + vz * uz < 0.f) sdf_proj = true;
        }
      }
    }
    // backward
    if (sdf_zero) { gx = gy = gz = 0.f; }
    if (sdf_proj) {
      const float gn = gx * ux + gy * uy + gz * uz;
      gx -= gn * ux; gy -= gn * uy; gz -= gn * uz;
    }
    if (wx) gx = 0.f;
    if (wy) gy = 0.f;
    if (wz) gz = 0.f;
    if (capped) {
      const float hx = u0x / sp, hy = u0y / sp, hz = u0z / sp;
      const float gh = gx * hx + gy * hy + gz * hz;
      const float kk = 15.f / sp;
      gx = kk * (gx - gh * hx); gy = kk * (gy - gh * hy); gz = kk * (gz - gh * hz);
    }
    const float im = 1.f / mass;
    grid_adj[nid] = make_float4(gx * im, gy * im, gz * im, -(gx * u0x + gy * u0y + gz * u0z) * im);
  }
}

// ---------------------------------------------------------------------------
// P2G backward
// ---------------------------------------------------------------------------
template <bool BATCH>
__global__ __launch_bounds__(256) void k_p2g_bwd(SubParams p, const float4* __restrict__ rec,
                                                 float4* __restrict__ adj_out, const nb_material* __restrict__ mats,
                                                 const uint32_t* __restrict__ sorted,
                                                 const float4* __restrict__ grid_adj) {
  const uint32_t k = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (k >= p.n) return;
  const uint32_t q = sorted[k];
  const float4* r = rec + size_t(q) * kRec;
  const SceneGeo sg = particle_geo<BATCH>(p, q);
  const float4 pp = r[R_P], vv = r[R_V];
  uint32_t mid = __float_as_uint(vv.w);
  mid = mid > p.material_count - 1u ? p.material_count - 1u : mid;
  const float mp = mats[mid].rho0 * pp.w;
  const bool apic = p.flip_ratio < -0.5f;
  const float sc = apic ? mp * (4.f / (sg.h * sg.h)) : 0.f;
  const m3 B = rec_B(r);
  const m3 T = rec_tau(r);
  const float st = p.dt * pp.w;
  const float xs[3] = {pp.x, pp.y, pp.z}, v[3] = {vv.x, vv.y, vv.z};
  int cc[3];
  AxisF ax[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    // the bin's cell (shaders/mpm_common.glsl:190-195); P2G's stencil is the cell's 27 nodes
    const float fq = cell_coord(sg, xs[a], a);
    const float f = fminf(fmaxf(fq, 0.f), float(sg.grid_res[a]) - 1.f);
    cc[a] = int(f);
    AxisD d;
    axis_stencil_d(fq, xs[a], sg.wmin[a], sg.h, cc[a] - 1, d);
    axis_factors(d, ax[a]);
  }
  const uint32_t cube = sg.cube_res;
  // P2G's weight cut (shaders/mpm_p2g.comp:60-61) as its forward takes it: a
  // particle whose smallest positive per-axis weights multiply below
  // 1.0001e-12 skips the nodes with (wx wy) wz < 1e-12
  const uint32_t cutm = cut_mask(ax[0].w, ax[1].w, ax[2].w, 1.0001e-12f);
  // grid adjoint (dL/dp, dL/dm) of node (ix, iy, iz); nodes outside the cube
  // and cut nodes read as zero
  auto node_g = [&](int ix, int iy, int iz) -> float4 {
    const uint32_t nx = uint32_t(cc[0] - 1 + ix), ny = uint32_t(cc[1] - 1 + iy), nz = uint32_t(cc[2] - 1 + iz);
    const bool cut = (cutm >> ((iz * 3 + iy) * 3 + ix)) & 1u;
    const bool in = (nx < cube) & (ny < cube) & (nz < cube) & !cut;  // negative wraps to large
    const float4 g = grid_adj[node_id(sg, min(nx, cube - 1u), min(ny, cube - 1u), min(nz, cube - 1u))];
    return in ? g : make_float4(0.f, 0.f, 0.f, 0.f);
  };
  // the gradient-times-offset and Hessian sums enter only the position
  // adjoint: contracted with sc B and -dt Vp tau as they arrive
  float S[kNSum][3], Sm[3];
  float gxo[3] = {0.f, 0.f, 0.f};
  auto sink = [&](int k, float z, const float qv[3]) {
    const int a = sep_a(k), b = sep_b(k);
    if (k < kHD) {  // (b, c) = (a, b): sc sum_r B[c][r] S[gd(b, c)][r] into d/dx_b
      gxo[a] = fmaf(z, sc * (B.m[b][0] * qv[0] + B.m[b][1] * qv[1] + B.m[b][2] * qv[2]), gxo[a]);
    } else {        // -st sum_r tau[c][r] S[h(c, b)][r] into d/dx_b, both orders of (a, b)
      gxo[b] = fmaf(z, -st * (T.m[a][0] * qv[0] + T.m[a][1] * qv[1] + T.m[a][2] * qv[2]), gxo[b]);
      if (a != b) gxo[a] = fmaf(z, -st * (T.m[b][0] * qv[0] + T.m[b][1] * qv[1] + T.m[b][2] * qv[2]), gxo[a]);
    }
  };
  sep_gather<true>(ax[0], ax[1], ax[2], node_g, S, Sm, sink);
  // P2G adds  w (m v + C d) - dt Vp tau grad w  to the node momentum and  w m  to its mass
  float gv[3], gx[3];
  m3 gC, gT;
#pragma unroll
  for (int rr = 0; rr < 3; rr++) {
    gv[rr] = mp * S[kW][rr];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      gC.m[c][rr] = S[kWD + c][rr];
      gT.m[c][rr] = -st * S[kG + c][rr];
    }
  }
  //   d/dx_b = m dL/dm . grad w + (dL/dp . (m v + C d)) grad w - w C^T dL/dp - dt Vp H tau dL/dp
#pragma unroll
  for (int b = 0; b < 3; b++) {
    float acc = fmaf(mp, Sm[b], gxo[b]);
#pragma unroll
    for (int rr = 0; rr < 3; rr++) {
      acc = fmaf(mp * v[rr], S[kG + b][rr], acc);
      acc = fmaf(-sc * B.m[b][rr], S[kW][rr], acc);
    }
    gx[b] = acc;
  }
  float4* go = adj_out + size_t(q) * kRec;
  const float4 g0 = go[R_P];
  const float4 gt = go[R_T];
  go[R_P] = make_float4(g0.x + gx[0], g0.y + gx[1], g0.z + gx[2], 0.f);
  go[R_V] = make_float4(gv[0], gv[1], gv[2], 0.f);
  // packed tau: diagonal, and off-diagonal values that fill two slots each
  go[R_B0] = make_float4(sc * gC.m[0][0], sc * gC.m[0][1], sc * gC.m[0][2], gT.m[1][0] + gT.m[0][1]);
  go[R_B1] = make_float4(sc * gC.m[1][0], sc * gC.m[1][1], sc * gC.m[1][2], gT.m[2][0] + gT.m[0][2]);
  go[R_B2] = make_float4(sc * gC.m[2][0], sc * gC.m[2][1], sc * gC.m[2][2], gT.m[2][1] + gT.m[1][2]);
  go[R_T] = make_float4(gT.m[0][0], gT.m[1][1], gT.m[2][2], gt.w);
}

// Fixed-order sum of the per-block rows in two levels.  Level 1: one wave per
// (chunk of 64 source blocks, material); lane l takes block chunk * 64 + l; a
// fixed shuffle tree sums the lanes into chunk_part[material][chunk].  Level
// 2: one block per material adds its chunk sums (thread t: chunks t, t + 256,
// ..., then a fixed LDS tree).
constexpr uint32_t kPgChunk = 64;
__global__ __launch_bounds__(256) void k_pgrad_chunks(const float4* __restrict__ pg_acc, uint32_t n_blocks,
                                                      uint32_t n_materials, uint32_t n_chunks,
                                                      float4* __restrict__ chunk_part) {
  const uint32_t chunk = blockIdx.x, m = blockIdx.y * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (m >= n_materials) return;  // whole waves
  const uint32_t b = chunk * kPgChunk + lane;
  float4 acc = b < n_blocks ? pg_acc[size_t(b) * n_materials + m] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    acc.x += __shfl_xor(acc.x, o);
    acc.y += __shfl_xor(acc.y, o);
    acc.z += __shfl_xor(acc.z, o);
    acc.w += __shfl_xor(acc.w, o);
  }
  if (lane == 0) chunk_part[size_t(m) * n_chunks + chunk] = acc;
}
__global__ __launch_bounds__(256) void k_pgrad_total(const float4* __restrict__ chunk_part, uint32_t n_chunks,
                                                     float* __restrict__ pgrad) {
  __shared__ float4 s[256];
  const uint32_t m = blockIdx.x;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t c = threadIdx.x; c < n_chunks; c += 256) {
    const float4 v = chunk_part[size_t(m) * n_chunks + c];
    acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) {
      const float4 v = s[threadIdx.x + o];
      s[threadIdx.x].x += v.x; s[threadIdx.x].y += v.y; s[threadIdx.x].z += v.z; s[threadIdx.x].w += v.w;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pgrad[4 * m + 0] = s[0].x;
    pgrad[4 * m + 1] = s[0].y;
    pgrad[4 * m + 2] = s[0].z;
    pgrad[4 * m + 3] = s[0].w;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv_a(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
static inline uint32_t tile_grid_a(uint32_t n, uint32_t cap) {
  const uint32_t c = cap ? cap : 2048u;
  return n < c ? (n ? n : 1u) : c;
}

template <bool BATCH>
static void adjoint_kernels(const DevState& d, const AdjState& a, const SubParams& p, uint32_t n_tiles,
                            const float4* adj_in, float4* adj_out, hipStream_t s) {
  const uint32_t nblk = cdiv_a(p.n, 256);
  k_g2p_bwd<BATCH><<<nblk, 256, 0, s>>>(p, d.rec, adj_in, adj_out, a.coef, a.pg_acc,
                                         d.materials, d.sorted, d.grid_node, d.sdf);
  k_adj_scatter<BATCH><<<tile_grid_a(n_tiles, p.tile_blocks), kAdjThreads, 0, s>>>(p, d.rec, a.coef, d.cell_count, d.cell_end,
                                                                    d.sorted, d.act_tiles, d.tile_order, d.counters,
                                                                    a.partial_adj);
  if (p.noise_forces)
    k_grid_bwd<BATCH, true><<<tile_grid_a(n_tiles, p.tile_blocks), 512, 0, s>>>(
        p, d.grid_pre, a.partial_adj, d.tile_act, d.node_tiles, d.counters + CNT_NODE_TILES, a.grid_adj, d.sdf, d.forces);
  else
    k_grid_bwd<BATCH, false><<<tile_grid_a(n_tiles, p.tile_blocks), 512, 0, s>>>(
        p, d.grid_pre, a.partial_adj, d.tile_act, d.node_tiles, d.counters + CNT_NODE_TILES, a.grid_adj, d.sdf, d.forces);
  k_p2g_bwd<BATCH><<<nblk, 256, 0, s>>>(p, d.rec, adj_out, d.materials, d.sorted, a.grid_adj);
}

void launch_adjoint_substep(const DevState& d, const AdjState& a, const SubParams& p, uint32_t n_tiles,
                            const float4* adj_in, float4* adj_out, hipStream_t s) {
  if (p.n_scenes > 0)
    adjoint_kernels<true>(d, a, p, n_tiles, adj_in, adj_out, s);
  else
    adjoint_kernels<false>(d, a, p, n_tiles, adj_in, adj_out, s);
}

void launch_adjoint_params(const AdjState& a, uint32_t n_particles, uint32_t n_materials, hipStream_t s) {
  if (n_particles == 0 || n_materials == 0) return;  // pgrad stays zero
  const uint32_t nblk = cdiv_a(n_particles, 256);
  const uint32_t n_chunks = cdiv_a(nblk, kPgChunk);
  k_pgrad_chunks<<<dim3(n_chunks, cdiv_a(n_materials, 4u)), 256, 0, s>>>(a.pg_acc, nblk, n_materials, n_chunks,
                                                                          a.pg_chunk);
  k_pgrad_total<<<n_materials, 256, 0, s>>>(a.pg_chunk, n_chunks, a.pgrad);
}

}  // namespace nb
