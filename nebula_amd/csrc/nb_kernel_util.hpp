// nb_kernel_util.hpp — device helpers shared by the substep kernels
// (mpm_kernels.hip) and their adjoints (mpm_adjoint.hip): tile coordinates,
// block scans, the XCD-aware chunk map, cell keys, the per-axis B-spline
// stencil, and the SDF / analytic-collider boundary functions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nb_device_math.hpp"
#include "nb_engine.hpp"

namespace nb {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tile_coords(uint32_t T, int& tx, int& ty, int& tz) {
  tx = int(morton_compact(T));
  ty = int(morton_compact(T >> 1));
  tz = int(morton_compact(T >> 2));
}

// Exclusive scan across a block of BS threads (BS a multiple of 64) of a
// 32- or 64-bit unsigned value.  s_wave must hold BS/64 + 1 values.  Returns
// the exclusive prefix; *total is the block sum.  Contains barriers: call from
// uniform control flow.
template <int BS, class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* s_wave, T& total) {
  constexpr int NW = BS / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  if (wave == 0) {
    T w = lane < NW ? s_wave[lane] : T(0);
    T ws = w;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      T y = __shfl_up(ws, d);
      if (lane >= d) ws += y;
    }
    if (lane < NW) s_wave[lane] = ws - w;
    if (lane == NW - 1) s_wave[NW] = ws;
  }
  __syncthreads();
  T excl = x - v + s_wave[wave];
  total = s_wave[NW];
  __syncthreads();
  return excl;
}

// Workgroups are dealt to the 8 XCDs (each with its own L2) round-robin by
// id.  Block b is given chunk xcd_chunk(b, nb) of the work, so that each XCD
// walks one contiguous eighth of the chunks: neighbouring particles, and the
// grid nodes they share, stay in one L2.  A bijection of [0, nb).
constexpr uint32_t kXCDs = 8;
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb) {
  const uint32_t x = b % kXCDs, i = b / kXCDs, q = nb / kXCDs, r = nb % kXCDs;
  return x < r ? x * (q + 1u) + i : r * (q + 1u) + (x - r) * q + i;
}

// ---- scene geometry ----
// the plain engine's one scene, from the kernel arguments (stays in scalar registers)
__device__ __forceinline__ SceneGeo geo_of(const SubParams& p) {
  SceneGeo g;
  g.h = p.h;
  g.inv_h = 1.f / p.h;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    g.wmin[a] = p.wmin[a];
    g.wmax[a] = p.wmax[a];
    g.grid_res[a] = p.grid_res[a];
  }
  g.cube_res = p.cube_res;
  g.tile_dim = p.tile_dim;
  g.key_base = 0u;
  g.h_pow2 = p.h_pow2;
  g.pad = 0u;
  return g;
}
// the scene of tile T (uniform across a block) and of particle q
template <bool BATCH>
__device__ __forceinline__ SceneGeo tile_geo(const SubParams& p, uint32_t T) {
  if constexpr (BATCH) return p.scenes[p.tile_scene[T]];
  else return geo_of(p);
}
template <bool BATCH>
__device__ __forceinline__ SceneGeo particle_geo(const SubParams& p, uint32_t q) {
  if constexpr (BATCH) return p.scenes[p.pscene[q]];
  else return geo_of(p);
}
// tile coordinates within the scene's cube, and the tile id of coordinates
__device__ __forceinline__ void tile_coords(const SceneGeo& g, uint32_t T, int& tx, int& ty, int& tz) {
  tile_coords(T - (g.key_base >> kTileBits), tx, ty, tz);
}
__device__ __forceinline__ uint32_t tile_id(const SceneGeo& g, int tx, int ty, int tz) {
  return (g.key_base >> kTileBits) + morton3(uint32_t(tx), uint32_t(ty), uint32_t(tz));
}
// index of node (x, y, z) of the scene in the Morton-ordered node arrays
__device__ __forceinline__ uint32_t node_id(const SceneGeo& g, uint32_t x, uint32_t y, uint32_t z) {
  return g.key_base + morton3(x, y, z);
}

// (x - wmin[a]) / h, the IEEE quotient the reference's cell index takes
// (shaders/mpm_common.glsl:190-195, mpm_g2p.comp:78).  When h is a power of
// two the product with 1/h is the same number (an exact scaling), one
// instruction instead of the division's sequence.
__device__ __forceinline__ float cell_coord(const SceneGeo& g, float x, int a) {
  const float d = x - g.wmin[a];
  if (g.h_pow2) return d * g.inv_h;
  return d / g.h;
}

// cell key of a position: shaders/mpm_common.glsl:190-195 (IEEE quotient,
// clamp to [0, gridRes-1], truncation; NaN maps to 0 as fmaxf returns the
// non-NaN operand), offset by the scene's key base
__device__ __forceinline__ uint32_t cell_key(const SceneGeo& g, float x, float y, float z) {
  float lx = fminf(fmaxf(cell_coord(g, x, 0), 0.f), float(g.grid_res[0]) - 1.f);
  float ly = fminf(fmaxf(cell_coord(g, y, 1), 0.f), float(g.grid_res[1]) - 1.f);
  float lz = fminf(fmaxf(cell_coord(g, z, 2), 0.f), float(g.grid_res[2]) - 1.f);
  return g.key_base + morton3(uint32_t(lx), uint32_t(ly), uint32_t(lz));
}


// Per-axis stencil data of a particle relative to the three nodes c-1, c, c+1.
struct Axis3 {
  float w[3], g[3], dx[3];  // weight, weight derivative / h, x_node - x_p
};

// smallest positive entry (+inf if none)
__device__ __forceinline__ float min_positive3(const float w[3]) {
  const float a = w[0] > 0.f ? w[0] : INFINITY, b = w[1] > 0.f ? w[1] : INFINITY, c = w[2] > 0.f ? w[2] : INFINITY;
  return fminf(a, fminf(b, c));
}

// Quadratic B-spline weight and derivative (shaders/mpm_common.glsl:198-210),
// evaluated with selects instead of branches; inv_h = 1/h replaces the
// division of the derivative by h.
__device__ __forceinline__ void axis_stencil_f(float f, float xp, float wmin, float h, float inv_h, int first_node,
                                               Axis3& a) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float nf = float(first_node + i) + 0.5f;
    const float d = f - nf;
    const float ad = fabsf(d);
    const float t = 1.5f - ad;
    const bool in = ad < 0.5f, mid = ad < 1.5f;
    a.w[i] = in ? 0.75f - ad * ad : (mid ? 0.5f * t * t : 0.f);
    a.g[i] = (in ? -2.f * d : (mid ? (d > 0.f ? -t : t) : 0.f)) * inv_h;
    a.dx[i] = nf * h + wmin - xp;
  }
}
// The same stencil for a particle known to lie in cell c (c - 1 the first
// node): when f is in [c, c+1) and c >= 1, the differences f - (c + i - 0.5)
// are exact and fall in [0.5, 1.5), [-0.5, 0.5) and [-1.5, -0.5), so each
// node's piece of the B-spline is fixed and the weights and derivatives come
// out with the bits axis_stencil_f gives, without its compares and selects
// (at the ends of the ranges both pieces give the same values: the spline
// is C1).  Returns false when f is not in that range; a is then not valid.
__device__ __forceinline__ bool axis_stencil_cell(float f, float xp, float wmin, float h, float inv_h, int c,
                                                  Axis3& a) {
  const float n0 = float(c - 1) + 0.5f, n1 = float(c) + 0.5f, n2 = float(c + 1) + 0.5f;
  const float d0 = f - n0, d1 = f - n1, d2 = f - n2;
  const float t0 = 1.5f - d0, t2 = 1.5f + d2;  // 1.5 - |d|
  a.w[0] = 0.5f * t0 * t0;
  a.g[0] = -t0 * inv_h;
  a.w[1] = 0.75f - d1 * d1;
  a.g[1] = (-2.f * d1) * inv_h;
  a.w[2] = 0.5f * t2 * t2;
  a.g[2] = t2 * inv_h;
  a.dx[0] = n0 * h + wmin - xp;
  a.dx[1] = n1 * h + wmin - xp;
  a.dx[2] = n2 * h + wmin - xp;
  return (c >= 1) & (f >= float(c)) & (f < float(c + 1));
}
// f = (x - wmin) / h, the IEEE quotient the cell keys use
__device__ __forceinline__ void axis_stencil(float xp, float wmin, float h, float inv_h, int first_node, Axis3& a) {
  axis_stencil_f((xp - wmin) / h, xp, wmin, h, inv_h, first_node, a);
}

// the dense SDF (plain engine only: Morton order over its cube)
__device__ __forceinline__ float sample_sdf(const float* sdf, const SceneGeo& g, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= g.grid_res[0] || y >= g.grid_res[1] || z >= g.grid_res[2]) return 1e6f;
  return sdf[morton3(x, y, z)];
}

__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// ---- procedural noise of the turbulence and noise forces ----
// shaders/force_noise.glsl:7-84 with GLSL's fract(x) = x - floor(x),
// mix(a, b, t) = a + (b - a) t and dot summed left to right.  The hash
// fract(sin(a) 43758.5453) turns one rounding step of a into a different
// value, so these functions are compiled without FMA contraction, every
// product and sum rounded as written.  GLSL leaves sin's accuracy to the
// driver; here the hash's sin is the correctly rounded FP32 sine, taken as
// sin in FP64 rounded once to FP32 (any two FP64 sines within an ulp round
// to the same float unless the value lies within 2^-29 ulp of a rounding
// boundary), so every platform that does the same gets the same hash.
__device__ __forceinline__ float hash_sin(float a) { return float(sin(double(a))); }
__device__ __forceinline__ float gl_fract(float x) { return x - floorf(x); }
__device__ __forceinline__ float gl_mix(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ float3 noise_hash33(float x, float y, float z) {  // force_noise.glsl:7-12
#pragma clang fp contract(off)
  const float a = x * 127.1f + y * 311.7f + z * 74.7f;
  const float b = x * 269.5f + y * 183.3f + z * 246.1f;
  const float c = x * 113.5f + y * 271.9f + z * 124.6f;
  return make_float3(-1.f + 2.f * gl_fract(hash_sin(a) * 43758.5453123f),
                     -1.f + 2.f * gl_fract(hash_sin(b) * 43758.5453123f),
                     -1.f + 2.f * gl_fract(hash_sin(c) * 43758.5453123f));
}
__device__ inline float noise_value3(float px, float py, float pz) {  // force_noise.glsl:14-37
#pragma clang fp contract(off)
  const float ix = floorf(px), iy = floorf(py), iz = floorf(pz);
  const float fx = gl_fract(px), fy = gl_fract(py), fz = gl_fract(pz);
  const float ux = fx * fx * (3.f - 2.f * fx), uy = fy * fy * (3.f - 2.f * fy), uz = fz * fz * (3.f - 2.f * fz);
  float n[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const float ox = float(c & 1), oy = float((c >> 1) & 1), oz = float((c >> 2) & 1);
    const float3 g = noise_hash33(ix + ox, iy + oy, iz + oz);
    n[c] = g.x * (fx - ox) + g.y * (fy - oy) + g.z * (fz - oz);
  }
  const float nx00 = gl_mix(n[0], n[1], ux), nx10 = gl_mix(n[2], n[3], ux);
  const float nx01 = gl_mix(n[4], n[5], ux), nx11 = gl_mix(n[6], n[7], ux);
  return gl_mix(gl_mix(nx00, nx10, uy), gl_mix(nx01, nx11, uy), uz);
}
__device__ inline float noise_fbm3(float px, float py, float pz, uint32_t octaves) {  // force_noise.glsl:40-50
#pragma clang fp contract(off)
  float sum = 0.f, amp = 0.5f, freq = 1.f;
  const uint32_t n = octaves > 1u ? octaves : 1u;
  for (uint32_t i = 0; i < n; i++) {
    sum += amp * noise_value3(px * freq, py * freq, pz * freq);
    freq *= 2.f;
    amp *= 0.5f;
  }
  return sum;
}
__device__ inline float3 noise_curl3(float px, float py, float pz, uint32_t oct) {  // force_noise.glsl:54-82
#pragma clang fp contract(off)
  const float e = 0.1f;
  const float px1 = noise_fbm3(px + 19.1f, py + e, pz, oct), px2 = noise_fbm3(px + 19.1f, py - e, pz, oct);
  const float py3 = noise_fbm3(px + e, py + 47.3f, pz, oct), py4 = noise_fbm3(px - e, py + 47.3f, pz, oct);
  const float py1 = noise_fbm3(px, py + 47.3f, pz + e, oct), py2 = noise_fbm3(px, py + 47.3f, pz - e, oct);
  const float pz3 = noise_fbm3(px, py + e, pz + 71.7f, oct), pz4 = noise_fbm3(px, py - e, pz + 71.7f, oct);
  const float px3 = noise_fbm3(px + 19.1f, py, pz + e, oct), px4 = noise_fbm3(px + 19.1f, py, pz - e, oct);
  const float pz1 = noise_fbm3(px + e, py, pz + 71.7f, oct), pz2 = noise_fbm3(px - e, py, pz + 71.7f, oct);
  const float s = 2.f * e;
  return make_float3((pz3 - pz4) / s - (py1 - py2) / s, (px3 - px4) / s - (pz1 - pz2) / s,
                     (py3 - py4) / s - (px1 - px2) / s);
}

// The grid's force block (src/core/ForceShaderCompiler.cpp:72-93 for the MPM
// grid, typeFlag 0; Force.h:136-256): gravity and constant wind add
// dir strength dt, turbulence its curl noise, noise its fBm along dir; the
// z clamp constrains positions only and does nothing here.  (x, y, z) is
// the node's world position.  NOISE = false compiles the turbulence and
// noise branches out: the caller launches that form when the force list has
// neither (the noise code's registers would otherwise cost every launch
// occupancy).
template <bool NOISE>
__device__ inline void apply_forces(float& vx, float& vy, float& vz, float x, float y, float z,
                                    const nb_force* forces, uint32_t count, float dt) {
  for (uint32_t fi = 0; fi < count; fi++) {
    const nb_force f = forces[fi];
    const uint32_t mask = f.affect_mask;
    if (mask != 0u && (mask & 1u) == 0u) continue;
    if (f.type == NB_FORCE_GRAVITY || f.type == NB_FORCE_CONSTANT_WIND) {
      const float s = f.data[3];
      vx += (f.data[0] * s) * dt;
      vy += (f.data[1] * s) * dt;
      vz += (f.data[2] * s) * dt;
    } else if (NOISE && f.type == NB_FORCE_TURBULENCE) {  // data: strength, frequency, octaves bits, seed
#pragma clang fp contract(off)
      const float fr = f.data[1], sd = f.data[3];
      const float3 c = noise_curl3(x * fr + sd, y * fr + sd, z * fr + sd, __float_as_uint(f.data[2]));
      vx += (c.x * f.data[0]) * dt;
      vy += (c.y * f.data[0]) * dt;
      vz += (c.z * f.data[0]) * dt;
    } else if (NOISE && f.type == NB_FORCE_NOISE) {  // data: dir xyz, strength, frequency, octaves bits, seed
#pragma clang fp contract(off)
      const float fr = f.data[4], sd = f.data[6];
      const float n = noise_fbm3(x * fr + sd, y * fr + sd, z * fr + sd, __float_as_uint(f.data[5]));
      const float k = (n * f.data[3]) * dt;
      vx += f.data[0] * k;
      vy += f.data[1] * k;
      vz += f.data[2] * k;
    }
  }
}

// applyColliderBC — shaders/collider_common.glsl:62-73
__device__ __forceinline__ void collider_bc(float& vx, float& vy, float& vz, float nx, float ny, float nz,
                                            const nb_collider& c) {
  float rx = vx - c.vx, ry = vy - c.vy, rz = vz - c.vz;
  float vn = rx * nx + ry * ny + rz * nz;
  if (vn < 0.f) {
    float nvx = vn * nx, nvy = vn * ny, nvz = vn * nz;
    float tx = rx - nvx, ty = ry - nvy, tz = rz - nvz;
    float tl = len3(tx, ty, tz);
    if (tl > 1e-8f) {
      float sc = fmaxf(0.f, 1.f - c.friction * fabsf(vn) / tl);
      tx *= sc; ty *= sc; tz *= sc;
    }
    vx = c.vx + tx - c.restitution * nvx;
    vy = c.vy + ty - c.restitution * nvy;
    vz = c.vz + tz - c.restitution * nvz;
  }
}

// analytic SDFs — shaders/collider_common.glsl:22-58; returns sdf, writes normal
__device__ __forceinline__ float collider_sdf(const nb_collider& c, float x, float y, float z, float& nx, float& ny,
                                              float& nz) {
  if (c.type == NB_COLL_SPHERE) {
    float dx = x - c.px, dy = y - c.py, dz = z - c.pz;
    float dist = len3(dx, dy, dz);
    if (dist > 1e-8f) { nx = dx / dist; ny = dy / dist; nz = dz / dist; } else { nx = 0.f; ny = 1.f; nz = 0.f; }
    return dist - c.radius;
  } else if (c.type == NB_COLL_BOX) {
    float rx = x - c.px, ry = y - c.py, rz = z - c.pz;
    float qx = fabsf(rx) - c.nx, qy = fabsf(ry) - c.ny, qz = fabsf(rz) - c.nz;
    float px = fmaxf(qx, 0.f), py = fmaxf(qy, 0.f), pz = fmaxf(qz, 0.f);
    float ql = len3(px, py, pz);
    float sdf = ql + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.f);
    if (sdf < 0.f) {
      float ex = c.nx - fabsf(rx), ey = c.ny - fabsf(ry), ez = c.nz - fabsf(rz);
      nx = ny = nz = 0.f;
      if (ex < ey && ex < ez) nx = glsl_sign(rx + 1e-8f);
      else if (ey < ez) ny = glsl_sign(ry + 1e-8f);
      else nz = glsl_sign(rz + 1e-8f);
    } else {
      float sx = glsl_sign(rx + 1e-8f), sy = glsl_sign(ry + 1e-8f), sz = glsl_sign(rz + 1e-8f);
      float ux, uy, uz;
      if (ql > 1e-8f) { ux = px * sx; uy = py * sy; uz = pz * sz; } else { ux = sx; uy = sy; uz = sz; }
      float ul = len3(ux, uy, uz);
      nx = ux / ul; ny = uy / ul; nz = uz / ul;
    }
    return sdf;
  } else if (c.type == NB_COLL_CAPSULE) {
    float al = len3(c.nx, c.ny, c.nz);
    float adx, ady, adz;
    if (al > 1e-8f) { adx = c.nx / al; ady = c.ny / al; adz = c.nz / al; } else { adx = 0.f; ady = 1.f; adz = 0.f; }
    float t = fminf(fmaxf((x - c.px) * adx + (y - c.py) * ady + (z - c.pz) * adz, 0.f), al);
    float dx = x - (c.px + t * adx), dy = y - (c.py + t * ady), dz = z - (c.pz + t * adz);
    float dist = len3(dx, dy, dz);
    if (dist > 1e-8f) { nx = dx / dist; ny = dy / dist; nz = dz / dist; } else { nx = 0.f; ny = 1.f; nz = 0.f; }
    return dist - c.radius;
  }
  // plane
  float l = len3(c.nx, c.ny, c.nz);
  nx = c.nx / l; ny = c.ny / l; nz = c.nz / l;
  return (x - c.px) * nx + (y - c.py) * ny + (z - c.pz) * nz;
}

}  // namespace nb
