// Emitter sampling on the host, shared by the engine (nb_mpm_emit) and the
// C++ facade's Emitter::sample: the shapes of src/core/Emitter.h:39-224 over
// one nb_emitter record, drawn with std::mt19937 and
// std::uniform_real_distribution<float>.  Where the reference draws a
// vector's three components inside one constructor call (an order C++ leaves
// to the compiler) they are drawn x, y, z here.
#pragma once

#include <algorithm>
#include <cmath>
#include <random>

#include "nebula_mpm.h"

namespace nebula {

// ---- emitter sampling (src/core/Emitter.h) ----
// A point of the emitter's shape: AABB uniform in the box (Emitter.h:100-107),
// sphere r = R cbrt(u), theta, phi uniform (:147-155), ellipse by rejection in
// the x-y bounding box with z at the center (:198-208).
inline void emitter_sample(const nb_emitter& e, std::mt19937& rng, float& x, float& y, float& z) {
  const float cx = e.center[0], cy = e.center[1], cz = e.center[2];
  if (e.shape == NB_EMITTER_AABB) {
    const float hx = e.size[0] * 0.5f, hy = e.size[1] * 0.5f, hz = e.size[2] * 0.5f;
    std::uniform_real_distribution<float> dx(-hx, hx), dy(-hy, hy), dz(-hz, hz);
    const float ox = dx(rng);
    const float oy = dy(rng);
    const float oz = dz(rng);
    x = cx + ox; y = cy + oy; z = cz + oz;
  } else if (e.shape == NB_EMITTER_SPHERE) {
    std::uniform_real_distribution<float> dr(0.0f, 1.0f), dt(0.0f, 3.14159265f), dp(0.0f, 6.28318530f);
    const float r = e.size[0] * std::cbrt(dr(rng));
    const float th = dt(rng);
    const float ph = dp(rng);
    x = cx + r * std::sin(th) * std::cos(ph);
    y = cy + r * std::sin(th) * std::sin(ph);
    z = cz + r * std::cos(th);
  } else {
    const float a = e.size[0], b = e.size[1];
    std::uniform_real_distribution<float> dx(-a, a), dy(-b, b);
    float ex, ey;
    do {
      ex = dx(rng);
      ey = dy(rng);
    } while ((ex * ex) / (a * a) + (ey * ey) / (b * b) > 1.0f);
    x = cx + ex; y = cy + ey; z = cz;
  }
}

// Emitter::sample_velocity (Emitter.h:66-74) with the outward normals of
// Emitter::surface_normal (:59-63; the box's own at :109-117).
inline void emitter_velocity(const nb_emitter& e, float px, float py, float pz, std::mt19937& rng, float& vx, float& vy,
                      float& vz) {
  vx = e.vel[0]; vy = e.vel[1]; vz = e.vel[2];
  if (e.emit_along_normal) {
    float nx, ny, nz;
    const float dx = px - e.center[0], dy = py - e.center[1], dz = pz - e.center[2];
    if (e.shape == NB_EMITTER_AABB) {
      const float hx = std::max(e.size[0] * 0.5f, 1e-6f), hy = std::max(e.size[1] * 0.5f, 1e-6f),
                  hz = std::max(e.size[2] * 0.5f, 1e-6f);
      const float ux = dx / hx, uy = dy / hy, uz = dz / hz;
      const float ax = std::fabs(ux), ay = std::fabs(uy), az = std::fabs(uz);
      nx = ny = nz = 0.0f;
      if (ax >= ay && ax >= az) nx = ux >= 0.0f ? 1.0f : -1.0f;
      else if (ay >= az) ny = uy >= 0.0f ? 1.0f : -1.0f;
      else nz = uz >= 0.0f ? 1.0f : -1.0f;
    } else {
      const float len = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (len > 1e-6f) { nx = dx / len; ny = dy / len; nz = dz / len; }
      else { nx = 0.0f; ny = 0.0f; nz = 1.0f; }
    }
    const float speed = std::sqrt(e.vel[0] * e.vel[0] + e.vel[1] * e.vel[1] + e.vel[2] * e.vel[2]);
    vx = nx * speed; vy = ny * speed; vz = nz * speed;
  }
  if (e.velocity_randomness > 0.0f) {
    std::uniform_real_distribution<float> d(-e.velocity_randomness, e.velocity_randomness);
    const float rx = d(rng);
    const float ry = d(rng);
    const float rz = d(rng);
    vx += rx; vy += ry; vz += rz;
  }
}

}  // namespace nebula
