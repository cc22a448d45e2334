// C++ host test of the facade's emitters (include/nebula_mpm.hpp), no GPU:
// the reference's emitter tests (tests/test_mpm.cpp:175-226) written against
// nebula::AABBEmitter / SphereEmitter / EllipseEmitter the way the reference's
// own tests use its Emitter classes.  Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <random>

#include "nebula_mpm.hpp"

#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                           \
    }                                                                     \
  } while (0)

static int test_aabb() {
  nebula::AABBEmitter e;
  e.center = {1.0f, 2.0f, 3.0f, 0.0f};
  e.size = {2.0f, 4.0f, 6.0f, 0.0f};
  std::mt19937 rng{42};
  for (int i = 0; i < 500; i++) {
    const nebula::vec4 p = e.sample(rng);
    for (int a = 0; a < 3; a++) CHECK(std::fabs(p[a] - e.center[a]) <= e.size[a] * 0.5f + 1e-5f);
  }
  return 0;
}

static int test_sphere() {
  nebula::SphereEmitter e;
  e.center = {0.0f, 0.0f, 0.0f, 0.0f};
  e.radius = 2.5f;
  std::mt19937 rng{42};
  double sum = 0.0;
  const int n = 500;
  for (int i = 0; i < n; i++) {
    const nebula::vec4 p = e.sample(rng);
    const float r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    CHECK(r <= e.radius + 1e-4f);
    sum += r;
  }
  const double mean = sum / n;  // uniform in the ball: 3/4 of the radius
  CHECK(mean > e.radius * 0.5 && mean < e.radius * 0.95);
  return 0;
}

static int test_ellipse() {
  nebula::EllipseEmitter e;
  e.center = {1.0f, -2.0f, 5.0f, 0.0f};
  e.semiA = 3.0f;
  e.semiB = 1.5f;
  std::mt19937 rng{42};
  for (int i = 0; i < 500; i++) {
    const nebula::vec4 p = e.sample(rng);
    const float dx = p[0] - e.center[0], dy = p[1] - e.center[1];
    CHECK((dx * dx) / (e.semiA * e.semiA) + (dy * dy) / (e.semiB * e.semiB) <= 1.0f + 1e-5f);
    CHECK(std::fabs(p[2] - e.center[2]) < 1e-6f);
  }
  return 0;
}

static int test_velocity_along_normal() {
  nebula::SphereEmitter e;
  e.radius = 1.0f;
  e.vel = {0.0f, 3.0f, 0.0f, 0.0f};
  e.emitAlongNormal = true;
  std::mt19937 rng{7};
  const nebula::vec4 v = e.sample_velocity({0.0f, 0.0f, 0.5f, 0.0f}, rng);
  CHECK(std::fabs(v[0]) < 1e-6f && std::fabs(v[1]) < 1e-6f && std::fabs(v[2] - 3.0f) < 1e-5f);
  return 0;
}

int main() {
  if (test_aabb() || test_sphere() || test_ellipse() || test_velocity_along_normal()) return 1;
  std::printf("emitter tests passed\n");
  return 0;
}
