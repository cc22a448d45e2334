// C++ host test through the drop-in facade (include/nebula_mpm.hpp): the
// reference's tests 8-1 (rectangular domain, tests/test_mpm.cpp:749-791) and
// 7-1 (one-step gravity, :519-557) written against nebula::MPMEngine the way
// the reference's own tests use MPMEngine.  Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <vector>

#include "nebula_mpm.hpp"

#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                       \
    }                                                                 \
  } while (0)

static int test_rect_domain() {
  nebula::MPMConfig cfg;
  cfg.nx = cfg.ny = cfg.nz = 0;
  cfg.max_particles = 4;
  cfg.domain_size[0] = 20.0f;
  cfg.domain_size[1] = 5.0f;
  cfg.domain_size[2] = 20.0f;
  cfg.cell_size = 1.25f;
  nebula::MPMEngine engine;
  engine.init(cfg);
  const float Vp = cfg.cell_size * cfg.cell_size * cfg.cell_size;
  engine.appendParticles({{10.0f, 2.5f, 10.0f, Vp}}, {{0.0f, 50.0f, 0.0f, 0.0f}});
  for (int f = 0; f < 30; ++f) engine.step(1.0f / 60.0f);
  engine.sync();
  auto p = engine.getField(NB_FIELD_P)[0];
  CHECK(!std::isnan(p[1]));
  CHECK(p[1] <= cfg.domain_size[1]);
  CHECK(p[1] > cfg.domain_size[1] * 0.3f);
  CHECK(std::fabs(p[0] - 10.0f) < 2.0f);
  CHECK(std::fabs(p[2] - 10.0f) < 2.0f);
  return 0;
}

static int test_gravity_one_step() {
  nebula::MPMConfig cfg;
  cfg.nx = cfg.ny = cfg.nz = 0;
  cfg.max_particles = 1;
  cfg.domain_size[0] = cfg.domain_size[1] = cfg.domain_size[2] = 10.0f;
  cfg.cell_size = 1.25f;
  nebula::MPMEngine engine;
  engine.init(cfg);
  nb_material m{};
  m.mu = 5000.0f;
  m.lambda = 11538.0f;
  m.rho0 = 1000.0f;
  m.q_max = 1e5f;
  engine.setMaterials({m});
  engine.addForce(nebula::GravityForce::FromDirection({0.0f, 1.0f, 0.0f, 0.0f}, -9.8f));
  const float h = cfg.cell_size;
  engine.appendParticles({{5.0f, 5.0f, 5.0f, h * h * h}}, {{0.0f, 0.0f, 0.0f, 0.0f}});
  engine.numSubsteps = 1;
  engine.step(0.01f);
  engine.sync();
  auto v = engine.getField(NB_FIELD_V)[0];
  auto p = engine.getField(NB_FIELD_P)[0];
  CHECK(std::fabs(v[0]) < 1e-4f && std::fabs(v[1] + 0.098f) < 1e-4f && std::fabs(v[2]) < 1e-4f);
  CHECK(std::fabs(p[1] - (5.0f - 0.01f * 0.098f)) < 1e-4f);
  // errors surface as exceptions: 65 colliders exceed the 64 slots
  nebula::ColliderSet cs;
  for (int i = 0; i < 65; i++) cs.addSphere({0, 0, 0, 0}, 1.0f);
  bool threw = false;
  try {
    engine.setColliders(cs);
  } catch (const nebula::Error& e) {
    threw = e.code == NB_ERR_RANGE;
  }
  CHECK(threw);
  return 0;
}

int main() {
  int rc = test_rect_domain();
  if (rc == 0) rc = test_gravity_one_step();
  std::printf(rc == 0 ? "engine_cpp_test: ok\n" : "engine_cpp_test: FAILED\n");
  return rc;
}
