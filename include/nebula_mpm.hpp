// nebula_mpm.hpp — C++ drop-in for the reference's MPMEngine over the C ABI.
//
// Same class name, member names and argument meaning as the reference's
// src/engine/MPMEngine.h:59-186 (and EngineBase's force API,
// src/engine/EngineBase.h:27-31), with HIP in place of Vulkan:
//   init(VkDevice, VmaAllocator, ..., shaderDir, cfg)  ->  init(cfg, hipStream)
//   step(VkCommandBuffer, dt)                          ->  step(dt) (enqueued on the stream)
//   getPositionBuffer() : VkBuffer                     ->  getPositionBuffer() : device float4*
// Errors that the reference reports by throwing std::runtime_error
// (allocation, bad arguments) throw nebula::Error here; capacity overflows the
// reference ignores are reported too.  Header-only; link libnebula_mpm.so.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nebula_mpm.h"
#include "nebula_emitter.hpp"

namespace nebula {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != NB_OK) throw Error(rc, std::string("nebula_mpm: ") + nb_last_error());
}

using vec4 = std::array<float, 4>;
using uvec3 = std::array<uint32_t, 3>;
using MaterialParams = nb_material;
using ColliderPrimitive = nb_collider;
using ForceGPU = nb_force;

enum class MaterialModel : uint32_t {
  ELASTIC = NB_MAT_ELASTIC,
  VON_MISES = NB_MAT_VON_MISES,
  DRUCKER_PRAGER = NB_MAT_DRUCKER_PRAGER,
  GRANULAR_POWDER = NB_MAT_GRANULAR_POWDER,
  FLUID = NB_MAT_FLUID,
  VISCOPLASTIC_MUD = NB_MAT_VISCOPLASTIC_MUD,
};

// src/core/MaterialParams.h:119-180
inline float calcMu(float E, float nu) { return E / (2.0f * (1.0f + nu)); }
inline float calcLambda(float E, float nu) { return E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu)); }
inline MaterialParams presetJelly(float E = 1e4f, float nu = 0.4f, float rho0 = 1000.0f) {
  MaterialParams m{};
  m.mu = calcMu(E, nu);
  m.lambda = calcLambda(E, nu);
  m.rho0 = rho0;
  m.model = NB_MAT_ELASTIC;
  return m;
}
inline MaterialParams presetSand(float E = 1e5f, float nu = 0.3f, float rho0 = 1600.0f) {
  MaterialParams m = presetJelly(E, nu, rho0);
  m.model = NB_MAT_DRUCKER_PRAGER;
  m.M_friction = 0.577f;
  m.q_cohesion = 0.0f;
  return m;
}
inline MaterialParams presetWater(float rho0 = 1000.0f, float bulkK = 5e3f) {
  MaterialParams m{};
  m.rho0 = rho0;
  m.model = NB_MAT_FLUID;
  m.bulkK = bulkK;
  m.fluidGamma = 7.0f;
  m.viscosity = 0.01f;
  return m;
}
inline MaterialParams presetMud(float E = 8e4f, float nu = 0.35f, float rho0 = 1800.0f) {
  MaterialParams m = presetJelly(E, nu, rho0);
  m.model = NB_MAT_VISCOPLASTIC_MUD;
  m.M_friction = 0.4f;
  m.q_cohesion = 5e2f;
  m.viscosity = 0.1f;
  return m;
}
inline MaterialParams presetPowder(float E = 5e4f, float nu = 0.25f, float rho0 = 1200.0f) {
  MaterialParams m = presetJelly(E, nu, rho0);
  m.model = NB_MAT_GRANULAR_POWDER;
  m.M_friction = 0.8f;
  m.q_cohesion = 50.0f;
  return m;
}

// src/core/Collider.h:31-120
class ColliderSet {
 public:
  void addPlane(vec4 point, vec4 normal, float restitution = 0.3f, float friction = 0.0f, vec4 vel = {0, 0, 0, 0}) {
    float l = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    add(NB_COLL_PLANE, point, {normal[0] / l, normal[1] / l, normal[2] / l, 0}, 0.f, restitution, friction, vel);
  }
  void addSphere(vec4 center, float radius, float restitution = 0.3f, float friction = 0.0f, vec4 vel = {0, 0, 0, 0}) {
    add(NB_COLL_SPHERE, center, {0, 0, 0, 0}, radius, restitution, friction, vel);
  }
  void addBox(vec4 center, vec4 halfExt, float restitution = 0.3f, float friction = 0.0f, vec4 vel = {0, 0, 0, 0}) {
    add(NB_COLL_BOX, center, halfExt, 0.f, restitution, friction, vel);
  }
  void addCapsule(vec4 start, vec4 axisVec, float radius, float restitution = 0.3f, float friction = 0.0f,
                  vec4 vel = {0, 0, 0, 0}) {
    add(NB_COLL_CAPSULE, start, axisVec, radius, restitution, friction, vel);
  }
  void clear() { p_.clear(); }
  uint32_t count() const { return uint32_t(p_.size()); }
  bool empty() const { return p_.empty(); }
  const std::vector<ColliderPrimitive>& data() const { return p_; }

 private:
  void add(uint32_t t, vec4 p, vec4 n, float r, float rest, float fric, vec4 v) {
    ColliderPrimitive c{};
    c.type = t;
    c.px = p[0]; c.py = p[1]; c.pz = p[2];
    c.nx = n[0]; c.ny = n[1]; c.nz = n[2];
    c.radius = r;
    c.vx = v[0]; c.vy = v[1]; c.vz = v[2];
    c.restitution = rest;
    c.friction = fric;
    p_.push_back(c);
  }
  std::vector<ColliderPrimitive> p_;
};

// src/core/Force.h:48-150 (the forces that act on the MPM grid)
struct Force {
  virtual ~Force() = default;
  virtual uint32_t type() const = 0;
  virtual void packData(float* out) const { (void)out; }
  virtual void advanceTime(float dt) { (void)dt; }
  uint32_t affectMask = 0u;
  ForceGPU pack() const {
    ForceGPU g{};
    g.type = uint16_t(type());
    g.affect_mask = uint16_t(affectMask);
    packData(g.data);
    return g;
  }
};
struct GravityForce : Force {
  vec4 direction{0.0f, 0.0f, -1.0f, 0.0f};
  float strength = 9.8f;
  uint32_t type() const override { return NB_FORCE_GRAVITY; }
  void packData(float* out) const override {
    out[0] = direction[0]; out[1] = direction[1]; out[2] = direction[2]; out[3] = strength;
  }
  static std::shared_ptr<GravityForce> FromDirection(vec4 dir, float magnitude) {
    auto f = std::make_shared<GravityForce>();
    f->direction = dir;
    f->strength = magnitude;
    return f;
  }
};
struct ConstantWindForce : GravityForce {
  uint32_t type() const override { return NB_FORCE_CONSTANT_WIND; }
};
// src/core/Force.h:180-256: curl-noise turbulence and fBm noise, their seeds
// advanced by speed * dt on every step
struct TurbulenceForce : Force {
  float strength = 1.0f, frequency = 0.5f;
  uint32_t octaves = 3u;
  float seed = 0.0f, speed = 0.3f;
  uint32_t type() const override { return NB_FORCE_TURBULENCE; }
  void packData(float* out) const override {
    out[0] = strength; out[1] = frequency;
    std::memcpy(&out[2], &octaves, 4);
    out[3] = seed;
  }
  void advanceTime(float dt) override { seed += speed * dt; }
};
struct NoiseForce : Force {
  vec4 direction{0.0f, 1.0f, 0.0f, 0.0f};
  float strength = 1.0f, frequency = 1.0f;
  uint32_t octaves = 1u;
  float seed = 0.0f, speed = 0.0f;
  uint32_t type() const override { return NB_FORCE_NOISE; }
  void packData(float* out) const override {
    out[0] = direction[0]; out[1] = direction[1]; out[2] = direction[2];
    out[3] = strength; out[4] = frequency;
    std::memcpy(&out[5], &octaves, 4);
    out[6] = seed;
  }
  void advanceTime(float dt) override { seed += speed * dt; }
};

// src/core/Emitter.h:39-224 (the MPM-side fields): shapes as nb_emitter records
struct Emitter {
  virtual ~Emitter() = default;
  vec4 center{0.f, 0.f, 0.f, 0.f};
  vec4 vel{0.f, 0.f, 0.f, 0.f};
  vec4 center_vel{0.f, 0.f, 0.f, 0.f};
  int particles_per_step = 1024;
  uint32_t particleType = 1u;
  bool emitAlongNormal = false;
  float velocityRandomness = 0.0f;
  int step_count = -1;
  virtual nb_emitter pack() const = 0;
  // Emitter::sample / sample_velocity (src/core/Emitter.h:66-74, 100-208):
  // the draws the engine makes for each emitted particle
  vec4 sample(std::mt19937& rng) const {
    const nb_emitter e = pack();
    float x, y, z;
    emitter_sample(e, rng, x, y, z);
    return {x, y, z, 0.f};
  }
  vec4 sample_velocity(const vec4& p, std::mt19937& rng) const {
    const nb_emitter e = pack();
    float vx, vy, vz;
    emitter_velocity(e, p[0], p[1], p[2], rng, vx, vy, vz);
    return {vx, vy, vz, 0.f};
  }

 protected:
  nb_emitter packCommon(uint32_t shape) const {
    nb_emitter e{};
    e.shape = shape;
    for (int a = 0; a < 3; a++) {
      e.center[a] = center[a];
      e.vel[a] = vel[a];
      e.center_vel[a] = center_vel[a];
    }
    e.particles_per_step = particles_per_step;
    e.particle_type = particleType;
    e.emit_along_normal = emitAlongNormal ? 1u : 0u;
    e.velocity_randomness = velocityRandomness;
    e.step_count = step_count;
    return e;
  }
};
struct AABBEmitter : Emitter {
  vec4 size{1.f, 1.f, 1.f, 0.f};  // edge lengths
  nb_emitter pack() const override {
    nb_emitter e = packCommon(NB_EMITTER_AABB);
    for (int a = 0; a < 3; a++) e.size[a] = size[a];
    return e;
  }
  static std::shared_ptr<AABBEmitter> FromAABB(vec4 mn, vec4 mx, vec4 vel, int particles_per_step) {
    auto e = std::make_shared<AABBEmitter>();
    for (int a = 0; a < 3; a++) {
      e->center[a] = (mn[a] + mx[a]) * 0.5f;
      e->size[a] = mx[a] - mn[a];
    }
    e->vel = vel;
    e->particles_per_step = particles_per_step;
    return e;
  }
};
struct SphereEmitter : Emitter {
  float radius = 1.0f;
  nb_emitter pack() const override {
    nb_emitter e = packCommon(NB_EMITTER_SPHERE);
    e.size[0] = e.size[1] = e.size[2] = radius;
    return e;
  }
  static std::shared_ptr<SphereEmitter> FromSphere(vec4 center, float radius, vec4 vel, int particles_per_step) {
    auto e = std::make_shared<SphereEmitter>();
    e->center = center;
    e->radius = radius;
    e->vel = vel;
    e->particles_per_step = particles_per_step;
    return e;
  }
};
struct EllipseEmitter : Emitter {
  float semiA = 1.0f, semiB = 1.0f;  // x and y semi-axes (z fixed at the center)
  nb_emitter pack() const override {
    nb_emitter e = packCommon(NB_EMITTER_ELLIPSE);
    e.size[0] = semiA;
    e.size[1] = semiB;
    return e;
  }
};

// MPMConfig (src/engine/MPMEngine.h:22-55)
struct MPMConfig : nb_mpm_config {
  MPMConfig() { nb_mpm_config_default(this); }
  uint32_t particleCount() const { return nx * ny * nz; }
  uint32_t maxParticleCount() const { return max_particles > 0 ? max_particles : particleCount(); }
};

class MPMEngine {
 public:
  MPMEngine() = default;
  MPMEngine(const MPMEngine&) = delete;
  MPMEngine& operator=(const MPMEngine&) = delete;
  ~MPMEngine() { cleanup(); }

  void init(const MPMConfig& cfg = {}, void* hipStream = nullptr) {
    cleanup();
    check(nb_mpm_create(&cfg, hipStream, &h_));
    cfg_ = cfg;
    nb_mpm_knobs_default(&k_);
  }
  void cleanup() {
    if (h_) nb_mpm_destroy(h_);
    h_ = nullptr;
  }

  void step(float dt) {
    pushKnobs();
    if (!forces_.empty()) uploadForces(dt);
    check(nb_mpm_step(h_, dt));
    syncEmitters();
  }
  // step(VkCommandBuffer, dt) records into the caller's buffer; this enqueues
  // on the caller's hipStream_t (passed as void*), and later calls stay there
  void step(void* hip_stream, float dt) {
    setStream(hip_stream);
    step(dt);
  }
  void setStream(void* hip_stream) { check(nb_mpm_set_stream(h_, hip_stream)); }

  // emitters (src/engine/MPMEngine.cpp:271-352); their centers follow the engine's
  void addEmitter(std::shared_ptr<Emitter> e) {
    const nb_emitter rec = e->pack();
    check(nb_mpm_add_emitter(h_, &rec));
    emitters_.push_back(std::move(e));
  }
  void clearEmitters() {
    check(nb_mpm_clear_emitters(h_));
    emitters_.clear();
  }
  void emitFromEmitters(float dt) {
    check(nb_mpm_emit(h_, dt));
    syncEmitters();
  }
  void sync() { check(nb_mpm_sync(h_)); }

  const MPMConfig& config() const { return cfg_; }
  uint32_t liveParticleCount() const {
    uint32_t n = 0;
    check(nb_mpm_live_particle_count(h_, &n));
    return n;
  }
  uvec3 gridRes() const {
    uvec3 r{};
    check(nb_mpm_grid_res(h_, r.data()));
    return r;
  }
  uint32_t totalCells() const {
    uint32_t n = 0;
    check(nb_mpm_total_cells(h_, &n));
    return n;
  }
  void* getPositionBuffer() const { return field(NB_FIELD_P); }
  void* getVelocityBuffer() const { return field(NB_FIELD_V); }

  // public knobs (src/engine/MPMEngine.h:72-80); pushed on every step()
  float restitution = 0.3f;
  float wall_friction = 0.0f;
  int numSubsteps = 20;
  uint32_t plasticModel = 0;
  float q_max = 1e5f;
  float M_friction = 0.577f;
  float q_cohesion = 0.0f;
  float flip_ratio = 0.0f;

  void setColliderSphere(float radius, float cx, float cy, float cz) {
    check(nb_mpm_set_collider_sphere(h_, radius, cx, cy, cz));
  }
  void clearCollider() { check(nb_mpm_clear_collider(h_)); }
  void setMaterials(const std::vector<MaterialParams>& mats) {
    check(nb_mpm_set_materials(h_, mats.data(), uint32_t(mats.size())));
  }
  void setParticleMaterialIds(const std::vector<uint32_t>& ids) {
    check(nb_mpm_set_particle_material_ids(h_, ids.data(), uint32_t(ids.size())));
  }
  void setColliders(const ColliderSet& cols) {
    check(nb_mpm_set_colliders(h_, cols.data().data(), cols.count()));
  }
  void clearAnalyticColliders() { check(nb_mpm_clear_analytic_colliders(h_)); }
  void resetParticles() { check(nb_mpm_reset_particles(h_)); }
  void appendParticles(const std::vector<vec4>& pos, const std::vector<vec4>& vel) {
    if (pos.size() != vel.size()) throw Error(NB_ERR_ARG, "appendParticles: pos/vel size mismatch");
    check(nb_mpm_append_particles(h_, pos.empty() ? nullptr : pos[0].data(), vel.empty() ? nullptr : vel[0].data(),
                                  uint32_t(pos.size()), nullptr));
  }
  void setColliderSDF(const std::vector<float>& mortonSDF) {
    check(nb_mpm_set_collider_sdf(h_, mortonSDF.data(), mortonSDF.size()));
  }
  // buildMeshSDF (src/core/MeshSDF.h:85-126) + setColliderSDF in one call, on
  // the GPU: tris holds 12 floats per triangle (normal, v0, v1, v2).
  void setColliderMesh(const std::vector<float>& tris, bool exact = false) {
    check(nb_mpm_set_collider_mesh(h_, tris.data(), uint32_t(tris.size() / 12), exact ? 1 : 0));
  }
  std::vector<float> getColliderSDF() {
    uint32_t cells = 0;
    check(nb_mpm_total_cells(h_, &cells));
    std::vector<float> out(cells);
    check(nb_mpm_get_collider_sdf(h_, out.data()));
    return out;
  }

  // EngineBase force list (src/engine/EngineBase.cpp:22-40)
  void addForce(std::shared_ptr<Force> f) {
    forces_.push_back(std::move(f));
    uploadForces(0.f);
  }
  void removeForce(const std::shared_ptr<Force>& f) {
    std::vector<std::shared_ptr<Force>> keep;
    for (auto& x : forces_)
      if (x != f) keep.push_back(x);
    forces_ = keep;
    uploadForces(0.f);
  }
  void setForces(std::vector<std::shared_ptr<Force>> forces) {
    forces_ = std::move(forces);
    uploadForces(0.f);
  }
  void clearForces() {
    forces_.clear();
    uploadForces(0.f);
  }
  const std::vector<std::shared_ptr<Force>>& forces() const { return forces_; }

  // state access in the reference's float4 layout
  std::vector<vec4> getField(int field, uint32_t first = 0, uint32_t count = ~0u) const {
    if (count == ~0u) count = liveParticleCount() - first;
    std::vector<vec4> out(count);
    if (count) check(nb_mpm_get_field(h_, field, out[0].data(), first, count));
    return out;
  }
  nb_mpm* handle() const { return h_; }

 private:
  void* field(int f) const {
    void* p = nullptr;
    check(nb_mpm_field_device_ptr(h_, f, &p));
    return p;
  }
  void pushKnobs() {
    k_.restitution = restitution;
    k_.wall_friction = wall_friction;
    k_.num_substeps = numSubsteps;
    k_.plastic_model = plasticModel;
    k_.q_max = q_max;
    k_.M_friction = M_friction;
    k_.q_cohesion = q_cohesion;
    k_.flip_ratio = flip_ratio;
    check(nb_mpm_set_knobs(h_, &k_));
  }
  void uploadForces(float dt) {
    std::vector<ForceGPU> packed;
    for (auto& f : forces_) {
      f->advanceTime(dt);
      packed.push_back(f->pack());
    }
    check(nb_mpm_set_forces(h_, packed.empty() ? nullptr : packed.data(), uint32_t(packed.size())));
  }

  nb_mpm* h_ = nullptr;
  MPMConfig cfg_;
  nb_mpm_knobs k_{};
  std::vector<std::shared_ptr<Force>> forces_;
  std::vector<std::shared_ptr<Emitter>> emitters_;
  void syncEmitters() {
    for (uint32_t i = 0; i < emitters_.size(); i++) {
      nb_emitter rec{};
      check(nb_mpm_get_emitter(h_, i, &rec, nullptr));
      for (int a = 0; a < 3; a++) emitters_[i]->center[a] = rec.center[a];
    }
  }
};

}  // namespace nebula
