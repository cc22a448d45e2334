// oracle_capi.cpp — C entry points of the CPU oracle, loaded by the Python
// test harness through ctypes.  TEST INFRASTRUCTURE ONLY: the product library
// never links this file, and only tests/, smoke() and the cpu_baseline leg of
// bench.py load it.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "mpm_oracle.hpp"

using namespace oracle;

namespace {
template <class R> static M3<R> load_m3(const R* a) {
  M3<R> m;
  for (int c = 0; c < 3; c++) m.c[c] = {a[3 * c + 0], a[3 * c + 1], a[3 * c + 2]};
  return m;
}
template <class R> static void store_m3(const M3<R>& m, R* a) {
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) a[3 * c + r] = m.at(c, r);
}
}  // namespace

extern "C" {

// Scene description handed in by the harness (a flat POD copy of Params).
struct orc_scene {
  float dt;
  float cell_size;
  uint32_t grid_res[3];
  float world_min[3];
  float world_max[3];
  float flip_ratio;
  uint32_t cube_res;
  uint32_t material_count;
  uint32_t collider_count;
  uint32_t force_count;
  uint32_t has_sdf;
  uint32_t corrected_svd;
};

struct orc_sim {
  int precision;  // 32 or 64
  orc_scene sc;
  std::vector<Material> mats;
  std::vector<Collider> cols;
  std::vector<Force> forces;
  std::vector<float> sdf;
  State<float> s32;
  State<double> s64;

  Params params() const {
    Params p{};
    p.dt = sc.dt;
    p.cell_size = sc.cell_size;
    for (int a = 0; a < 3; a++) {
      p.grid_res[a] = sc.grid_res[a];
      p.world_min[a] = sc.world_min[a];
      p.world_max[a] = sc.world_max[a];
    }
    p.flip_ratio = sc.flip_ratio;
    p.cube_res = sc.cube_res;
    p.materials = mats.data();
    p.material_count = uint32_t(mats.size());
    p.colliders = cols.data();
    p.collider_count = uint32_t(cols.size());
    p.forces = forces.data();
    p.force_count = uint32_t(forces.size());
    p.sdf = sc.has_sdf ? sdf.data() : nullptr;
    p.corrected_svd = sc.corrected_svd != 0;
    return p;
  }
};

orc_sim* orc_sim_create(const orc_scene* sc, const void* mats, const void* cols, const void* forces, const float* sdf,
                        int precision) {
  orc_sim* s = new orc_sim();
  s->precision = precision;
  s->sc = *sc;
  s->mats.resize(sc->material_count);
  if (sc->material_count) std::memcpy(s->mats.data(), mats, sizeof(Material) * sc->material_count);
  s->cols.resize(sc->collider_count);
  if (sc->collider_count) std::memcpy(s->cols.data(), cols, sizeof(Collider) * sc->collider_count);
  s->forces.resize(sc->force_count);
  if (sc->force_count) std::memcpy(s->forces.data(), forces, sizeof(Force) * sc->force_count);
  uint32_t NC = sc->cube_res * sc->cube_res * sc->cube_res;
  if (sc->has_sdf) s->sdf.assign(sdf, sdf + NC);
  return s;
}

void orc_sim_destroy(orc_sim* s) { delete s; }

void orc_sim_set_dt(orc_sim* s, float dt) { s->sc.dt = dt; }
void orc_sim_set_flip_ratio(orc_sim* s, float r) { s->sc.flip_ratio = r; }
void orc_sim_set_corrected_svd(orc_sim* s, int on) { s->sc.corrected_svd = on ? 1u : 0u; }
void orc_sim_set_colliders(orc_sim* s, const void* cols, uint32_t n) {
  s->cols.resize(n);
  if (n) std::memcpy(s->cols.data(), cols, sizeof(Collider) * n);
  s->sc.collider_count = n;
}

// fields: 8 arrays of n float4 concatenated (P, v, F0, F1, F2, B0, B1, B2)
void orc_sim_set_state(orc_sim* s, const float* fields, uint32_t n) {
  uint32_t NC = s->sc.cube_res * s->sc.cube_res * s->sc.cube_res;
  if (s->precision == 64) {
    s->s64.resize(n, NC);
    for (int f = 0; f < 8; f++)
      for (size_t i = 0; i < size_t(n) * 4; i++) s->s64.f[f][i] = double(fields[size_t(f) * n * 4 + i]);
    for (uint32_t i = 0; i < n; i++) std::memcpy(&s->s64.matid[i], &fields[size_t(1) * n * 4 + 4 * i + 3], 4);
  } else {
    s->s32.resize(n, NC);
    for (int f = 0; f < 8; f++) std::memcpy(s->s32.f[f].data(), fields + size_t(f) * n * 4, sizeof(float) * n * 4);
    for (uint32_t i = 0; i < n; i++) std::memcpy(&s->s32.matid[i], &fields[size_t(1) * n * 4 + 4 * i + 3], 4);
  }
}

// FP64 state (precision 64 only), e.g. for finite differences; material ids
// come separately since they are bit patterns in the float layout
void orc_sim_set_state_f64(orc_sim* s, const double* fields, const uint32_t* matid, uint32_t n) {
  uint32_t NC = s->sc.cube_res * s->sc.cube_res * s->sc.cube_res;
  s->s64.resize(n, NC);
  for (int f = 0; f < 8; f++)
    for (size_t i = 0; i < size_t(n) * 4; i++) s->s64.f[f][i] = fields[size_t(f) * n * 4 + i];
  for (uint32_t i = 0; i < n; i++) s->s64.matid[i] = matid[i];
}

void orc_sim_get_state(orc_sim* s, float* fields) {
  uint32_t n = s->precision == 64 ? s->s64.n : s->s32.n;
  for (int f = 0; f < 8; f++)
    for (size_t i = 0; i < size_t(n) * 4; i++)
      fields[size_t(f) * n * 4 + i] = s->precision == 64 ? float(s->s64.f[f][i]) : s->s32.f[f][i];
  for (uint32_t i = 0; i < n; i++) {
    uint32_t id = s->precision == 64 ? s->s64.matid[i] : s->s32.matid[i];
    std::memcpy(&fields[size_t(1) * n * 4 + 4 * i + 3], &id, 4);
  }
}

// double-precision state export (FP64 twin only); matid lane holds the id as a number
void orc_sim_get_state_f64(orc_sim* s, double* fields) {
  uint32_t n = s->s64.n;
  for (int f = 0; f < 8; f++)
    for (size_t i = 0; i < size_t(n) * 4; i++) fields[size_t(f) * n * 4 + i] = s->s64.f[f][i];
  for (uint32_t i = 0; i < n; i++) fields[size_t(1) * n * 4 + 4 * i + 3] = double(s->s64.matid[i]);
}

void orc_sim_bin(orc_sim* s) {
  Params p = s->params();
  if (s->precision == 64)
    bin<double>(p, s->s64.f[0].data(), s->s64.n, s->s64.keys.data(), s->s64.sorted.data(), s->s64.cell_count.data(),
                s->s64.cell_end.data());
  else
    bin<float>(p, s->s32.f[0].data(), s->s32.n, s->s32.keys.data(), s->s32.sorted.data(), s->s32.cell_count.data(),
               s->s32.cell_end.data());
}

void orc_sim_get_bins(orc_sim* s, uint32_t* keys, uint32_t* sorted, uint32_t* cell_count, uint32_t* cell_end) {
  auto cp = [](const std::vector<uint32_t>& v, uint32_t* out) {
    if (out) std::memcpy(out, v.data(), v.size() * 4);
  };
  if (s->precision == 64) {
    cp(s->s64.keys, keys); cp(s->s64.sorted, sorted); cp(s->s64.cell_count, cell_count); cp(s->s64.cell_end, cell_end);
  } else {
    cp(s->s32.keys, keys); cp(s->s32.sorted, sorted); cp(s->s32.cell_count, cell_count); cp(s->s32.cell_end, cell_end);
  }
}

void orc_sim_p2g(orc_sim* s) {
  Params p = s->params();
  if (s->precision == 64) p2g<double>(p, s->s64); else p2g<float>(p, s->s32);
}
void orc_sim_grid_update(orc_sim* s) {
  Params p = s->params();
  if (s->precision == 64) grid_update<double>(p, s->s64); else grid_update<float>(p, s->s32);
}
void orc_sim_sdf_bc(orc_sim* s) {
  Params p = s->params();
  if (s->precision == 64) sdf_bc<double>(p, s->s64); else sdf_bc<float>(p, s->s32);
}
void orc_sim_g2p(orc_sim* s) {
  Params p = s->params();
  if (s->precision == 64) g2p<double>(p, s->s64); else g2p<float>(p, s->s32);
}
// n substeps; returns the wall time in seconds
double orc_sim_substeps(orc_sim* s, uint32_t n) {
  Params p = s->params();
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t i = 0; i < n; i++) {
    if (s->precision == 64) substep<double>(p, s->s64); else substep<float>(p, s->s32);
  }
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// grid: mom4 = 2*NC float4, mass = NC
void orc_sim_get_grid(orc_sim* s, float* mom4, float* mass) {
  if (s->precision == 64) {
    for (size_t i = 0; i < s->s64.grid_mom.size(); i++) mom4[i] = float(s->s64.grid_mom[i]);
    for (size_t i = 0; i < s->s64.grid_mass.size(); i++) mass[i] = float(s->s64.grid_mass[i]);
  } else {
    std::memcpy(mom4, s->s32.grid_mom.data(), s->s32.grid_mom.size() * 4);
    std::memcpy(mass, s->s32.grid_mass.data(), s->s32.grid_mass.size() * 4);
  }
}
void orc_sim_get_grid_f64(orc_sim* s, double* mom4, double* mass) {
  std::memcpy(mom4, s->s64.grid_mom.data(), s->s64.grid_mom.size() * 8);
  std::memcpy(mass, s->s64.grid_mass.data(), s->s64.grid_mass.size() * 8);
}
// mom4 = NC float4 (first half only), mass = NC; the second half is zeroed
void orc_sim_set_grid(orc_sim* s, const float* mom4, const float* mass) {
  uint32_t NC = s->sc.cube_res * s->sc.cube_res * s->sc.cube_res;
  if (s->precision == 64) {
    std::fill(s->s64.grid_mom.begin(), s->s64.grid_mom.end(), 0.0);
    for (size_t i = 0; i < size_t(NC) * 4; i++) s->s64.grid_mom[i] = mom4[i];
    for (size_t i = 0; i < NC; i++) s->s64.grid_mass[i] = mass[i];
  } else {
    std::fill(s->s32.grid_mom.begin(), s->s32.grid_mom.end(), 0.0f);
    std::memcpy(s->s32.grid_mom.data(), mom4, size_t(NC) * 16);
    std::memcpy(s->s32.grid_mass.data(), mass, size_t(NC) * 4);
  }
}
// The saved pre-force velocities (the grid's [NC, 2NC) half, what FLIP's
// vpic_old gathers): old4 = NC float4.
void orc_sim_set_grid_old(orc_sim* s, const float* old4) {
  uint32_t NC = s->sc.cube_res * s->sc.cube_res * s->sc.cube_res;
  if (s->precision == 64) {
    for (size_t i = 0; i < size_t(NC) * 4; i++) s->s64.grid_mom[size_t(NC) * 4 + i] = old4[i];
  } else {
    std::memcpy(s->s32.grid_mom.data() + size_t(NC) * 4, old4, size_t(NC) * 16);
  }
}
// Set the bins directly (tests inject host-built bins, as MPMHarness::buildHashGridCPU does).
void orc_sim_set_bins(orc_sim* s, const uint32_t* sorted, const uint32_t* cell_count, const uint32_t* cell_end) {
  uint32_t NC = s->sc.cube_res * s->sc.cube_res * s->sc.cube_res;
  if (s->precision == 64) {
    std::memcpy(s->s64.sorted.data(), sorted, s->s64.n * 4);
    std::memcpy(s->s64.cell_count.data(), cell_count, NC * 4);
    std::memcpy(s->s64.cell_end.data(), cell_end, NC * 4);
  } else {
    std::memcpy(s->s32.sorted.data(), sorted, s->s32.n * 4);
    std::memcpy(s->s32.cell_count.data(), cell_count, NC * 4);
    std::memcpy(s->s32.cell_end.data(), cell_end, NC * 4);
  }
}

// ---- scalar kernels, for property tests ----
uint32_t orc_morton_encode(int x, int y, int z) { return morton_encode(x, y, z); }
void orc_morton_decode(uint32_t c, int* x, int* y, int* z) { morton_decode(c, *x, *y, *z); }
uint32_t orc_cell_id(const orc_scene* sc, float x, float y, float z) {
  Params p{};
  p.cell_size = sc->cell_size;
  for (int a = 0; a < 3; a++) {
    p.grid_res[a] = sc->grid_res[a];
    p.world_min[a] = sc->world_min[a];
  }
  return cell_id_from_pos<float>(p, x, y, z);
}
float orc_bspline2_f32(float d) { return bspline2<float>(d); }
float orc_bspline2g_f32(float d) { return bspline2g<float>(d); }
double orc_bspline2_f64(double d) { return bspline2<double>(d); }
double orc_bspline2g_f64(double d) { return bspline2g<double>(d); }

// matrices are column-major 9-arrays
void orc_svd3_f32(const float* F, float* U, float* sigma, float* V) {
  M3<float> u, v;
  V3<float> s;
  svd3(load_m3(F), u, s, v);
  store_m3(u, U);
  store_m3(v, V);
  sigma[0] = s.x; sigma[1] = s.y; sigma[2] = s.z;
}
void orc_svd3_f64(const double* F, double* U, double* sigma, double* V) {
  M3<double> u, v;
  V3<double> s;
  svd3(load_m3(F), u, s, v);
  store_m3(u, U);
  store_m3(v, V);
  sigma[0] = s.x; sigma[1] = s.y; sigma[2] = s.z;
}
void orc_hencky_f32(const float* F, float mu, float lam, float* tau) { store_m3(hencky_stress(load_m3(F), mu, lam), tau); }
void orc_hencky_f64(const double* F, double mu, double lam, double* tau) { store_m3(hencky_stress(load_m3(F), mu, lam), tau); }
void orc_project_vm_f32(const float* F, float mu, float lam, float qmax, float* out) {
  store_m3(project_von_mises(load_m3(F), mu, lam, qmax), out);
}
void orc_project_vm_f64(const double* F, double mu, double lam, double qmax, double* out) {
  store_m3(project_von_mises(load_m3(F), mu, lam, qmax), out);
}
void orc_project_dp_f32(const float* F, float mu, float lam, float M, float qc, float* out) {
  store_m3(project_drucker_prager(load_m3(F), mu, lam, M, qc), out);
}
void orc_project_dp_f64(const double* F, double mu, double lam, double M, double qc, double* out) {
  store_m3(project_drucker_prager(load_m3(F), mu, lam, M, qc), out);
}


// One analytic collider (a 64-byte ColliderPrimitive record) at point x, as
// the grid update evaluates it (shaders/collider_common.glsl:22-68,
// mpm_grid_update.comp:84-93): returns the SDF, writes the outward normal.
}  // extern "C"
template <class R> static R collider_sdf_at(const Collider& c, const V3<R>& x, V3<R>& on) {
  V3<R> cpos(c.px, c.py, c.pz), cn(c.nx, c.ny, c.nz);
  if (c.type == 1u) return sphere_sdf(x, cpos, R(c.radius), on);
  if (c.type == 2u) return box_sdf(x, cpos, cn, on);
  if (c.type == 3u) return capsule_sdf(x, cpos, cn, R(c.radius), on);
  return plane_sdf(x, cpos, cn, on);
}
extern "C" {
double orc_collider_sdf(const void* rec, double x, double y, double z, int precision, double n_out[3]) {
  Collider c;
  std::memcpy(&c, rec, sizeof(Collider));
  if (precision == 64) {
    V3<double> on;
    const double d = collider_sdf_at<double>(c, V3<double>(x, y, z), on);
    n_out[0] = on.x; n_out[1] = on.y; n_out[2] = on.z;
    return d;
  }
  V3<float> on;
  const float d = collider_sdf_at<float>(c, V3<float>(float(x), float(y), float(z)), on);
  n_out[0] = on.x; n_out[1] = on.y; n_out[2] = on.z;
  return double(d);
}
// applyColliderBC (shaders/collider_common.glsl:73-85) on one velocity v (in/out).
void orc_collider_bc(double v[3], const double n[3], const double vc[3], double rest, double fric, int precision) {
  if (precision == 64) {
    V3<double> vv(v[0], v[1], v[2]);
    apply_collider_bc(vv, V3<double>(n[0], n[1], n[2]), V3<double>(vc[0], vc[1], vc[2]), rest, fric);
    v[0] = vv.x; v[1] = vv.y; v[2] = vv.z;
  } else {
    V3<float> vv(static_cast<float>(v[0]), static_cast<float>(v[1]), static_cast<float>(v[2]));
    apply_collider_bc(vv, V3<float>(static_cast<float>(n[0]), static_cast<float>(n[1]), static_cast<float>(n[2])),
                      V3<float>(static_cast<float>(vc[0]), static_cast<float>(vc[1]), static_cast<float>(vc[2])),
                      static_cast<float>(rest), static_cast<float>(fric));
    v[0] = vv.x; v[1] = vv.y; v[2] = vv.z;
  }
}

// The spatial hash of the PBF and cloth engines (zero_cells, hash_count,
// hash_scan_*, hash_add_base, hash_sort; dispatched at
// src/engine/FluidEngine.cpp:444-474): the stable counting sort of the
// particles whose typeFlag is not 6 (shaders/hash_count.comp:15,
// hash_sort.comp:16) by cellId (shaders/common.glsl:100-105), END offsets,
// ascending index within a cell.  type_flag may be null.  Returns the number
// of particles sorted.
uint32_t orc_hash_build(const float* pred4, const uint32_t* type_flag, uint32_t n, const float* world_min,
                        float cell_size, const uint32_t* grid_res, uint32_t hash_cells, uint32_t* cell_count,
                        uint32_t* cell_end, uint32_t* sorted) {
  Params p{};
  p.cell_size = cell_size;
  for (int a = 0; a < 3; a++) {
    p.grid_res[a] = grid_res[a];
    p.world_min[a] = world_min[a];
  }
  std::vector<uint32_t> key(n);
  std::fill(cell_count, cell_count + hash_cells, 0u);
  for (uint32_t i = 0; i < n; i++) {
    if (type_flag && type_flag[i] == 6u) continue;
    key[i] = cell_id_from_pos<float>(p, pred4[4 * i + 0], pred4[4 * i + 1], pred4[4 * i + 2]);
    cell_count[key[i]]++;
  }
  uint32_t acc = 0;
  for (uint32_t c = 0; c < hash_cells; c++) {
    acc += cell_count[c];
    cell_end[c] = acc;
  }
  std::vector<uint32_t> wp(hash_cells);
  for (uint32_t c = 0; c < hash_cells; c++) wp[c] = cell_end[c] - cell_count[c];
  for (uint32_t i = 0; i < n; i++) {
    if (type_flag && type_flag[i] == 6u) continue;
    sorted[wp[key[i]]++] = i;
  }
  return acc;
}

// procedural noise of the turbulence / noise forces (shaders/force_noise.glsl)
double orc_noise_value(double x, double y, double z, int precision) {
  if (precision == 64) return oracle::noise_value3(oracle::V3<double>(x, y, z));
  return double(oracle::noise_value3(oracle::V3<float>(float(x), float(y), float(z))));
}
double orc_noise_fbm(double x, double y, double z, uint32_t octaves, int precision) {
  if (precision == 64) return oracle::noise_fbm3(oracle::V3<double>(x, y, z), octaves);
  return double(oracle::noise_fbm3(oracle::V3<float>(float(x), float(y), float(z)), octaves));
}
void orc_noise_curl(double x, double y, double z, uint32_t octaves, int precision, double out[3]) {
  if (precision == 64) {
    const oracle::V3<double> c = oracle::noise_curl3(oracle::V3<double>(x, y, z), octaves);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
  } else {
    const oracle::V3<float> c = oracle::noise_curl3(oracle::V3<float>(float(x), float(y), float(z)), octaves);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
  }
}
void orc_noise_hash(double x, double y, double z, int precision, double out[3]) {
  if (precision == 64) {
    const oracle::V3<double> c = oracle::noise_hash33(oracle::V3<double>(x, y, z));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
  } else {
    const oracle::V3<float> c = oracle::noise_hash33(oracle::V3<float>(float(x), float(y), float(z)));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
  }
}


// MPMEngine::emitFromEmitters (src/engine/MPMEngine.cpp:289-352) over n_calls
// calls with dt: the particles it would append, in order (positions (x, Vp),
// velocities (v, bits(type))).  Records are the library's 80-byte emitter
// records; rng: std::mt19937 seeded 12345 (MPMEngine.h:165); the components
// of one glm::vec3 constructed from three draws are drawn x, y, z.
struct OrcEmitter {
  uint32_t shape;
  float center[3], size[3], vel[3], center_vel[3];
  int32_t particles_per_step;
  uint32_t particle_type, emit_along_normal;
  float velocity_randomness;
  int32_t step_count;
  uint32_t reserved[2];
};
static_assert(sizeof(OrcEmitter) == 80, "emitter record");

uint32_t orc_emit(const void* records, uint32_t n_emitters, uint32_t n_calls, float dt, uint32_t capacity,
                  uint32_t live, float Vp, float* pos4, float* vel4) {
  std::vector<OrcEmitter> em(n_emitters);
  std::memcpy(em.data(), records, sizeof(OrcEmitter) * n_emitters);
  std::vector<int> done(n_emitters, 0);
  std::mt19937 rng(12345u);
  uint32_t out = 0;
  for (uint32_t call = 0; call < n_calls; call++) {
    for (uint32_t si = 0; si < n_emitters; si++) {
      OrcEmitter& e = em[si];
      bool emit;
      if (e.step_count == -1) emit = done[si] == 0;
      else if (e.step_count == 0) emit = true;
      else emit = done[si] < e.step_count;
      if (!emit) continue;
      const int n = std::min(e.particles_per_step, int(capacity) - int(live));
      if (n <= 0) {
        done[si]++;
        continue;
      }
      float idf;
      std::memcpy(&idf, &e.particle_type, 4);
      std::vector<float> px(n), py(n), pz(n);
      for (int j = 0; j < n; j++) {  // Emitter::sample, Emitter.h:100-107, 147-155, 198-208
        if (e.shape == 0u) {
          std::uniform_real_distribution<float> ux(-e.size[0] * 0.5f, e.size[0] * 0.5f);
          std::uniform_real_distribution<float> uy(-e.size[1] * 0.5f, e.size[1] * 0.5f);
          std::uniform_real_distribution<float> uz(-e.size[2] * 0.5f, e.size[2] * 0.5f);
          const float a = ux(rng);
          const float b = uy(rng);
          const float c = uz(rng);
          px[j] = e.center[0] + a; py[j] = e.center[1] + b; pz[j] = e.center[2] + c;
        } else if (e.shape == 1u) {
          std::uniform_real_distribution<float> u01(0.0f, 1.0f), uth(0.0f, 3.14159265f), uph(0.0f, 6.28318530f);
          const float r = e.size[0] * std::cbrt(u01(rng));
          const float theta = uth(rng);
          const float phi = uph(rng);
          px[j] = e.center[0] + r * std::sin(theta) * std::cos(phi);
          py[j] = e.center[1] + r * std::sin(theta) * std::sin(phi);
          pz[j] = e.center[2] + r * std::cos(theta);
        } else {
          const float A = e.size[0], B = e.size[1];
          std::uniform_real_distribution<float> ux(-A, A), uy(-B, B);
          float a, b;
          do {
            a = ux(rng);
            b = uy(rng);
          } while ((a * a) / (A * A) + (b * b) / (B * B) > 1.0f);
          px[j] = e.center[0] + a; py[j] = e.center[1] + b; pz[j] = e.center[2];
        }
      }
      for (int j = 0; j < n; j++) {  // Emitter::sample_velocity, Emitter.h:66-74
        float v[3] = {e.vel[0], e.vel[1], e.vel[2]};
        if (e.emit_along_normal) {
          const float d[3] = {px[j] - e.center[0], py[j] - e.center[1], pz[j] - e.center[2]};
          float nrm[3] = {0.0f, 0.0f, 0.0f};
          if (e.shape == 0u) {  // AABBEmitter::surface_normal, Emitter.h:109-117
            float q[3], aq[3];
            for (int k = 0; k < 3; k++) {
              q[k] = d[k] / std::max(e.size[k] * 0.5f, 1e-6f);
              aq[k] = std::fabs(q[k]);
            }
            if (aq[0] >= aq[1] && aq[0] >= aq[2]) nrm[0] = q[0] >= 0.0f ? 1.0f : -1.0f;
            else if (aq[1] >= aq[2]) nrm[1] = q[1] >= 0.0f ? 1.0f : -1.0f;
            else nrm[2] = q[2] >= 0.0f ? 1.0f : -1.0f;
          } else {  // Emitter::surface_normal, Emitter.h:59-63
            const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (len > 1e-6f) { nrm[0] = d[0] / len; nrm[1] = d[1] / len; nrm[2] = d[2] / len; }
            else nrm[2] = 1.0f;
          }
          const float sp = std::sqrt(e.vel[0] * e.vel[0] + e.vel[1] * e.vel[1] + e.vel[2] * e.vel[2]);
          for (int k = 0; k < 3; k++) v[k] = nrm[k] * sp;
        }
        if (e.velocity_randomness > 0.0f) {
          std::uniform_real_distribution<float> ur(-e.velocity_randomness, e.velocity_randomness);
          const float a = ur(rng);
          const float b = ur(rng);
          const float c = ur(rng);
          v[0] += a; v[1] += b; v[2] += c;
        }
        pos4[4 * out + 0] = px[j]; pos4[4 * out + 1] = py[j]; pos4[4 * out + 2] = pz[j]; pos4[4 * out + 3] = Vp;
        vel4[4 * out + 0] = v[0]; vel4[4 * out + 1] = v[1]; vel4[4 * out + 2] = v[2]; vel4[4 * out + 3] = idf;
        out++;
      }
      live += uint32_t(n);
      done[si]++;
      for (int k = 0; k < 3; k++) e.center[k] += e.center_vel[k] * dt;
    }
  }
  return out;
}


// buildMeshSDF (src/core/MeshSDF.h:85-126) with closestPtOnTriangle
// (MeshSDF.h:42-75): cells (i + 0.5) cs of the res grid, Morton order, cube
// cells beyond the grid 1e9; sign from the normal of the first closest
// triangle.  tris: n x 12 floats (normal, v0, v1, v2).  exact = 0 keeps
// the reference's missing region for edge b-c (a cell there gets the plane
// projection); exact = 1 adds it (Ericson, Real-Time Collision Detection,
// 5.1.5).
static void orc_closest(const float* p, const float* a, const float* b, const float* c, bool exact, float* q) {
  auto sub = [](const float* x, const float* y, float* o) { o[0] = x[0] - y[0]; o[1] = x[1] - y[1]; o[2] = x[2] - y[2]; };
  auto dot = [](const float* x, const float* y) { return (x[0] * y[0] + x[1] * y[1]) + x[2] * y[2]; };
  float ab[3], ac[3], ap[3];
  sub(b, a, ab); sub(c, a, ac); sub(p, a, ap);
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) { q[0] = a[0]; q[1] = a[1]; q[2] = a[2]; return; }
  float bp[3];
  sub(p, b, bp);
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) { q[0] = b[0]; q[1] = b[1]; q[2] = b[2]; return; }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    for (int k = 0; k < 3; k++) q[k] = a[k] + v * ab[k];
    return;
  }
  float cp[3];
  sub(p, c, cp);
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) { q[0] = c[0]; q[1] = c[1]; q[2] = c[2]; return; }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    for (int k = 0; k < 3; k++) q[k] = a[k] + w * ac[k];
    return;
  }
  const float va = d3 * d6 - d5 * d4;
  if (exact && va <= 0.0f && d4 - d3 >= 0.0f && d5 - d6 >= 0.0f) {  // edge b-c, absent from the reference
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; k++) q[k] = b[k] + w * (c[k] - b[k]);
    return;
  }
  const float sum = va + vb + vc;
  const float v = vb / sum, w = vc / sum;
  for (int k = 0; k < 3; k++) q[k] = a[k] + v * ab[k] + w * ac[k];
}

void orc_mesh_sdf(const float* tris, uint32_t n, uint32_t rx, uint32_t ry, uint32_t rz, uint32_t cube, float cs,
                  int exact, float* sdf) {
  const size_t nc = size_t(cube) * cube * cube;
  for (size_t i = 0; i < nc; i++) sdf[i] = 1e9f;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t iz = 0; iz < int64_t(rz); iz++)
    for (uint32_t iy = 0; iy < ry; iy++)
      for (uint32_t ix = 0; ix < rx; ix++) {
        const float p[3] = {(float(ix) + 0.5f) * cs, (float(iy) + 0.5f) * cs, (float(iz) + 0.5f) * cs};
        float best = 1e9f;
        bool outer = true;
        for (uint32_t t = 0; t < n; t++) {
          const float* tr = tris + 12 * size_t(t);
          float q[3];
          orc_closest(p, tr + 3, tr + 6, tr + 9, exact != 0, q);
          const float d3[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
          const float d = std::sqrt((d3[0] * d3[0] + d3[1] * d3[1]) + d3[2] * d3[2]);
          if (d < best) {
            best = d;
            outer = (d3[0] * tr[0] + d3[1] * tr[1]) + d3[2] * tr[2] >= 0.0f;
          }
        }
        sdf[oracle::morton_encode(int(ix), int(iy), int(iz))] = outer ? best : -best;
      }
}


}  // extern "C"

// Forward-mode substeps over a dual-number type D (Dual or Dual32).
template <class D>
static void dual_substeps_impl(const orc_scene* sc, const void* mats, const void* cols, const void* forces,
                               const float* sdf, const double* fields, const uint32_t* matid, uint32_t n,
                               const double* tan_fields, const double* tan_mats, uint32_t nsub, double* out_fields,
                               double* out_tan) {
  orc_sim* sim = orc_sim_create(sc, mats, cols, forces, sdf, 64);
  Params p = sim->params();
  p.mat_tangent = tan_mats;
  const uint32_t NC = sc->cube_res * sc->cube_res * sc->cube_res;
  State<D> st;
  st.resize(n, NC);
  const size_t per = size_t(n) * 4;
  for (int f = 0; f < 8; f++)
    for (size_t i = 0; i < per; i++) {
      D& x = st.f[f][i];
      x = D(fields[size_t(f) * per + i]);
      for (int k = 0; k < kTangents; k++) x.d[k] = decltype(x.v)(tan_fields[(size_t(k) * 8 + f) * per + i]);
    }
  for (uint32_t i = 0; i < n; i++) st.matid[i] = matid[i];
  for (uint32_t i = 0; i < nsub; i++) substep<D>(p, st);
  for (int f = 0; f < 8; f++)
    for (size_t i = 0; i < per; i++) {
      const D& x = st.f[f][i];
      out_fields[size_t(f) * per + i] = double(x.v);
      for (int k = 0; k < kTangents; k++) out_tan[(size_t(k) * 8 + f) * per + i] = double(x.d[k]);
    }
  orc_sim_destroy(sim);
}

extern "C" {

// Forward-mode derivatives of nsub substeps (the FP64 twin with kTangents
// tangents; oracle/dual.hpp).  fields: 8 arrays of n double4 (the reference
// layout, the material-id lanes ignored), matid: n ids; tan_fields:
// kTangents x (8 x n x 4) input tangents; tan_mats: [material][mu, lambda,
// M_friction, q_cohesion][tangent] or null.  Writes the values after nsub
// substeps to out_fields and their tangents to out_tan (same layouts).
int orc_dual_tangents(void) { return kTangents; }
void orc_dual_substeps(const orc_scene* sc, const void* mats, const void* cols, const void* forces, const float* sdf,
                       const double* fields, const uint32_t* matid, uint32_t n, const double* tan_fields,
                       const double* tan_mats, uint32_t nsub, double* out_fields, double* out_tan) {
  dual_substeps_impl<Dual>(sc, mats, cols, forces, sdf, fields, matid, n, tan_fields, tan_mats, nsub, out_fields,
                           out_tan);
}
// The same with FP32 values and tangents (Dual32): the FP32 oracle's map
// differentiated in FP32.
void orc_dual32_substeps(const orc_scene* sc, const void* mats, const void* cols, const void* forces, const float* sdf,
                         const double* fields, const uint32_t* matid, uint32_t n, const double* tan_fields,
                         const double* tan_mats, uint32_t nsub, double* out_fields, double* out_tan) {
  dual_substeps_impl<Dual32>(sc, mats, cols, forces, sdf, fields, matid, n, tan_fields, tan_mats, nsub, out_fields,
                             out_tan);
}

}  // extern "C"
