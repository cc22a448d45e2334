// mpm_engine.hip — host side of the MI355X MPM engine: lifecycle, state
// setters/getters and substep orchestration behind the C ABI declared in
// include/nebula_mpm.h.  Counterpart of the reference's MPMEngine
// (src/engine/MPMEngine.cpp) with HIP allocations and launches in place of the
// bindless Vulkan buffer table and pipelines.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <random>
#include <vector>

#include "../../include/nebula_mpm.h"
#include "../../include/nebula_emitter.hpp"
#include "nb_device_math.hpp"
#include "nb_engine.hpp"

using namespace nb;

static_assert(sizeof(nb_material) == 64, "nb_material mirrors the 64-byte MaterialParams");
static_assert(sizeof(nb_collider) == 64, "nb_collider mirrors the 64-byte ColliderPrimitive");
static_assert(sizeof(nb_force) == 32, "nb_force mirrors the 32-byte ForceGPU");

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define NB_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(NB_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

enum class GridMode { NONE, SPARSE, DENSE };

}  // namespace

namespace nb {
// the thread-local message nb_last_error() returns, for the other entry-point files
int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace nb

struct nb_mpm {
  nb_mpm_config cfg{};
  nb_mpm_knobs knobs{};
  uint32_t grid_res[3] = {1, 1, 1};
  uint32_t cube_res = 1, NC = 1, tile_dim = 1, NT = 1, cap = 0, cap_tiles = 1;
  uint32_t n_live = 0;
  hipStream_t stream = nullptr;  // where work is enqueued
  hipStream_t owned = nullptr;   // the stream the engine created, if any
  DevState d{};
  uint32_t material_count = 1, collider_count = 0, force_count = 0, noise_force_count = 0;
  std::vector<nb_material> mats_host;  // host copy of the material table
  nb_mpm_svd_fallback svd_fb{1u, 1e-4f, 1e-6f, 0u};
  bool fuse_g2p = false;  // the grid update inside G2P for all but a call's last substep (nb_mpm_set_fused_g2p)
  uint32_t prof_calls[kNumProfiledKernels] = {};  // launches per kernel in the last nb_mpm_profile_substeps
  unsigned long long* d_svd_stats = nullptr;  // NB_SVD_FB_STATS counters
  bool has_sdf = false;
  bool bins_valid = false;
  // The kernels work on d.rec (one record per particle); the ABI exposes the
  // reference layout d.field.  rec_stale: the API wrote d.field since the
  // records were last filled.  ptr_exported: device pointers to d.field were
  // handed out, so the caller may write them between calls.  keys_valid: the
  // last G2P wrote every live particle's cell key and nothing moved since.
  bool rec_stale = true;
  bool ptr_exported = false;
  bool keys_valid = false;
  GridMode grid_mode = GridMode::NONE;
  bool old_valid = false;
  std::vector<void*> allocs;

  // batch of scenes (n_scenes > 0): the scene table, the scene of every tile
  // and of every particle row; the single-scene grid fields above are unused
  uint32_t n_scenes = 0;
  std::vector<SceneGeo> scenes_host;
  std::vector<uint32_t> scene_first;  // first particle row of each scene
  SceneGeo* d_scenes = nullptr;
  uint32_t* d_tile_scene = nullptr;
  uint32_t* d_pscene = nullptr;
  bool batch() const { return n_scenes > 0; }

  // emitters (MPMEngine.h:162-166)
  std::vector<nb_emitter> emitters;
  std::vector<int> emit_done;
  std::mt19937 emit_rng{12345u};

  // launch shapes and the captured substep (nb_mpm_graph_substeps)
  nb_mpm_tuning tuning{2048u, 128u, 12u, 0.f};  // measured best on the 4.19 M-particle snow scene
  hipGraphExec_t graph_exec = nullptr;
  SubParams graph_p{};
  DevState graph_d{};
  uint32_t max_materials() const { return batch() ? kBatchMaxMaterials : NB_MAX_MATERIALS; }

  // reverse mode: checkpoints of a recorded run of n substeps every `seg`,
  // the substep states of one segment, the adjoint in both layouts
  struct Adjoint {
    bool on = false, recorded = false;
    uint32_t n = 0, seg = 0, n_ckpt = 0;
    float dt = 0.f;
    std::vector<void*> allocs;
    // The state before substep s = sg * seg + j is slot(sg, j): ckpt[sg] for
    // j = 0, tape t = sg % n_tapes at j - 1 otherwise.  The recorded run
    // writes every slot; the tapes then hold the last n_tapes segments, and
    // the backward recomputes a segment's tape from its checkpoint when the
    // tape holds another segment.
    uint32_t n_tapes = 1;
    float4* ckpt = nullptr;   // n_ckpt * cap * kRec
    float4* tape = nullptr;   // n_tapes * (seg - 1) * cap * kRec
    uint32_t* ckpt_keys = nullptr;  // bin keys of the same states: n_ckpt * cap
    uint32_t* tape_keys = nullptr;  // n_tapes * (seg - 1) * cap
    // grids of substep j of tape t's segment: node velocities after the
    // update at grid_tape[2 (t seg + j)], (momentum / mass, mass) before it
    // at [2 (t seg + j) + 1]
    float4* grid_tape = nullptr;    // n_tapes * seg * 2 * NT * 512
    std::vector<int> tape_seg;      // n_tapes: the segment whose states each tape holds (-1: none)
    // the bins substep j of tape t's segment built (sorted order, cell counts
    // and ends, tile lists) at t seg + j: the backward reads them instead of
    // binning again
    std::vector<DevState> slot_bins;  // n_tapes * seg
    float4* adj_a = nullptr;  // cap * kRec
    float4* adj_b = nullptr;
    float4* field[NB_NUM_FIELDS] = {};  // adjoint in the reference layout
    AdjState a{};
  } adj;

  template <class T>
  int alloc(T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(NB_ERR_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    allocs.push_back(q);
    *p = reinterpret_cast<T*>(q);
    return NB_OK;
  }

  // device pointers for a launch: the SDF only while a collider SDF is set
  DevState dev() const {
    DevState v = d;
    if (!has_sdf) v.sdf = nullptr;
    return v;
  }

  SubParams params(float dt, bool write_old) const {
    SubParams p;
    std::memset(&p, 0, sizeof(p));  // padding too: captured substeps are matched bytewise
    p.dt = dt;
    p.h = cfg.cell_size;
    p.h_pow2 = cell_size_is_pow2(cfg.cell_size);
    for (int a = 0; a < 3; a++) {
      p.wmin[a] = 0.f;  // worldMin = 0, worldMax = domainSize (src/engine/MPMEngine.cpp:373-374)
      p.wmax[a] = cfg.domain_size[a];
      p.grid_res[a] = int32_t(grid_res[a]);
    }
    p.cube_res = cube_res;
    p.tile_dim = tile_dim;
    p.n = n_live;
    p.flip_ratio = knobs.flip_ratio;
    p.material_count = material_count;
    p.collider_count = collider_count;
    p.force_count = force_count;
    p.noise_forces = noise_force_count;
    p.write_old = write_old ? 1u : 0u;
    p.corrected_svd = knobs.corrected_svd ? 1u : 0u;
    // the fallback serves the Drucker-Prager family only: with no such
    // material G2P queues nothing and its fix-up kernel is not launched
    bool dp_family = false;
    for (uint32_t m = 0; m < material_count && m < mats_host.size(); m++) {
      const uint32_t model = mats_host[m].model;
      dp_family = dp_family || model == NB_MAT_DRUCKER_PRAGER || model == NB_MAT_GRANULAR_POWDER ||
                  model == NB_MAT_VISCOPLASTIC_MUD;
    }
    p.svd_fp64 = (svd_fb.enabled && dp_family) ? 1u : 0u;
    p.svd_gap_rel = svd_fb.gap_rel;
    p.svd_zero_rel = svd_fb.zero_rel;
    p.svd_stats = svd_fb.count_stats ? d_svd_stats : nullptr;
    p.fb_queue = d.fb_queue;
    p.fb_ctl = d.fb_ctl;
    p.fb_region = (n_live + kFbSlots - 1) / kFbSlots;
    p.tile_blocks = tuning.tile_blocks;
    p.g2p_block = tuning.g2p_block;
    p.p2g_item = tuning.p2g_item;
    if (batch()) {
      p.n_scenes = n_scenes;
      p.scenes = d_scenes;
      p.tile_scene = d_tile_scene;
      p.pscene = d_pscene;
    }
    return p;
  }
};

namespace {

// src/core/Domain.h:14-37
void domain_dims(const nb_mpm_config& c, uint32_t res[3], uint32_t& cube) {
  const float h = std::max(c.cell_size, 1e-6f);
  uint32_t m = 1;
  for (int a = 0; a < 3; a++) {
    float r = std::ceil(c.domain_size[a] / h);
    res[a] = r < 1.f ? 1u : uint32_t(r);
    m = std::max(m, res[a]);
  }
  cube = 1;
  while (cube < m) cube <<= 1;
}

nb_material preset_jelly(float E, float nu, float rho0) {  // src/core/MaterialParams.h:42-58
  nb_material m{};
  m.mu = E / (2.0f * (1.0f + nu));
  m.lambda = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
  m.rho0 = rho0;
  m.model = NB_MAT_ELASTIC;
  return m;
}

int upload(nb_mpm* h, void* dst, const void* src, size_t bytes) {
  if (!bytes) return NB_OK;
  NB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, h->stream));
  NB_HIP(hipStreamSynchronize(h->stream));
  return NB_OK;
}
int download(nb_mpm* h, void* dst, const void* src, size_t bytes) {
  if (!bytes) return NB_OK;
  NB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, h->stream));
  NB_HIP(hipStreamSynchronize(h->stream));
  return NB_OK;
}

// Refresh the particle records from the reference-layout arrays when those may
// have changed since the records were last written back.
int sync_in(nb_mpm* h) {
  if ((h->rec_stale || h->ptr_exported) && h->n_live) {
    launch_soa_to_aos(h->d, h->n_live, h->stream);
    NB_HIP(hipGetLastError());
    h->keys_valid = false;
  }
  h->rec_stale = false;
  return NB_OK;
}

// Write the particle records back to the reference-layout arrays.
int sync_out(nb_mpm* h) {
  if (h->n_live) {
    launch_aos_to_soa(h->d, h->n_live, h->stream);
    NB_HIP(hipGetLastError());
  }
  return NB_OK;
}

// one substep on the records (the caller brackets a run of them with
// sync_in / sync_out)
// Where a recorded substep reads and writes: records and bin keys before
// and after, and the grids it keeps.  Null members mean the engine's own
// buffers (records and keys updated in place); with g2p = false the substep
// stops after the grid update.
// the bin buffers of dv replaced by those of b
void use_bins(DevState& dv, const DevState& b) {
  dv.rank = b.rank;
  dv.sorted = b.sorted;
  dv.cell_count = b.cell_count;
  dv.cell_end = b.cell_end;
  dv.tile_flag = b.tile_flag;
  dv.tile_mark = b.tile_mark;
  dv.blk_cnt = b.blk_cnt;
  dv.tile_act = b.tile_act;
  dv.act_tiles = b.act_tiles;
  dv.node_tiles = b.node_tiles;
  dv.tile_total = b.tile_total;
  dv.tile_start = b.tile_start;
  dv.tile_order = b.tile_order;
  dv.counters = b.counters;
}

struct TapeIO {
  float4* rec_in = nullptr;
  float4* rec_out = nullptr;
  uint32_t* keys_in = nullptr;
  uint32_t* keys_out = nullptr;
  float4* grid_node = nullptr;
  float4* grid_pre = nullptr;
  bool g2p = true;
  const DevState* bins = nullptr;  // bin buffers to use instead of the engine's
};

// One substep.  fused: the grid update runs inside G2P (k_g2p_tile) and the
// grid arrays are left unwritten; only for plain substeps with no tape, and
// the same results bit for bit.
int enqueue_substep(nb_mpm* h, float dt, Prof* pf, const TapeIO& io = TapeIO{}, bool fused = false) {
  if (h->n_live == 0) return NB_OK;
  const bool flip = h->knobs.flip_ratio > 0.01f;
  SubParams p = h->params(dt, flip);
  DevState dv = h->dev();
  if (fused && p.noise_forces == 0 && !io.bins && !io.rec_in && !io.keys_in && !io.grid_node && !io.grid_pre &&
      io.g2p && !io.keys_out && !io.rec_out) {
    launch_bin(dv, p, h->NT, h->keys_valid, h->stream, pf);
    launch_p2g(dv, p, h->NT, h->stream, pf);
    launch_g2p_fused(dv, p, h->NT, h->stream, pf);
    NB_HIP(hipGetLastError());
    h->bins_valid = true;
    h->keys_valid = true;
    h->grid_mode = GridMode::NONE;  // the fused substep keeps its grid in LDS
    h->old_valid = false;
    return NB_OK;
  }
  if (io.bins) use_bins(dv, *io.bins);
  if (io.rec_in) dv.rec = io.rec_in;
  if (io.keys_in) dv.keys = io.keys_in;
  if (io.grid_node) dv.grid_node = io.grid_node;
  dv.grid_pre = io.grid_pre;
  launch_bin(dv, p, h->NT, h->keys_valid, h->stream, pf);
  launch_p2g(dv, p, h->NT, h->stream, pf);
  launch_grid(dv, p, h->NT, true, true, dv.node_tiles, dv.counters + CNT_NODE_TILES, h->stream, pf);
  if (!io.g2p) {
    NB_HIP(hipGetLastError());
    h->bins_valid = true;
    h->keys_valid = true;
    return NB_OK;
  }
  dv.keys = io.keys_out ? io.keys_out : h->d.keys;  // G2P writes the next substep's keys
  launch_g2p(dv, p, true, h->stream, pf, io.rec_out);
  NB_HIP(hipGetLastError());
  h->bins_valid = true;
  h->keys_valid = true;
  h->grid_mode = GridMode::SPARSE;
  h->old_valid = flip;
  return NB_OK;
}

int engine_alloc(nb_mpm* h, void* hip_stream);

int run_substeps(nb_mpm* h, float dt, uint32_t n, Prof* pf) {
  if (n == 0 || h->n_live == 0) return NB_OK;
  int rc = sync_in(h);
  // every substep but the last fuses the grid update into G2P; the last one
  // writes the grid, so that the grid a caller reads after the call is the
  // last substep's, as after the reference's step
  for (uint32_t i = 0; i < n && rc == NB_OK; i++)
    rc = enqueue_substep(h, dt, pf, TapeIO{}, h->fuse_g2p && i + 1 < n);
  if (rc == NB_OK) rc = sync_out(h);
  return rc;
}

}  // namespace

extern "C" {

const char* nb_last_error(void) { return g_err.c_str(); }
int nb_mpm_abi_version(void) { return NB_MPM_ABI_VERSION; }

void nb_mpm_config_default(nb_mpm_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->nx = c->ny = c->nz = 32;
  c->domain_size[0] = c->domain_size[1] = c->domain_size[2] = 10.0f;
  c->cell_size = 10.0f / 64.0f;
  c->max_particles = 0;
  c->E = 1e4f;
  c->nu = 0.3f;
  c->rho0 = 1000.0f;
}

void nb_mpm_knobs_default(nb_mpm_knobs* k) {
  if (!k) return;
  std::memset(k, 0, sizeof(*k));
  k->restitution = 0.3f;
  k->wall_friction = 0.0f;
  k->num_substeps = 20;
  k->plastic_model = 0;
  k->q_max = 1e5f;
  k->M_friction = 0.577f;
  k->q_cohesion = 0.0f;
  k->flip_ratio = 0.0f;
  k->corrected_svd = 0;
}

int nb_mpm_create(const nb_mpm_config* cfg, void* hip_stream, nb_mpm** out) {
  if (!cfg || !out) return fail(NB_ERR_ARG, "nb_mpm_create: null argument");
  *out = nullptr;
  if (!(cfg->cell_size > 0.f)) return fail(NB_ERR_ARG, "cell_size must be > 0");
  for (int a = 0; a < 3; a++)
    if (!(cfg->domain_size[a] >= 0.f)) return fail(NB_ERR_ARG, "domain_size must be >= 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NB_ERR_HIP, "no HIP device available");

  nb_mpm* h = new nb_mpm();
  h->cfg = *cfg;
  nb_mpm_knobs_default(&h->knobs);
  domain_dims(*cfg, h->grid_res, h->cube_res);
  if (h->cube_res > 1024u) {
    const uint32_t cr = h->cube_res;
    delete h;
    return fail(NB_ERR_RANGE, "grid of %u cells per axis exceeds the 10-bit Morton range", cr);
  }
  h->NC = h->cube_res * h->cube_res * h->cube_res;
  h->tile_dim = std::max(h->cube_res, 8u) / 8u;
  h->NT = h->tile_dim * h->tile_dim * h->tile_dim;
  const uint64_t lattice = uint64_t(cfg->nx) * cfg->ny * cfg->nz;
  const uint64_t cap = cfg->max_particles ? cfg->max_particles : lattice;
  if (cap > 0xFFFFFFF0ull || lattice > cap) {
    delete h;
    return fail(NB_ERR_RANGE, "particle capacity %llu invalid (lattice %llu)", (unsigned long long)cap,
                (unsigned long long)lattice);
  }
  h->cap = uint32_t(cap);
  h->cap_tiles = std::max<uint32_t>(1u, std::min<uint32_t>(h->NT, std::max<uint32_t>(h->cap, 1u)));

  auto cleanup_fail = [&](int rc) {
    std::string keep = g_err;
    nb_mpm_destroy(h);
    g_err = keep;
    return rc;
  };

  int rc = engine_alloc(h, hip_stream);
  if (rc != NB_OK) return cleanup_fail(rc);
  DevState& d = h->d;

  // default material (src/engine/MPMEngine.cpp:67-78)
  {
    nb_material m = preset_jelly(cfg->E, cfg->nu, cfg->rho0);
    m.model = NB_MAT_ELASTIC;
    m.M_friction = h->knobs.M_friction;
    m.q_cohesion = h->knobs.q_cohesion;
    m.q_max = h->knobs.q_max;
    if ((rc = upload(h, d.materials, &m, sizeof(m))) != NB_OK) return cleanup_fail(rc);
    h->material_count = 1;
    h->mats_host.assign(1, m);
  }

  // default particle block (src/engine/MPMEngine.cpp:86-137)
  h->n_live = uint32_t(lattice);
  if (lattice) {
    const float block_frac = 0.40f;
    const float min_dom = std::min({cfg->domain_size[0], cfg->domain_size[1], cfg->domain_size[2]});
    const float sp = min_dom * block_frac / float(std::max({cfg->nx, cfg->ny, cfg->nz}));
    const float Vp = sp * sp * sp;
    const float cx = cfg->domain_size[0] * 0.5f, cz = cfg->domain_size[2] * 0.5f, cy = cfg->domain_size[1] * 0.70f;
    const float hx = sp * float(cfg->nx - 1) / 2.0f, hy = sp * float(cfg->ny - 1) / 2.0f,
                hz = sp * float(cfg->nz - 1) / 2.0f;
    std::vector<float> pos(size_t(lattice) * 4);
    size_t idx = 0;
    for (uint32_t iz = 0; iz < cfg->nz; iz++)
      for (uint32_t iy = 0; iy < cfg->ny; iy++)
        for (uint32_t ix = 0; ix < cfg->nx; ix++, idx++) {
          pos[4 * idx + 0] = cx - hx + ix * sp;
          pos[4 * idx + 1] = cy - hy + iy * sp;
          pos[4 * idx + 2] = cz - hz + iz * sp;
          pos[4 * idx + 3] = Vp;
        }
    if ((rc = upload(h, d.field[NB_FIELD_P], pos.data(), pos.size() * 4)) != NB_OK) return cleanup_fail(rc);
    launch_init_particles(d, 0, h->n_live, h->stream);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup_fail(fail(NB_ERR_HIP, "init failed"));
  }
  *out = h;
  return NB_OK;
}

int nb_mpm_create_batch(const nb_mpm_scene* scenes, uint32_t n_scenes, void* hip_stream, nb_mpm** out) {
  if (!scenes || !out) return fail(NB_ERR_ARG, "nb_mpm_create_batch: null argument");
  *out = nullptr;
  if (n_scenes == 0) return fail(NB_ERR_ARG, "a batch needs at least one scene");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NB_ERR_HIP, "no HIP device available");
  // scene table: each scene's Morton cube at the next multiple of 512 keys
  std::vector<SceneGeo> geo(n_scenes);
  std::vector<uint32_t> first(n_scenes);
  uint64_t tiles = 0, particles = 0;
  for (uint32_t s = 0; s < n_scenes; s++) {
    const nb_mpm_scene& sc = scenes[s];
    if (!(sc.cell_size > 0.f)) return fail(NB_ERR_ARG, "scene %u: cell_size must be > 0", s);
    nb_mpm_config c{};
    for (int a = 0; a < 3; a++) {
      if (!(sc.domain_size[a] >= 0.f)) return fail(NB_ERR_ARG, "scene %u: domain_size must be >= 0", s);
      c.domain_size[a] = sc.domain_size[a];
    }
    c.cell_size = sc.cell_size;
    uint32_t res[3], cube = 1;
    domain_dims(c, res, cube);
    if (cube > 1024u) return fail(NB_ERR_RANGE, "scene %u: %u cells per axis exceed the 10-bit Morton range", s, cube);
    SceneGeo& g = geo[s];
    g = SceneGeo{};
    g.h = sc.cell_size;
    g.inv_h = 1.0f / sc.cell_size;
    g.h_pow2 = cell_size_is_pow2(sc.cell_size);
    for (int a = 0; a < 3; a++) {
      g.wmin[a] = 0.f;
      g.wmax[a] = sc.domain_size[a];
      g.grid_res[a] = int32_t(res[a]);
    }
    g.cube_res = cube;
    g.tile_dim = std::max(cube, 8u) / 8u;
    g.key_base = uint32_t(tiles * kTileCells);
    tiles += uint64_t(g.tile_dim) * g.tile_dim * g.tile_dim;
    if (tiles * kTileCells > 0xFFFFFFFFull) return fail(NB_ERR_RANGE, "the scenes' cells exceed the 32-bit key space");
    first[s] = uint32_t(particles);
    particles += sc.particle_count;
    if (particles > 0xFFFFFFF0ull) return fail(NB_ERR_RANGE, "too many particles in the batch");
  }

  nb_mpm* h = new nb_mpm();
  auto cleanup_fail = [&](int rc) {
    std::string keep = g_err;
    nb_mpm_destroy(h);
    g_err = keep;
    return rc;
  };
  nb_mpm_config_default(&h->cfg);
  h->cfg.cell_size = geo[0].h;
  for (int a = 0; a < 3; a++) {
    h->cfg.domain_size[a] = geo[0].wmax[a];
    h->grid_res[a] = uint32_t(geo[0].grid_res[a]);
  }
  h->cfg.nx = h->cfg.ny = h->cfg.nz = 0;
  h->cfg.max_particles = uint32_t(particles);
  nb_mpm_knobs_default(&h->knobs);
  h->cube_res = geo[0].cube_res;
  h->tile_dim = geo[0].tile_dim;
  h->NT = uint32_t(tiles);
  h->NC = h->NT * kTileCells;
  h->cap = uint32_t(particles);
  h->cap_tiles = std::max<uint32_t>(1u, std::min<uint32_t>(h->NT, std::max<uint32_t>(h->cap, 1u)));
  h->n_scenes = n_scenes;
  h->scenes_host = geo;
  h->scene_first = first;
  int rc = engine_alloc(h, hip_stream);
  if (rc == NB_OK) rc = h->alloc(&h->d_scenes, n_scenes);
  if (rc == NB_OK) rc = h->alloc(&h->d_tile_scene, h->NT);
  if (rc == NB_OK) rc = h->alloc(&h->d_pscene, h->cap);
  if (rc != NB_OK) return cleanup_fail(rc);
  {
    std::vector<uint32_t> ts(h->NT), ps(std::max(h->cap, 1u));
    for (uint32_t s = 0; s < n_scenes; s++) {
      const uint32_t t0 = geo[s].key_base / kTileCells, nt = geo[s].tile_dim * geo[s].tile_dim * geo[s].tile_dim;
      std::fill(ts.begin() + t0, ts.begin() + t0 + nt, s);
      std::fill(ps.begin() + first[s], ps.begin() + first[s] + scenes[s].particle_count, s);
    }
    if ((rc = upload(h, h->d_scenes, geo.data(), geo.size() * sizeof(SceneGeo))) != NB_OK) return cleanup_fail(rc);
    if ((rc = upload(h, h->d_tile_scene, ts.data(), size_t(h->NT) * 4)) != NB_OK) return cleanup_fail(rc);
    if ((rc = upload(h, h->d_pscene, ps.data(), size_t(h->cap) * 4)) != NB_OK) return cleanup_fail(rc);
  }
  // default material as nb_mpm_create's, from nb_mpm_config_default's E, nu, rho0
  {
    nb_material m = preset_jelly(h->cfg.E, h->cfg.nu, h->cfg.rho0);
    m.model = NB_MAT_ELASTIC;
    m.M_friction = h->knobs.M_friction;
    m.q_cohesion = h->knobs.q_cohesion;
    m.q_max = h->knobs.q_max;
    if ((rc = upload(h, h->d.materials, &m, sizeof(m))) != NB_OK) return cleanup_fail(rc);
    h->material_count = 1;
    h->mats_host.assign(1, m);
  }
  // every particle starts at rest with F = I, B = 0, tau = 0 and material 0;
  // positions are the caller's (nb_mpm_set_field)
  h->n_live = h->cap;
  if (h->n_live) launch_init_particles(h->d, 0, h->n_live, h->stream);
  if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup_fail(fail(NB_ERR_HIP, "init failed"));
  *out = h;
  return NB_OK;
}

int nb_mpm_scene_count(const nb_mpm* h, uint32_t* n) {
  if (!h || !n) return fail(NB_ERR_ARG, "null argument");
  *n = h->n_scenes;
  return NB_OK;
}

int nb_mpm_scene_info(const nb_mpm* h, uint32_t scene, uint32_t res[3], uint32_t* first_particle,
                      uint32_t* particle_count) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (!h->batch()) return fail(NB_ERR_STATE, "not a batch");
  if (scene >= h->n_scenes) return fail(NB_ERR_RANGE, "scene %u of %u", scene, h->n_scenes);
  const SceneGeo& g = h->scenes_host[scene];
  if (res)
    for (int a = 0; a < 3; a++) res[a] = uint32_t(g.grid_res[a]);
  const uint32_t f = h->scene_first[scene];
  const uint32_t e = scene + 1 < h->n_scenes ? h->scene_first[scene + 1] : h->cap;
  if (first_particle) *first_particle = f;
  if (particle_count) *particle_count = e - f;
  return NB_OK;
}

}  // extern "C"

namespace {
// Device buffers of an engine whose sizes (cap, NT, cap_tiles) are set.
int engine_alloc(nb_mpm* h, void* hip_stream) {
  if (hip_stream) {
    h->stream = reinterpret_cast<hipStream_t>(hip_stream);
  } else {
    if (hipStreamCreateWithFlags(&h->owned, hipStreamNonBlocking) != hipSuccess)
      return fail(NB_ERR_HIP, "hipStreamCreate failed");
    h->stream = h->owned;
  }

  DevState& d = h->d;
  const size_t cells = size_t(h->NT) * kTileCells;
  int rc = NB_OK;
  for (int f = 0; f < NB_NUM_FIELDS && rc == NB_OK; f++) rc = h->alloc(&d.field[f], h->cap);
  if (rc == NB_OK) rc = h->alloc(&d.rec, size_t(h->cap) * kRec);
  if (rc == NB_OK) rc = h->alloc(&d.keys, h->cap);
  if (rc == NB_OK) rc = h->alloc(&d.rank, h->cap);
  if (rc == NB_OK) rc = h->alloc(&d.sorted, h->cap);
  if (rc == NB_OK) rc = h->alloc(&d.cell_count, cells);
  if (rc == NB_OK) rc = h->alloc(&d.cell_end, cells);
  if (rc == NB_OK) rc = h->alloc(&d.tile_flag, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.tile_mark, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.blk_cnt, (h->NT + 255) / 256);
  if (rc == NB_OK) rc = h->alloc(&d.tile_act, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.act_tiles, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.node_tiles, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.all_tiles, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.tile_total, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.tile_start, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.tile_order, h->NT);
  if (rc == NB_OK) rc = h->alloc(&d.counters, CNT_SLOTS);
  if (rc == NB_OK) rc = h->alloc(&d.partial, size_t(h->cap_tiles) * kBoxNodes);
  if (rc == NB_OK) rc = h->alloc(&d.grid_node, cells);
  if (rc == NB_OK) rc = h->alloc(&d.grid_old, cells);
  if (rc == NB_OK) rc = h->alloc(&d.materials, h->max_materials());
  if (rc == NB_OK) rc = h->alloc(&d.colliders, NB_MAX_COLLIDERS);
  if (rc == NB_OK) rc = h->alloc(&d.forces, NB_MAX_FORCES);
  if (rc == NB_OK) rc = h->alloc(&h->d_svd_stats, NB_SVD_FB_STATS);
  if (rc == NB_OK) rc = h->alloc(&d.units, size_t(h->NT) + h->cap / kG2PUnit + 1);
  if (rc == NB_OK) rc = h->alloc(&d.fb_queue, (size_t(h->cap) + kFbSlots) * kFbEntry);
  if (rc == NB_OK) rc = h->alloc(&d.fb_ctl, kFbSlots * kFbCtlStride + 1);
  if (rc != NB_OK) return rc;
  d.sdf = nullptr;

  auto zero = [&](void* p, size_t bytes) { return hipMemsetAsync(p, 0, bytes, h->stream); };
  if (zero(d.cell_count, cells * 4) != hipSuccess || zero(d.cell_end, cells * 4) != hipSuccess ||
      zero(d.tile_flag, h->NT) != hipSuccess || zero(d.counters, CNT_SLOTS * 4) != hipSuccess ||
      zero(d.grid_node, cells * 16) != hipSuccess || zero(d.grid_old, cells * 16) != hipSuccess ||
      zero(d.materials, size_t(h->max_materials()) * 64) != hipSuccess ||
      zero(d.colliders, NB_MAX_COLLIDERS * 64) != hipSuccess || zero(d.forces, NB_MAX_FORCES * 32) != hipSuccess ||
      zero(h->d_svd_stats, NB_SVD_FB_STATS * 8) != hipSuccess || zero(d.fb_ctl, (kFbSlots * kFbCtlStride + 1) * 4) != hipSuccess)
    return fail(NB_ERR_HIP, "hipMemsetAsync failed");
  for (int f = 0; f < NB_NUM_FIELDS; f++)
    if (zero(d.field[f], size_t(std::max(h->cap, 1u)) * 16) != hipSuccess)
      return fail(NB_ERR_HIP, "hipMemsetAsync failed");
  if (hipMemsetAsync(d.tile_act, 0xFF, size_t(h->NT) * 4, h->stream) != hipSuccess)
    return fail(NB_ERR_HIP, "hipMemsetAsync failed");
  launch_fill_identity(d.all_tiles, h->NT, h->stream);
  launch_fill_identity(d.tile_order, h->NT, h->stream);
  {
    uint32_t all = h->NT;
    if (hipMemcpyAsync(d.counters + CNT_ALL_TILES, &all, 4, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      return fail(NB_ERR_HIP, "counter upload failed");
  }
  return NB_OK;
}
}  // namespace

extern "C" {

int nb_mpm_destroy(nb_mpm* h) {
  if (!h) return NB_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  for (void* p : h->allocs) (void)hipFree(p);
  for (void* p : h->adj.allocs) (void)hipFree(p);
  if (h->owned) (void)hipStreamDestroy(h->owned);
  delete h;
  return NB_OK;
}

int nb_mpm_step(nb_mpm* h, float dt) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (!h->emitters.empty()) {  // src/engine/MPMEngine.cpp:405-406
    const int rc = nb_mpm_emit(h, dt);
    if (rc != NB_OK) return rc;
  }
  const int ns = std::max(1, h->knobs.num_substeps);
  return run_substeps(h, dt / float(ns), uint32_t(ns), nullptr);
}

int nb_mpm_substep(nb_mpm* h, float dt_sub) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  return run_substeps(h, dt_sub, 1, nullptr);
}

int nb_mpm_substeps(nb_mpm* h, float dt_sub, uint32_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  return run_substeps(h, dt_sub, n, nullptr);
}

int nb_mpm_set_stream(nb_mpm* h, void* hip_stream) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (!hip_stream && !h->owned && hipStreamCreateWithFlags(&h->owned, hipStreamNonBlocking) != hipSuccess)
    return fail(NB_ERR_HIP, "hipStreamCreate failed");
  hipStream_t next = hip_stream ? reinterpret_cast<hipStream_t>(hip_stream) : h->owned;
  if (next == h->stream) return NB_OK;
  NB_HIP(hipStreamSynchronize(h->stream));  // what the old stream holds comes first
  h->stream = next;
  return NB_OK;
}

int nb_mpm_sync(nb_mpm* h) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  NB_HIP(hipStreamSynchronize(h->stream));
  NB_HIP(hipGetLastError());
  return NB_OK;
}

int nb_mpm_get_tuning(const nb_mpm* h, nb_mpm_tuning* t) {
  if (!h || !t) return fail(NB_ERR_ARG, "null argument");
  *t = h->tuning;
  return NB_OK;
}

int nb_mpm_set_tuning(nb_mpm* h, const nb_mpm_tuning* t) {
  if (!h || !t) return fail(NB_ERR_ARG, "null argument");
  if (t->tile_blocks < 1u || t->tile_blocks > 65536u) return fail(NB_ERR_RANGE, "tile_blocks %u not in [1, 65536]", t->tile_blocks);
  if (t->g2p_block != 64u && t->g2p_block != 128u && t->g2p_block != 256u)
    return fail(NB_ERR_RANGE, "g2p_block %u: 64, 128 or 256", t->g2p_block);
  if (t->p2g_item < 4u || t->p2g_item > 32u) return fail(NB_ERR_RANGE, "p2g_item %u not in [4, 32]", t->p2g_item);
  h->tuning = *t;
  return NB_OK;
}

int nb_mpm_graph_substeps(nb_mpm* h, float dt_sub, uint32_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (n == 0 || h->n_live == 0) return NB_OK;
  int rc = sync_in(h);
  // the captured substep reads the keys the previous G2P wrote
  if (rc == NB_OK && !h->keys_valid) {
    rc = enqueue_substep(h, dt_sub, nullptr);
    n--;
  }
  if (rc == NB_OK && n) {
    const SubParams p = h->params(dt_sub, h->knobs.flip_ratio > 0.01f);
    const DevState dv = h->dev();
    if (!h->graph_exec || std::memcmp(&p, &h->graph_p, sizeof(p)) != 0 ||
        std::memcmp(&dv, &h->graph_d, sizeof(dv)) != 0) {
      if (h->graph_exec) {
        (void)hipGraphExecDestroy(h->graph_exec);
        h->graph_exec = nullptr;
      }
      NB_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_substep(h, dt_sub, nullptr);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(h->stream, &g);
      if (rc != NB_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc != NB_OK ? rc : fail(NB_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
      }
      const hipError_t ei = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        h->graph_exec = nullptr;
        return fail(NB_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
      }
      h->graph_p = p;
      h->graph_d = dv;
    }
    for (uint32_t i = 0; i < n; i++) NB_HIP(hipGraphLaunch(h->graph_exec, h->stream));
  }
  if (rc == NB_OK) rc = sync_out(h);
  return rc;
}

int nb_mpm_autotune(nb_mpm* h, float dt_sub, uint32_t trial_substeps, uint32_t flags, nb_mpm_tuning* chosen) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->n_live == 0) return fail(NB_ERR_STATE, "no particles to time");
  if (h->adj.on) return fail(NB_ERR_STATE, "not during a recorded run (nb_mpm_adjoint_end first)");
  const uint32_t reps = trial_substeps ? trial_substeps : 3u;
  int rc = sync_in(h);
  if (rc != NB_OK) return rc;
  const size_t rec_bytes = size_t(h->n_live) * kRec * sizeof(float4), key_bytes = size_t(h->n_live) * 4;
  void* save_rec = nullptr;
  void* save_keys = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto release = [&]() {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (save_rec) (void)hipFree(save_rec);
    if (save_keys) (void)hipFree(save_keys);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  if (hipMalloc(&save_rec, rec_bytes) != hipSuccess || hipMalloc(&save_keys, key_bytes) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    release();
    return fail(NB_ERR_HIP, "autotune: allocation failed");
  }
  const bool keys0 = h->keys_valid;
  const nb_mpm_tuning start = h->tuning;
  auto copy = [&](void* dst, const void* src, size_t b) {
    return hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
  };
  bool ok = copy(save_rec, h->d.rec, rec_bytes) && copy(save_keys, h->d.keys, key_bytes);
  auto restore = [&]() {
    ok = ok && copy(h->d.rec, save_rec, rec_bytes) && copy(h->d.keys, save_keys, key_bytes);
    h->keys_valid = keys0;
  };
  // one trial: from the saved state, one warm substep, then reps timed
  auto trial = [&](const nb_mpm_tuning& t) -> float {
    h->tuning = t;
    restore();
    int r = enqueue_substep(h, dt_sub, nullptr);
    ok = ok && r == NB_OK && hipEventRecord(e0, h->stream) == hipSuccess;
    for (uint32_t i = 0; i < reps && ok; i++) ok = enqueue_substep(h, dt_sub, nullptr) == NB_OK;
    ok = ok && hipEventRecord(e1, h->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0.f;
    ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    return ok ? ms / float(reps) : 1e30f;
  };
  // A candidate replaces the best shape only when it is 2 % faster, so that
  // timing noise does not move the choice; the final choice is timed again
  // against the starting shape below.
  nb_mpm_tuning best = start;
  float best_ms = trial(best);
  auto consider = [&](nb_mpm_tuning t) {
    const float ms = trial(t);
    if (ms < 0.98f * best_ms) {
      best_ms = ms;
      best = t;
    }
  };
  for (uint32_t tb : {256u, 512u, 1024u, 2048u, 4096u, 8192u}) {
    nb_mpm_tuning t = best;
    t.tile_blocks = tb;
    if (tb != best.tile_blocks) consider(t);
  }
  for (uint32_t gb : {64u, 128u, 256u}) {
    nb_mpm_tuning t = best;
    t.g2p_block = gb;
    if (gb != best.g2p_block) consider(t);
  }
  if (flags & NB_TUNE_ROUNDING) {
    for (uint32_t it : {6u, 8u, 12u, 16u, 24u}) {
      nb_mpm_tuning t = best;
      t.p2g_item = it;
      if (it != best.p2g_item) consider(t);
    }
  }
  // confirmation: the choice and the starting shape, timed alternately twice
  // each; the choice stands only if it is at least 1 % faster
  if (ok && std::memcmp(&best, &start, sizeof(best)) != 0) {
    float t_best = 0.f, t_start = 0.f;
    for (int r = 0; r < 2 && ok; r++) {
      t_start += trial(start);
      t_best += trial(best);
    }
    if (!(t_best < 0.99f * t_start)) best = start, best_ms = 0.5f * t_start;
    else best_ms = 0.5f * t_best;
  }
  restore();
  release();
  if (!ok) {
    h->tuning = start;
    return fail(NB_ERR_HIP, "autotune: a trial failed");
  }
  best.ms_per_substep = best_ms;
  h->tuning = best;
  // the records and keys are back; bins and grid are those of the last trial
  h->bins_valid = false;
  h->grid_mode = GridMode::NONE;
  h->old_valid = false;
  if (chosen) *chosen = best;
  return NB_OK;
}

}  // extern "C"

namespace {
// Bins of the current state (plain engine or batch), for the occupied-tile count.
int launch_bin_only(nb_mpm* h) {
  SubParams p = h->params(0.f, false);
  launch_bin(h->dev(), p, h->NT, h->keys_valid, h->stream, nullptr);
  NB_HIP(hipGetLastError());
  NB_HIP(hipStreamSynchronize(h->stream));
  h->bins_valid = true;
  return NB_OK;
}

// The cache key of nb_mpm_autotune_cached (see the header).
int tuning_key(nb_mpm* h, uint32_t flags, std::string& key) {
  int dev = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(NB_ERR_HIP, "autotune cache: device properties unavailable");
  uint32_t tiles = 0;
  const int rc = nb_mpm_active_tile_count(h, &tiles);
  if (rc != NB_OK) return rc;
  auto pow2 = [](uint32_t v) {  // nearest power of two (ties up)
    if (v == 0) return 0u;
    uint32_t lo = 1;
    while (lo * 2u <= v && lo < (1u << 31)) lo *= 2u;
    return (v - lo >= lo * 2u - v && lo < (1u << 31)) ? lo * 2u : lo;
  };
  unsigned models = 0;
  for (uint32_t m = 0; m < h->material_count && m < h->mats_host.size(); m++)
    models |= 1u << std::min<uint32_t>(h->mats_host[m].model, 31u);
  const float fr = h->knobs.flip_ratio;
  const char* mode = fr < -0.5f ? "apic" : (fr > 0.01f ? "flip" : "pic");
  char buf[384];
  std::snprintf(buf, sizeof(buf), "%s|cu%d|n%u|tiles%u|scenes%u|%s|models%x|svd%u|flags%u", prop.gcnArchName,
                prop.multiProcessorCount, pow2(h->n_live), pow2(tiles), h->n_scenes, mode, models,
                h->knobs.corrected_svd ? 1u : 0u, flags);
  key = buf;
  return NB_OK;
}

// One cache line's fields; false when the line is not such an object.
bool parse_tuning_line(const std::string& ln, std::string& key, nb_mpm_tuning& t) {
  auto str_field = [&](const char* name, std::string& out) {
    const std::string pat = std::string("\"") + name + "\"";
    size_t i = ln.find(pat);
    if (i == std::string::npos) return false;
    i = ln.find('"', ln.find(':', i + pat.size()) + 1);
    if (i == std::string::npos) return false;
    const size_t j = ln.find('"', i + 1);
    if (j == std::string::npos) return false;
    out = ln.substr(i + 1, j - i - 1);
    return true;
  };
  auto num_field = [&](const char* name, double& out) {
    const std::string pat = std::string("\"") + name + "\"";
    size_t i = ln.find(pat);
    if (i == std::string::npos) return false;
    i = ln.find(':', i + pat.size());
    if (i == std::string::npos) return false;
    char* end = nullptr;
    out = std::strtod(ln.c_str() + i + 1, &end);
    return end != ln.c_str() + i + 1;
  };
  double tb = 0, gb = 0, it = 0, ms = 0;
  if (!str_field("key", key) || !num_field("tile_blocks", tb) || !num_field("g2p_block", gb) ||
      !num_field("p2g_item", it))
    return false;
  if (!num_field("ms_per_substep", ms)) ms = 0.0;
  t.tile_blocks = uint32_t(tb);
  t.g2p_block = uint32_t(gb);
  t.p2g_item = uint32_t(it);
  t.ms_per_substep = float(ms);
  return true;
}
}  // namespace

extern "C" {

int nb_mpm_autotune_cached(nb_mpm* h, float dt_sub, uint32_t trial_substeps, uint32_t flags, const char* cache_path,
                           nb_mpm_tuning* chosen, int* hit) {
  if (!h || !cache_path) return fail(NB_ERR_ARG, "null argument");
  if (hit) *hit = 0;
  if (h->n_live == 0) return fail(NB_ERR_STATE, "no particles to time");
  int rc = sync_in(h);
  if (rc == NB_OK && !h->bins_valid) rc = launch_bin_only(h);  // the key's occupied-tile count
  std::string key;
  if (rc == NB_OK) rc = tuning_key(h, flags, key);
  if (rc != NB_OK) return rc;
  if (FILE* f = std::fopen(cache_path, "r")) {
    char line[1024];
    bool found = false;
    nb_mpm_tuning t{};
    while (std::fgets(line, sizeof(line), f)) {
      std::string k;
      nb_mpm_tuning c{};
      if (parse_tuning_line(line, k, c) && k == key) {
        t = c;
        found = true;  // the last entry of a key wins
      }
    }
    std::fclose(f);
    if (found) {
      rc = nb_mpm_set_tuning(h, &t);
      if (rc != NB_OK) return rc;
      h->tuning.ms_per_substep = t.ms_per_substep;
      if (chosen) *chosen = h->tuning;
      if (hit) *hit = 1;
      return NB_OK;
    }
  }
  nb_mpm_tuning t{};
  rc = nb_mpm_autotune(h, dt_sub, trial_substeps, flags, &t);
  if (rc != NB_OK) return rc;
  FILE* f = std::fopen(cache_path, "a");
  if (!f) return fail(NB_ERR_ARG, "autotune cache: cannot append to %s", cache_path);
  std::fprintf(f, "{\"key\": \"%s\", \"tile_blocks\": %u, \"g2p_block\": %u, \"p2g_item\": %u, "
                  "\"ms_per_substep\": %.5f}\n",
               key.c_str(), t.tile_blocks, t.g2p_block, t.p2g_item, double(t.ms_per_substep));
  std::fclose(f);
  if (chosen) *chosen = t;
  return NB_OK;
}

int nb_mpm_set_knobs(nb_mpm* h, const nb_mpm_knobs* k) {
  if (!h || !k) return fail(NB_ERR_ARG, "null argument");
  h->knobs = *k;
  return NB_OK;
}
int nb_mpm_get_knobs(const nb_mpm* h, nb_mpm_knobs* k) {
  if (!h || !k) return fail(NB_ERR_ARG, "null argument");
  *k = h->knobs;
  return NB_OK;
}
int nb_mpm_set_svd_fallback(nb_mpm* h, const nb_mpm_svd_fallback* f) {
  if (!h || !f) return fail(NB_ERR_ARG, "null argument");
  if (!(f->gap_rel >= 0.f && f->gap_rel < 1.f) || !(f->zero_rel >= 0.f && f->zero_rel < 1.f))
    return fail(NB_ERR_RANGE, "svd fallback thresholds gap_rel %g, zero_rel %g: each must lie in [0, 1)",
                double(f->gap_rel), double(f->zero_rel));
  h->svd_fb = *f;
  h->svd_fb.enabled = f->enabled ? 1u : 0u;
  h->svd_fb.count_stats = f->count_stats ? 1u : 0u;
  return NB_OK;
}
int nb_mpm_get_svd_fallback(const nb_mpm* h, nb_mpm_svd_fallback* f) {
  if (!h || !f) return fail(NB_ERR_ARG, "null argument");
  *f = h->svd_fb;
  return NB_OK;
}
int nb_mpm_svd_fallback_stats(nb_mpm* h, uint64_t* out, int reset) {
  if (!h || !out) return fail(NB_ERR_ARG, "null argument");
  int rc = download(h, out, h->d_svd_stats, NB_SVD_FB_STATS * 8);
  if (rc == NB_OK && reset) {
    NB_HIP(hipMemsetAsync(h->d_svd_stats, 0, NB_SVD_FB_STATS * 8, h->stream));
    NB_HIP(hipStreamSynchronize(h->stream));
  }
  return rc;
}
int nb_mpm_set_fused_g2p(nb_mpm* h, int on) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  h->fuse_g2p = on != 0;
  return NB_OK;
}
int nb_mpm_get_config(const nb_mpm* h, nb_mpm_config* c) {
  if (!h || !c) return fail(NB_ERR_ARG, "null argument");
  *c = h->cfg;
  return NB_OK;
}
int nb_mpm_grid_res(const nb_mpm* h, uint32_t res[3]) {
  if (!h || !res) return fail(NB_ERR_ARG, "null argument");
  for (int a = 0; a < 3; a++) res[a] = h->grid_res[a];
  return NB_OK;
}
int nb_mpm_total_cells(const nb_mpm* h, uint32_t* nc) {
  if (!h || !nc) return fail(NB_ERR_ARG, "null argument");
  *nc = h->NC;
  return NB_OK;
}
int nb_mpm_live_particle_count(const nb_mpm* h, uint32_t* n) {
  if (!h || !n) return fail(NB_ERR_ARG, "null argument");
  *n = h->n_live;
  return NB_OK;
}

int nb_mpm_set_materials(nb_mpm* h, const nb_material* mats, uint32_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (n == 0) return NB_OK;  // src/engine/MPMEngine.cpp:240
  if (!mats) return fail(NB_ERR_ARG, "null materials");
  if (n > h->max_materials()) return fail(NB_ERR_RANGE, "%u materials exceed the %u slots", n, h->max_materials());
  int rc = upload(h, h->d.materials, mats, sizeof(nb_material) * n);
  if (rc == NB_OK) {
    h->material_count = n;
    h->mats_host.assign(mats, mats + n);
  }
  return rc;
}

int nb_mpm_set_particle_material_ids(nb_mpm* h, const uint32_t* ids, uint32_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (n == 0) return NB_OK;
  if (!ids) return fail(NB_ERR_ARG, "null ids");
  if (n > h->cap) return fail(NB_ERR_RANGE, "%u ids exceed capacity %u", n, h->cap);
  std::vector<float> v(size_t(n) * 4, 0.f);
  for (uint32_t i = 0; i < n; i++) std::memcpy(&v[4 * size_t(i) + 3], &ids[i], 4);
  h->rec_stale = true;
  return upload(h, h->d.field[NB_FIELD_V], v.data(), v.size() * 4);
}

int nb_mpm_set_colliders(nb_mpm* h, const nb_collider* cols, uint32_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (n > NB_MAX_COLLIDERS) return fail(NB_ERR_RANGE, "%u colliders exceed the %u slots", n, NB_MAX_COLLIDERS);
  if (n && !cols) return fail(NB_ERR_ARG, "null colliders");
  int rc = upload(h, h->d.colliders, cols, sizeof(nb_collider) * n);
  if (rc == NB_OK) h->collider_count = n;
  return rc;
}
int nb_mpm_clear_analytic_colliders(nb_mpm* h) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  h->collider_count = 0;
  return NB_OK;
}

static int ensure_sdf(nb_mpm* h) {
  if (h->d.sdf) return NB_OK;
  return h->alloc(&h->d.sdf, h->NC);
}

int nb_mpm_set_collider_sdf(nb_mpm* h, const float* sdf, uint64_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  if (!sdf) return fail(NB_ERR_ARG, "null sdf");
  if (n != h->NC) return fail(NB_ERR_RANGE, "sdf has %llu values, the Morton cube has %u cells", (unsigned long long)n, h->NC);
  int rc = ensure_sdf(h);
  if (rc != NB_OK) return rc;
  rc = upload(h, h->d.sdf, sdf, size_t(n) * 4);
  if (rc == NB_OK) h->has_sdf = true;
  return rc;
}

int nb_mpm_set_collider_sphere(nb_mpm* h, float r, float cx, float cy, float cz) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  int rc = ensure_sdf(h);
  if (rc != NB_OK) return rc;
  launch_sdf_sphere(h->d.sdf, h->params(0.f, false), r, cx, cy, cz, h->stream);
  NB_HIP(hipGetLastError());
  NB_HIP(hipStreamSynchronize(h->stream));
  h->has_sdf = true;
  return NB_OK;
}

int nb_mpm_set_collider_mesh(nb_mpm* h, const float* tris, uint32_t n_tris, int exact) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  if (n_tris && !tris) return fail(NB_ERR_ARG, "null triangles");
  int rc = ensure_sdf(h);
  if (rc != NB_OK) return rc;
  float* d_tris = nullptr;
  const size_t bytes = std::max<size_t>(size_t(n_tris) * 12 * 4, 4);
  NB_HIP(hipMalloc(&d_tris, bytes));
  rc = n_tris ? upload(h, d_tris, tris, size_t(n_tris) * 12 * 4) : NB_OK;
  if (rc == NB_OK) {
    const int32_t res[3] = {int32_t(h->grid_res[0]), int32_t(h->grid_res[1]), int32_t(h->grid_res[2])};
    launch_mesh_sdf(d_tris, n_tris, res, h->cube_res, h->cfg.cell_size, exact != 0, h->d.sdf, h->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fail(NB_ERR_HIP, "mesh SDF build failed: %s", hipGetErrorString(e));
  }
  (void)hipFree(d_tris);
  if (rc == NB_OK) h->has_sdf = true;
  return rc;
}

int nb_mpm_get_collider_sdf(nb_mpm* h, float* out) {
  if (!h || !out) return fail(NB_ERR_ARG, "null argument");
  if (!h->has_sdf) return fail(NB_ERR_STATE, "no collider SDF is set");
  return download(h, out, h->d.sdf, size_t(h->NC) * 4);
}

int nb_mpm_clear_collider(nb_mpm* h) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  h->has_sdf = false;
  return NB_OK;
}

int nb_mpm_set_forces(nb_mpm* h, const nb_force* forces, uint32_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (n > NB_MAX_FORCES) return fail(NB_ERR_RANGE, "%u forces exceed the %u slots", n, NB_MAX_FORCES);
  if (n && !forces) return fail(NB_ERR_ARG, "null forces");
  uint32_t noise = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t t = forces[i].type;
    if (t > NB_FORCE_Z_CLAMP) return fail(NB_ERR_ARG, "unknown force type %u", t);
    if (t == NB_FORCE_TURBULENCE || t == NB_FORCE_NOISE) noise++;
  }
  int rc = upload(h, h->d.forces, forces, sizeof(nb_force) * n);
  if (rc == NB_OK) {
    h->force_count = n;
    h->noise_force_count = noise;
  }
  return rc;
}

int nb_mpm_append_particles(nb_mpm* h, const float* pos4, const float* vel4, uint32_t n, uint32_t* appended) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  if (appended) *appended = 0;
  const uint32_t nn = std::min(n, h->cap - h->n_live);
  if (nn == 0) return NB_OK;
  if (!pos4 || !vel4) return fail(NB_ERR_ARG, "null particle arrays");
  int rc = upload(h, h->d.field[NB_FIELD_P] + h->n_live, pos4, size_t(nn) * 16);
  if (rc == NB_OK) rc = upload(h, h->d.field[NB_FIELD_V] + h->n_live, vel4, size_t(nn) * 16);
  if (rc != NB_OK) return rc;
  launch_init_particles(h->d, h->n_live, nn, h->stream);
  NB_HIP(hipGetLastError());
  NB_HIP(hipStreamSynchronize(h->stream));
  h->n_live += nn;
  h->bins_valid = false;
  h->rec_stale = true;
  if (appended) *appended = nn;
  return NB_OK;
}

int nb_mpm_reset_particles(nb_mpm* h) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  NB_HIP(hipStreamSynchronize(h->stream));
  h->n_live = 0;
  h->bins_valid = false;
  std::fill(h->emit_done.begin(), h->emit_done.end(), 0);
  h->emit_rng.seed(12345u);
  return NB_OK;
}

int nb_mpm_add_emitter(nb_mpm* h, const nb_emitter* e) {
  if (!h || !e) return fail(NB_ERR_ARG, "null argument");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  if (e->shape > NB_EMITTER_ELLIPSE) return fail(NB_ERR_ARG, "unknown emitter shape %u", e->shape);
  if (e->shape == NB_EMITTER_ELLIPSE && !(e->size[0] > 0.f && e->size[1] > 0.f))
    return fail(NB_ERR_ARG, "an ellipse needs positive semi-axes");
  h->emitters.push_back(*e);
  h->emit_done.push_back(0);
  return NB_OK;
}

int nb_mpm_clear_emitters(nb_mpm* h) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  h->emitters.clear();
  h->emit_done.clear();
  return NB_OK;
}

int nb_mpm_get_emitter(const nb_mpm* h, uint32_t i, nb_emitter* e, uint32_t* count) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (count) *count = uint32_t(h->emitters.size());
  if (!e) return NB_OK;
  if (i >= h->emitters.size()) return fail(NB_ERR_RANGE, "emitter %u of %zu", i, h->emitters.size());
  *e = h->emitters[i];
  return NB_OK;
}

int nb_mpm_emit(nb_mpm* h, float dt) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  const float Vp = (h->cfg.domain_size[0] / float(h->cfg.nx)) * (h->cfg.domain_size[1] / float(h->cfg.ny)) *
                   (h->cfg.domain_size[2] / float(h->cfg.nz));  // MPMConfig::particleVolume
  for (size_t si = 0; si < h->emitters.size(); si++) {
    nb_emitter& e = h->emitters[si];
    int& done = h->emit_done[si];
    const bool go = e.step_count == -1 ? done == 0 : (e.step_count == 0 ? true : done < e.step_count);
    if (!go) continue;
    const int64_t avail = int64_t(h->cap) - int64_t(h->n_live);
    const int64_t n = std::min<int64_t>(e.particles_per_step, avail);
    if (n <= 0) {
      done++;
      continue;
    }
    std::vector<float> pos(size_t(n) * 4), vel(size_t(n) * 4);
    float id_bits;
    std::memcpy(&id_bits, &e.particle_type, 4);
    for (int64_t j = 0; j < n; j++) {  // all positions, then all velocities (MPMEngine.cpp:324-336)
      float x, y, z;
      nebula::emitter_sample(e, h->emit_rng, x, y, z);
      pos[4 * j + 0] = x; pos[4 * j + 1] = y; pos[4 * j + 2] = z; pos[4 * j + 3] = Vp;
    }
    for (int64_t j = 0; j < n; j++) {
      float vx, vy, vz;
      nebula::emitter_velocity(e, pos[4 * j + 0], pos[4 * j + 1], pos[4 * j + 2], h->emit_rng, vx, vy, vz);
      vel[4 * j + 0] = vx; vel[4 * j + 1] = vy; vel[4 * j + 2] = vz; vel[4 * j + 3] = id_bits;
    }
    uint32_t got = 0;
    const int rc = nb_mpm_append_particles(h, pos.data(), vel.data(), uint32_t(n), &got);
    if (rc != NB_OK) return rc;
    done++;
    for (int a = 0; a < 3; a++) e.center[a] += e.center_vel[a] * dt;
  }
  return NB_OK;
}

int nb_mpm_get_field(nb_mpm* h, int field, float* out4, uint32_t first, uint32_t count) {
  if (!h || !out4) return fail(NB_ERR_ARG, "null argument");
  if (field < 0 || field >= NB_NUM_FIELDS) return fail(NB_ERR_ARG, "bad field %d", field);
  if (uint64_t(first) + count > h->cap) return fail(NB_ERR_RANGE, "rows beyond capacity");
  return download(h, out4, h->d.field[field] + first, size_t(count) * 16);
}

int nb_mpm_set_field(nb_mpm* h, int field, const float* in4, uint32_t first, uint32_t count) {
  if (!h || !in4) return fail(NB_ERR_ARG, "null argument");
  if (field < 0 || field >= NB_NUM_FIELDS) return fail(NB_ERR_ARG, "bad field %d", field);
  if (uint64_t(first) + count > h->cap) return fail(NB_ERR_RANGE, "rows beyond capacity");
  h->bins_valid = false;
  h->rec_stale = true;
  return upload(h, h->d.field[field] + first, in4, size_t(count) * 16);
}

int nb_mpm_field_device_ptr(nb_mpm* h, int field, void** dptr) {
  if (!h || !dptr) return fail(NB_ERR_ARG, "null argument");
  if (field < 0 || field >= NB_NUM_FIELDS) return fail(NB_ERR_ARG, "bad field %d", field);
  *dptr = h->d.field[field];
  h->ptr_exported = true;
  return NB_OK;
}

int nb_mpm_set_live_particle_count(nb_mpm* h, uint32_t n) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch() && n != h->cap) return fail(NB_ERR_UNSUPPORTED, "a batch's particle count is fixed by its scenes");
  if (n > h->cap) return fail(NB_ERR_RANGE, "%u exceeds capacity %u", n, h->cap);
  h->n_live = n;
  h->bins_valid = false;
  h->rec_stale = true;
  return NB_OK;
}

// ---- pass-level access ----
int nb_mpm_run_bin(nb_mpm* h) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  SubParams p = h->params(0.f, false);
  int rc = sync_in(h);
  if (rc != NB_OK) return rc;
  if (h->n_live) launch_bin(h->dev(), p, h->NT, false, h->stream, nullptr);
  NB_HIP(hipGetLastError());
  NB_HIP(hipStreamSynchronize(h->stream));
  h->bins_valid = true;
  return NB_OK;
}

int nb_mpm_get_bins(nb_mpm* h, uint32_t* keys, uint32_t* sorted, uint32_t* cell_count, uint32_t* cell_end) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  if (!h->bins_valid) return fail(NB_ERR_STATE, "no bins: call nb_mpm_run_bin first");
  const uint32_t n = h->n_live;
  int rc = NB_OK;
  if (keys && n) rc = download(h, keys, h->d.keys, size_t(n) * 4);
  if (rc == NB_OK && sorted && n) rc = download(h, sorted, h->d.sorted, size_t(n) * 4);
  if (rc != NB_OK) return rc;
  if (cell_count || cell_end) {
    std::vector<uint32_t> cnt(h->NC), dev_end(h->NC);
    if (n) {
      if ((rc = download(h, cnt.data(), h->d.cell_count, size_t(h->NC) * 4)) != NB_OK) return rc;
      if ((rc = download(h, dev_end.data(), h->d.cell_end, size_t(h->NC) * 4)) != NB_OK) return rc;
    }
    if (cell_count) std::memcpy(cell_count, cnt.data(), size_t(h->NC) * 4);
    if (cell_end) {
      // dense END offsets: occupied cells carry the device value; an empty cell
      // holds the running prefix (the reference's cellOffset for count 0)
      uint32_t acc = 0;
      for (uint32_t c = 0; c < h->NC; c++) {
        acc += cnt[c];
        cell_end[c] = cnt[c] ? dev_end[c] : acc;
      }
    }
  }
  return NB_OK;
}

int nb_mpm_run_p2g(nb_mpm* h, float dt_sub) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  if (!h->bins_valid) return fail(NB_ERR_STATE, "no bins: call nb_mpm_run_bin first");
  SubParams p = h->params(dt_sub, false);
  int rc = sync_in(h);
  if (rc != NB_OK) return rc;
  if (h->n_live) {
    launch_p2g(h->dev(), p, h->NT, h->stream, nullptr);
    launch_grid(h->dev(), p, h->NT, true, false, h->d.node_tiles, h->d.counters + CNT_NODE_TILES, h->stream, nullptr);
  }
  NB_HIP(hipGetLastError());
  NB_HIP(hipStreamSynchronize(h->stream));
  h->grid_mode = h->n_live ? GridMode::SPARSE : GridMode::NONE;
  h->old_valid = false;
  return NB_OK;
}

int nb_mpm_run_grid_update(nb_mpm* h, float dt_sub) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  if (h->grid_mode == GridMode::NONE) return NB_OK;
  SubParams p = h->params(dt_sub, true);
  if (h->grid_mode == GridMode::DENSE)
    launch_grid(h->dev(), p, h->NT, false, true, h->d.all_tiles, h->d.counters + CNT_ALL_TILES, h->stream, nullptr);
  else
    launch_grid(h->dev(), p, h->NT, false, true, h->d.node_tiles, h->d.counters + CNT_NODE_TILES, h->stream, nullptr);
  NB_HIP(hipGetLastError());
  NB_HIP(hipStreamSynchronize(h->stream));
  h->old_valid = true;
  return NB_OK;
}

int nb_mpm_run_g2p(nb_mpm* h, float dt_sub) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  SubParams p = h->params(dt_sub, false);
  int rc = sync_in(h);
  if (rc != NB_OK) return rc;
  if (h->n_live) launch_g2p(h->dev(), p, h->bins_valid, h->stream, nullptr);
  NB_HIP(hipGetLastError());
  if ((rc = sync_out(h)) != NB_OK) return rc;
  NB_HIP(hipStreamSynchronize(h->stream));
  h->bins_valid = false;  // positions moved
  h->keys_valid = h->n_live != 0;
  return NB_OK;
}

int nb_mpm_get_grid(nb_mpm* h, float* mom4, float* mass) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  const size_t NC = h->NC;
  if (mom4) std::memset(mom4, 0, NC * 2 * 16);
  if (mass) std::memset(mass, 0, NC * 4);
  if (h->grid_mode == GridMode::NONE) return NB_OK;
  const size_t cells = size_t(h->NT) * kTileCells;
  std::vector<float> node(cells * 4), old(cells * 4);
  int rc = download(h, node.data(), h->d.grid_node, cells * 16);
  if (rc == NB_OK && h->old_valid) rc = download(h, old.data(), h->d.grid_old, cells * 16);
  if (rc != NB_OK) return rc;
  std::vector<uint32_t> tiles;
  std::vector<int32_t> tile_act;
  if (h->grid_mode == GridMode::DENSE) {
    tiles.resize(h->NT);
    for (uint32_t t = 0; t < h->NT; t++) tiles[t] = t;
  } else {
    uint32_t cnt[CNT_SLOTS];
    if ((rc = download(h, cnt, h->d.counters, sizeof(cnt))) != NB_OK) return rc;
    tiles.resize(cnt[CNT_NODE_TILES]);
    if (!tiles.empty() && (rc = download(h, tiles.data(), h->d.node_tiles, tiles.size() * 4)) != NB_OK) return rc;
    tile_act.resize(h->NT);
    if ((rc = download(h, tile_act.data(), h->d.tile_act, size_t(h->NT) * 4)) != NB_OK) return rc;
  }
  // sparse grids hold only the nodes inside some occupied tile's 10^3 box
  const int td = int(h->tile_dim);
  auto covered = [&](uint32_t T, uint32_t j) {
    if (tile_act.empty()) return true;
    const int t[3] = {int(morton_compact(T)), int(morton_compact(T >> 1)), int(morton_compact(T >> 2))};
    const int l[3] = {int(morton_compact(j)), int(morton_compact(j >> 1)), int(morton_compact(j >> 2))};
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int dd[3] = {dx, dy, dz};
          bool ok = true;
          int s3[3];
          for (int a = 0; a < 3; a++) {
            if ((dd[a] < 0 && l[a] != 0) || (dd[a] > 0 && l[a] != 7)) ok = false;
            s3[a] = t[a] + dd[a];
            if (s3[a] < 0 || s3[a] >= td) ok = false;
          }
          if (ok && tile_act[morton3(uint32_t(s3[0]), uint32_t(s3[1]), uint32_t(s3[2]))] >= 0) return true;
        }
    return false;
  };
  for (uint32_t T : tiles) {
    for (uint32_t j = 0; j < uint32_t(kTileCells); j++) {
      const size_t nid = (size_t(T) << kTileBits) | j;
      if (nid >= NC || !covered(T, j)) continue;
      const float* g = &node[nid * 4];
      if (mom4) {
        mom4[nid * 4 + 0] = g[0];
        mom4[nid * 4 + 1] = g[1];
        mom4[nid * 4 + 2] = g[2];
        if (h->old_valid && g[3] >= 1e-10f) {
          mom4[(NC + nid) * 4 + 0] = old[nid * 4 + 0];
          mom4[(NC + nid) * 4 + 1] = old[nid * 4 + 1];
          mom4[(NC + nid) * 4 + 2] = old[nid * 4 + 2];
        }
      }
      if (mass) mass[nid] = g[3];
    }
  }
  return NB_OK;
}

int nb_mpm_set_grid(nb_mpm* h, const float* mom4, const float* mass, const float* old4) {
  if (!h || !mom4 || !mass) return fail(NB_ERR_ARG, "null argument");
  if (h->batch()) return fail(NB_ERR_UNSUPPORTED, "%s: not available on a batch of scenes", __func__);
  const size_t cells = size_t(h->NT) * kTileCells;
  std::vector<float> node(cells * 4, 0.f);
  for (size_t c = 0; c < h->NC; c++) {
    node[4 * c + 0] = mom4[4 * c + 0];
    node[4 * c + 1] = mom4[4 * c + 1];
    node[4 * c + 2] = mom4[4 * c + 2];
    node[4 * c + 3] = mass[c];
  }
  int rc = upload(h, h->d.grid_node, node.data(), cells * 16);
  if (rc != NB_OK) return rc;
  if (old4) {
    std::fill(node.begin(), node.end(), 0.f);
    for (size_t c = 0; c < h->NC; c++)
      for (int k = 0; k < 3; k++) node[4 * c + k] = old4[4 * c + k];
    if ((rc = upload(h, h->d.grid_old, node.data(), cells * 16)) != NB_OK) return rc;
  } else {
    NB_HIP(hipMemsetAsync(h->d.grid_old, 0, cells * 16, h->stream));
    NB_HIP(hipStreamSynchronize(h->stream));
  }
  h->grid_mode = GridMode::DENSE;
  h->old_valid = true;
  return NB_OK;
}

int nb_mpm_profile_substeps(nb_mpm* h, float dt_sub, uint32_t n, float* kernel_ms, uint32_t kernel_cap,
                            uint32_t* n_kernels, char* names, uint32_t names_cap) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  Prof pf{};
  for (int i = 0; i < 2 * kNumProfiledKernels; i++) NB_HIP(hipEventCreate(&pf.ev[i]));
  int rc = run_substeps(h, dt_sub, n, &pf);
  (void)hipStreamSynchronize(h->stream);
  for (int i = 0; i < 2 * kNumProfiledKernels; i++) (void)hipEventDestroy(pf.ev[i]);
  if (rc != NB_OK) return rc;
  for (int i = 0; i < kNumProfiledKernels; i++) h->prof_calls[i] = pf.calls[i];
  if (n_kernels) *n_kernels = kNumProfiledKernels;
  if (kernel_ms)
    for (uint32_t i = 0; i < kernel_cap && i < uint32_t(kNumProfiledKernels); i++) kernel_ms[i] = float(pf.ms[i]);
  if (names && names_cap) {
    std::string s;
    for (int i = 0; i < kNumProfiledKernels; i++) {
      if (i) s += ",";
      s += kKernelNames[i];
    }
    std::snprintf(names, names_cap, "%s", s.c_str());
  }
  return NB_OK;
}

int nb_mpm_profile_calls(const nb_mpm* h, uint32_t* calls, uint32_t cap) {
  if (!h || !calls) return fail(NB_ERR_ARG, "null argument");
  for (uint32_t i = 0; i < cap && i < uint32_t(kNumProfiledKernels); i++) calls[i] = h->prof_calls[i];
  return NB_OK;
}

int nb_mpm_calibration_copy(nb_mpm* h, uint64_t bytes, uint32_t repeats) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  const uint64_t n4 = bytes / 16;
  if (!n4) return NB_OK;
  void *a = nullptr, *b = nullptr;
  NB_HIP(hipMalloc(&a, n4 * 16));
  if (hipMalloc(&b, n4 * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(NB_ERR_HIP, "hipMalloc failed");
  }
  hipError_t e = hipMemsetAsync(a, 0, n4 * 16, h->stream);
  for (uint32_t r = 0; r < repeats && e == hipSuccess; r++) {
    launch_copy(reinterpret_cast<float4*>(b), reinterpret_cast<const float4*>(a), n4, h->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(NB_ERR_HIP, "calibration copy failed: %s", hipGetErrorString(e));
  return NB_OK;
}

int nb_mpm_active_node_count(nb_mpm* h, uint64_t* g_act) {
  if (!h || !g_act) return fail(NB_ERR_ARG, "null argument");
  *g_act = 0;
  if (h->grid_mode == GridMode::NONE) return NB_OK;
  const size_t NC = h->NC;
  std::vector<float> mass(NC);
  int rc = nb_mpm_get_grid(h, nullptr, mass.data());
  if (rc != NB_OK) return rc;
  uint64_t c = 0;
  for (size_t i = 0; i < NC; i++) c += mass[i] >= 1e-10f;
  *g_act = c;
  return NB_OK;
}

int nb_mpm_active_tile_count(nb_mpm* h, uint32_t* tiles) {
  if (!h || !tiles) return fail(NB_ERR_ARG, "null argument");
  uint32_t cnt[CNT_SLOTS];
  int rc = download(h, cnt, h->d.counters, sizeof(cnt));
  if (rc != NB_OK) return rc;
  *tiles = cnt[CNT_ACTIVE];
  return NB_OK;
}

}  // extern "C"

// ---- reverse mode ----
namespace {
FieldPtrs adj_fields(nb_mpm* h) {
  FieldPtrs f;
  for (int i = 0; i < NB_NUM_FIELDS; i++) f.f[i] = h->adj.field[i];
  return f;
}
template <class T>
int adj_alloc(nb_mpm* h, T** p, size_t count) {
  void* q = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(NB_ERR_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  h->adj.allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return NB_OK;
}
// one set of bin buffers, in the state a fresh engine's bins start in
int adj_alloc_bins(nb_mpm* h, DevState& b) {
  const size_t cells = size_t(h->NT) * kTileCells;
  int rc = NB_OK;
  if (rc == NB_OK) rc = adj_alloc(h, &b.rank, h->cap);
  if (rc == NB_OK) rc = adj_alloc(h, &b.sorted, h->cap);
  if (rc == NB_OK) rc = adj_alloc(h, &b.cell_count, cells);
  if (rc == NB_OK) rc = adj_alloc(h, &b.cell_end, cells);
  if (rc == NB_OK) rc = adj_alloc(h, &b.tile_flag, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.tile_mark, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.blk_cnt, (h->NT + 255) / 256);
  if (rc == NB_OK) rc = adj_alloc(h, &b.tile_act, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.act_tiles, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.node_tiles, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.tile_total, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.tile_start, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.tile_order, h->NT);
  if (rc == NB_OK) rc = adj_alloc(h, &b.counters, CNT_SLOTS);
  if (rc == NB_OK &&
      (hipMemsetAsync(b.cell_count, 0, cells * 4, h->stream) != hipSuccess ||
       hipMemsetAsync(b.cell_end, 0, cells * 4, h->stream) != hipSuccess ||
       hipMemsetAsync(b.tile_flag, 0, h->NT, h->stream) != hipSuccess ||
       hipMemsetAsync(b.counters, 0, CNT_SLOTS * 4, h->stream) != hipSuccess ||
       hipMemsetAsync(b.tile_act, 0xFF, size_t(h->NT) * 4, h->stream) != hipSuccess))
    rc = fail(NB_ERR_HIP, "hipMemsetAsync failed");
  if (rc == NB_OK) launch_fill_identity(b.tile_order, h->NT, h->stream);
  return rc;
}
void adj_free(nb_mpm* h) {
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void* p : h->adj.allocs) (void)hipFree(p);
  h->adj = nb_mpm::Adjoint{};
}
int copy_rec(nb_mpm* h, float4* dst, const float4* src) {
  if (!h->n_live) return NB_OK;
  NB_HIP(hipMemcpyAsync(dst, src, size_t(h->n_live) * kRec * sizeof(float4), hipMemcpyDeviceToDevice, h->stream));
  return NB_OK;
}
// the tape that holds segment sg's inner states (see nb_mpm::Adjoint)
uint32_t tape_of(const nb_mpm* h, uint32_t sg) { return sg % h->adj.n_tapes; }
// the records before substep s
float4* adj_slot(nb_mpm* h, uint32_t s) {
  const auto& A = h->adj;
  const size_t rec = size_t(h->cap) * kRec;
  const uint32_t sg = s / A.seg, j = s % A.seg;
  return j == 0 ? A.ckpt + sg * rec : A.tape + (size_t(tape_of(h, sg)) * (A.seg - 1) + (j - 1)) * rec;
}
// where substep s writes its result: the next slot, or the engine's records after the last substep
float4* adj_next(nb_mpm* h, uint32_t s) { return s + 1 == h->adj.n ? h->d.rec : adj_slot(h, s + 1); }
// the bin keys of the records in adj_slot(h, s) / adj_next(h, s)
uint32_t* key_slot(nb_mpm* h, uint32_t s) {
  const auto& A = h->adj;
  const uint32_t sg = s / A.seg, j = s % A.seg;
  return j == 0 ? A.ckpt_keys + size_t(sg) * h->cap
                : A.tape_keys + (size_t(tape_of(h, sg)) * (A.seg - 1) + (j - 1)) * h->cap;
}
uint32_t* key_next(nb_mpm* h, uint32_t s) { return s + 1 == h->adj.n ? h->d.keys : key_slot(h, s + 1); }
// the slot of substep s in its tape (grids and bins)
size_t tape_slot(const nb_mpm* h, uint32_t s) {
  return size_t(tape_of(h, s / h->adj.seg)) * h->adj.seg + s % h->adj.seg;
}
// the grids kept for substep s
float4* grid_slot(nb_mpm* h, uint32_t s, int which) {
  return h->adj.grid_tape + (tape_slot(h, s) * 2 + which) * h->NT * kTileCells;
}
// substep s of the recorded run: slot to slot, keeping its grids
TapeIO tape_io(nb_mpm* h, uint32_t s, float4* rec_out, uint32_t* keys_out) {
  TapeIO io;
  io.rec_in = adj_slot(h, s);
  io.rec_out = rec_out;
  io.keys_in = key_slot(h, s);
  io.keys_out = keys_out;
  io.grid_node = grid_slot(h, s, 0);
  io.grid_pre = grid_slot(h, s, 1);
  io.bins = &h->adj.slot_bins[tape_slot(h, s)];
  return io;
}
}  // namespace

extern "C" {

int nb_mpm_adjoint_begin(nb_mpm* h, uint32_t n_substeps, uint32_t segment) {
  return nb_mpm_adjoint_begin_tapes(h, n_substeps, segment, 1);
}

int nb_mpm_adjoint_begin_tapes(nb_mpm* h, uint32_t n_substeps, uint32_t segment, uint32_t tapes) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  if (n_substeps == 0 || segment == 0) return fail(NB_ERR_ARG, "n_substeps and segment must be > 0");
  if (!h->knobs.corrected_svd)
    return fail(NB_ERR_STATE, "the adjoint is defined for corrected_svd = 1 (proper polar SVD)");
  if (h->knobs.flip_ratio > 0.01f) return fail(NB_ERR_UNSUPPORTED, "the adjoint supports PIC and APIC, not FLIP");
  if (h->collider_count) return fail(NB_ERR_UNSUPPORTED, "the adjoint does not support analytic colliders");
  for (uint32_t i = 0; i < h->material_count; i++) {
    const uint32_t m = h->mats_host[i].model;
    if (m == NB_MAT_FLUID || m == NB_MAT_VISCOPLASTIC_MUD)
      return fail(NB_ERR_UNSUPPORTED, "material %u: the adjoint supports elastic, Von Mises and Drucker-Prager", i);
  }
  adj_free(h);
  auto& A = h->adj;
  A.n = n_substeps;
  A.seg = std::min(segment, n_substeps);
  A.n_ckpt = (n_substeps + A.seg - 1) / A.seg;
  const size_t rec = size_t(h->cap) * kRec;
  if (tapes == 0) {
    // as many segments' tapes as fit in the free device memory, 4 GiB and a
    // tenth of it left over
    const size_t cells = size_t(h->NT) * kTileCells;
    const size_t per_tape = size_t(A.seg - 1) * (rec * sizeof(float4) + size_t(h->cap) * 4) +
                            size_t(A.seg) * (2 * cells * sizeof(float4) + 2 * cells * 4 + size_t(h->cap) * 8 +
                                             size_t(h->NT) * 40);
    size_t free_b = 0, total_b = 0;
    NB_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t fixed = rec * sizeof(float4) * (A.n_ckpt + 2) + size_t(h->cap) * 4 * A.n_ckpt + (size_t(4) << 30);
    const size_t avail = free_b > fixed ? (free_b - fixed) / 10 * 9 : 0;
    tapes = uint32_t(std::min<size_t>(A.n_ckpt, std::max<size_t>(1, avail / std::max<size_t>(per_tape, 1))));
  }
  A.n_tapes = std::min(tapes, A.n_ckpt);
  const uint32_t n_blocks = (h->cap + 255u) / 256u;  // blocks of the G2P backward
  int rc = NB_OK;
  if (rc == NB_OK) rc = adj_alloc(h, &A.ckpt, rec * A.n_ckpt);
  if (rc == NB_OK) rc = adj_alloc(h, &A.tape, rec * (A.seg - 1) * A.n_tapes);
  float4* kc = nullptr;
  float4* kt = nullptr;
  if (rc == NB_OK) rc = adj_alloc(h, &kc, (size_t(h->cap) * A.n_ckpt + 3) / 4);
  if (rc == NB_OK) rc = adj_alloc(h, &kt, (size_t(h->cap) * (A.seg - 1) * A.n_tapes + 3) / 4);
  if (rc == NB_OK) rc = adj_alloc(h, &A.grid_tape, size_t(A.seg) * A.n_tapes * 2 * h->NT * kTileCells);
  A.slot_bins.assign(size_t(A.seg) * A.n_tapes, DevState{});
  for (size_t j = 0; j < A.slot_bins.size() && rc == NB_OK; j++) rc = adj_alloc_bins(h, A.slot_bins[j]);
  A.tape_seg.assign(A.n_tapes, -1);
  if (rc == NB_OK) rc = adj_alloc(h, &A.adj_a, rec);
  if (rc == NB_OK) rc = adj_alloc(h, &A.adj_b, rec);
  for (int f = 0; f < NB_NUM_FIELDS && rc == NB_OK; f++) rc = adj_alloc(h, &A.field[f], h->cap);
  if (rc == NB_OK) rc = adj_alloc(h, &A.a.coef, size_t(h->cap) * 6);
  if (rc == NB_OK) rc = adj_alloc(h, &A.a.partial_adj, size_t(h->cap_tiles) * kBoxNodes);
  if (rc == NB_OK) rc = adj_alloc(h, &A.a.grid_adj, size_t(h->NT) * kTileCells);
  float4* pg = nullptr;
  A.a.pg_materials = std::max(h->material_count, 1u);
  if (rc == NB_OK) rc = adj_alloc(h, &A.a.pg_acc, size_t(n_blocks) * A.a.pg_materials);
  if (rc == NB_OK) rc = adj_alloc(h, &A.a.pg_chunk, size_t(h->max_materials()) * ((n_blocks + 63) / 64));
  if (rc == NB_OK) rc = adj_alloc(h, &pg, h->max_materials());
  if (rc != NB_OK) {
    std::string keep = g_err;
    adj_free(h);
    g_err = keep;
    return rc;
  }
  A.ckpt_keys = reinterpret_cast<uint32_t*>(kc);
  A.tape_keys = reinterpret_cast<uint32_t*>(kt);
  A.a.pgrad = reinterpret_cast<float*>(pg);
  for (int f = 0; f < NB_NUM_FIELDS; f++) NB_HIP(hipMemsetAsync(A.field[f], 0, size_t(h->cap) * 16, h->stream));
  NB_HIP(hipStreamSynchronize(h->stream));
  A.on = true;
  return NB_OK;
}

int nb_mpm_adjoint_tapes(const nb_mpm* h, uint32_t* tapes) {
  if (!h || !tapes) return fail(NB_ERR_ARG, "null argument");
  if (!h->adj.on) return fail(NB_ERR_STATE, "call nb_mpm_adjoint_begin first");
  *tapes = h->adj.n_tapes;
  return NB_OK;
}

int nb_mpm_adjoint_end(nb_mpm* h) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  adj_free(h);
  return NB_OK;
}

int nb_mpm_forward_recorded(nb_mpm* h, float dt_sub) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  auto& A = h->adj;
  if (!A.on) return fail(NB_ERR_STATE, "call nb_mpm_adjoint_begin first");
  int rc = sync_in(h);
  if (rc == NB_OK) rc = copy_rec(h, A.ckpt, h->d.rec);
  h->keys_valid = false;  // the first substep writes its keys into the first key slot
  for (uint32_t s = 0; s < A.n && rc == NB_OK; s++)
    rc = enqueue_substep(h, dt_sub, nullptr, tape_io(h, s, adj_next(h, s), key_next(h, s)));
  h->grid_mode = GridMode::NONE;  // the grids went to the tape
  h->bins_valid = false;          // and the bins to the tape's bin buffers
  if (rc == NB_OK) rc = sync_out(h);
  if (rc != NB_OK) return rc;
  NB_HIP(hipStreamSynchronize(h->stream));
  A.dt = dt_sub;
  A.recorded = true;
  // each tape holds the last segment the run wrote into it
  for (uint32_t sg = 0; sg < A.n_ckpt; sg++) A.tape_seg[tape_of(h, sg)] = int(sg);
  return NB_OK;
}

int nb_mpm_set_adjoint(nb_mpm* h, int field, const float* in4, uint32_t first, uint32_t count) {
  if (!h || !in4) return fail(NB_ERR_ARG, "null argument");
  if (!h->adj.on) return fail(NB_ERR_STATE, "call nb_mpm_adjoint_begin first");
  if (field < 0 || field >= NB_NUM_FIELDS) return fail(NB_ERR_ARG, "bad field %d", field);
  if (uint64_t(first) + count > h->cap) return fail(NB_ERR_RANGE, "rows beyond capacity");
  return upload(h, h->adj.field[field] + first, in4, size_t(count) * 16);
}

int nb_mpm_get_adjoint(nb_mpm* h, int field, float* out4, uint32_t first, uint32_t count) {
  if (!h || !out4) return fail(NB_ERR_ARG, "null argument");
  if (!h->adj.on) return fail(NB_ERR_STATE, "call nb_mpm_adjoint_begin first");
  if (field < 0 || field >= NB_NUM_FIELDS) return fail(NB_ERR_ARG, "bad field %d", field);
  if (uint64_t(first) + count > h->cap) return fail(NB_ERR_RANGE, "rows beyond capacity");
  return download(h, out4, h->adj.field[field] + first, size_t(count) * 16);
}

int nb_mpm_adjoint_device_ptr(nb_mpm* h, int field, void** dptr) {
  if (!h || !dptr) return fail(NB_ERR_ARG, "null argument");
  if (!h->adj.on) return fail(NB_ERR_STATE, "call nb_mpm_adjoint_begin first");
  if (field < 0 || field >= NB_NUM_FIELDS) return fail(NB_ERR_ARG, "bad field %d", field);
  *dptr = h->adj.field[field];
  return NB_OK;
}

int nb_mpm_backward(nb_mpm* h, float* d_params, uint32_t params_cap) {
  if (!h) return fail(NB_ERR_ARG, "null handle");
  auto& A = h->adj;
  if (!A.on || !A.recorded) return fail(NB_ERR_STATE, "call nb_mpm_forward_recorded first");
  if (h->material_count > A.a.pg_materials)
    return fail(NB_ERR_STATE, "materials were added after nb_mpm_adjoint_begin");
  NB_HIP(hipMemsetAsync(A.a.pgrad, 0, size_t(h->max_materials()) * 4 * sizeof(float), h->stream));
  NB_HIP(hipMemsetAsync(A.a.pg_acc, 0, size_t((h->n_live + 255) / 256) * h->material_count * sizeof(float4),
                        h->stream));
  launch_soa_to_aos_ptrs(adj_fields(h), A.adj_a, h->n_live, h->stream);
  float4* in = A.adj_a;
  float4* out = A.adj_b;
  int rc = NB_OK;
  for (int sg = int(A.n_ckpt) - 1; sg >= 0 && rc == NB_OK; sg--) {
    const uint32_t s0 = uint32_t(sg) * A.seg;
    const uint32_t len = std::min(A.seg, A.n - s0);
    if (A.tape_seg[tape_of(h, uint32_t(sg))] != sg) {
      // recompute the segment's records and grids from its checkpoint (its
      // last substep only up to the grid update: its result is the next checkpoint)
      h->keys_valid = true;
      for (uint32_t j = 0; j < len && rc == NB_OK; j++) {
        const bool last = j + 1 == len;
        TapeIO io = tape_io(h, s0 + j, last ? nullptr : adj_slot(h, s0 + j + 1), last ? nullptr : key_slot(h, s0 + j + 1));
        io.g2p = !last;
        rc = enqueue_substep(h, A.dt, nullptr, io);
      }
      A.tape_seg[tape_of(h, uint32_t(sg))] = sg;
    }
    for (int j = int(len) - 1; j >= 0 && rc == NB_OK; j--) {
      const uint32_t s = s0 + uint32_t(j);
      SubParams p = h->params(A.dt, false);
      DevState dv = h->dev();
      use_bins(dv, A.slot_bins[tape_slot(h, s)]);  // as the forward substep binned them
      dv.rec = adj_slot(h, s);
      dv.keys = key_slot(h, s);
      dv.grid_node = grid_slot(h, s, 0);
      dv.grid_pre = grid_slot(h, s, 1);
      launch_adjoint_substep(dv, A.a, p, h->NT, in, out, h->stream);
      NB_HIP(hipGetLastError());
      std::swap(in, out);
    }
  }
  if (rc != NB_OK) return rc;
  launch_adjoint_params(A.a, h->n_live, h->material_count, h->stream);
  NB_HIP(hipGetLastError());
  launch_aos_to_soa_ptrs(in, adj_fields(h), h->n_live, h->stream);
  // the engine's records still hold the state the recorded run ended in; the
  // bins, keys and grid are those of an earlier substep
  h->keys_valid = false;
  h->bins_valid = false;
  h->grid_mode = GridMode::NONE;
  if (d_params) {
    std::vector<float> pg(size_t(h->material_count) * 4);
    if ((rc = download(h, pg.data(), A.a.pgrad, pg.size() * sizeof(float))) != NB_OK) return rc;
    const uint32_t n = std::min(params_cap, h->material_count * 4u);
    std::memcpy(d_params, pg.data(), n * sizeof(float));
  }
  NB_HIP(hipStreamSynchronize(h->stream));
  return NB_OK;
}

}  // extern "C"
