/*
 * nebula_mpm.h — C ABI of the MI355X-native MPM substep.
 *
 * This is the drop-in boundary for the reference's `class MPMEngine`
 * (src/engine/MPMEngine.h:59-186, src/engine/EngineBase.h:23-66).  Every entry
 * point below names the reference member it replaces.  The ABI is plain C:
 * opaque handle, POD structs, host or device pointers plus sizes, an `int`
 * status (0 = ok, negative = error) and a thread-local error string.  No C++
 * exception crosses it and no torch type appears in it.
 *
 * Differences from the Vulkan surface, by design:
 *   - `step(VkCommandBuffer, dt)` becomes `nb_mpm_step(h, dt)`: work is enqueued
 *     on the engine's HIP stream (its own, or the one handed to create); the
 *     caller synchronises with `nb_mpm_sync` or its own stream wait.
 *   - the runtime shaderc rebuild of the grid-update shader on every force change
 *     (src/core/ForceShaderCompiler.cpp:139-166) becomes a compiled device-side
 *     switch over the 32-byte force records; `nb_mpm_set_forces` only uploads.
 *   - setters that the reference leaves unchecked (setColliders writes past its
 *     64-slot buffer, src/engine/MPMEngine.cpp:260-265) return NB_ERR_RANGE.
 *
 * Particle state uses the reference's layout (src/engine/MPMEngine.cpp:46-53,
 * shaders/mpm_common.glsl:98-122): eight float4 arrays of length max_particles
 *   P  = (x, y, z, Vp)            v  = (vx, vy, vz, bits(material id))
 *   F0 = (F col 0, tau_xx)        B0 = (B col 0, tau_xy)
 *   F1 = (F col 1, tau_yy)        B1 = (B col 1, tau_xz)
 *   F2 = (F col 2, tau_zz)        B2 = (B col 2, tau_yz)
 * The material id is a uint32 stored bit-for-bit in v.w.
 */
#ifndef NEBULA_MPM_H
#define NEBULA_MPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NB_MPM_ABI_VERSION 1

/* status codes */
#define NB_OK 0
#define NB_ERR_ARG (-1)      /* null handle/pointer or malformed argument     */
#define NB_ERR_RANGE (-2)    /* count or index beyond a documented capacity   */
#define NB_ERR_HIP (-3)      /* a HIP runtime call failed (message has detail) */
#define NB_ERR_STATE (-4)    /* call not valid in the engine's current state  */
#define NB_ERR_UNSUPPORTED (-5)

/* capacities, as allocated by the reference (src/engine/MPMEngine.cpp:75,82;
 * src/engine/EngineBase.h:34) */
#define NB_MAX_MATERIALS 16u
#define NB_MAX_COLLIDERS 64u
#define NB_MAX_FORCES 32u

/* material models — src/core/MaterialParams.h:5-12 */
#define NB_MAT_ELASTIC 0u
#define NB_MAT_VON_MISES 1u
#define NB_MAT_DRUCKER_PRAGER 2u
#define NB_MAT_GRANULAR_POWDER 3u
#define NB_MAT_FLUID 4u
#define NB_MAT_VISCOPLASTIC_MUD 5u

/* analytic collider types — src/core/Collider.h:8-13 */
#define NB_COLL_PLANE 0u
#define NB_COLL_SPHERE 1u
#define NB_COLL_BOX 2u
#define NB_COLL_CAPSULE 3u

/* force types — src/core/Force.h:13-19 */
#define NB_FORCE_GRAVITY 0u
#define NB_FORCE_CONSTANT_WIND 1u
#define NB_FORCE_TURBULENCE 2u
#define NB_FORCE_NOISE 3u
#define NB_FORCE_Z_CLAMP 4u

/* particle fields (the eight float4 arrays above) */
#define NB_FIELD_P 0
#define NB_FIELD_V 1
#define NB_FIELD_F0 2
#define NB_FIELD_F1 3
#define NB_FIELD_F2 4
#define NB_FIELD_B0 5
#define NB_FIELD_B1 6
#define NB_FIELD_B2 7
#define NB_NUM_FIELDS 8

/* 64-byte material record, layout of MaterialParams (src/core/MaterialParams.h:16-38) */
typedef struct nb_material {
  float mu;          /* 0  Lame mu = E/(2(1+nu))                 */
  float lambda;      /* 4  Lame lambda = E nu/((1+nu)(1-2nu))    */
  float rho0;        /* 8  rest density; particle mass = rho0*Vp */
  uint32_t model;    /* 12 NB_MAT_*                               */
  float M_friction;  /* 16 Drucker-Prager M = tan(friction angle) */
  float q_cohesion;  /* 20 Drucker-Prager cohesion               */
  float q_max;       /* 24 Von Mises yield (on Hencky dev strain) */
  float bulkK;       /* 28 fluid bulk modulus                     */
  float fluidGamma;  /* 32 */
  float viscosity;   /* 36 viscoplastic mud viscosity            */
  float hardening;   /* 40 carried, unread (as in the reference)  */
  float xi;          /* 44 carried, unread                        */
  float pad[4];      /* 48..63 */
} nb_material;

/* 64-byte analytic collider record, layout of ColliderPrimitive (src/core/Collider.h:19-29) */
typedef struct nb_collider {
  uint32_t type;      /* NB_COLL_*                                     */
  float px, py, pz;   /* point on plane / centre / capsule start        */
  float nx, ny, nz;   /* plane normal / box half extents / capsule axis */
  float radius;       /* sphere or capsule radius                       */
  float vx, vy, vz;   /* collider velocity                              */
  float restitution;
  float friction;
  float pad[3];
} nb_collider;

/* 32-byte force record, layout of ForceGPU (src/core/Force.h:27-32):
 * data[] meaning per type follows Force.h (gravity/wind: dir xyz in data[0..2],
 * strength in data[3]). */
typedef struct nb_force {
  uint16_t type;        /* NB_FORCE_* */
  uint16_t affect_mask; /* 0 = applies to every node */
  float data[7];
} nb_force;

/* MPMConfig (src/engine/MPMEngine.h:22-55).  Particles of the default block are
 * an nx*ny*nz lattice placed as MPMEngine::init does (src/engine/MPMEngine.cpp:86-137). */
typedef struct nb_mpm_config {
  uint32_t nx, ny, nz;
  float domain_size[3]; /* physical size [m]; worldMin = 0, worldMax = domain_size */
  float cell_size;      /* h, shared by all axes */
  uint32_t max_particles; /* 0 = nx*ny*nz */
  float E, nu, rho0;    /* default material (presetJelly) */
} nb_mpm_config;

/* The public knobs of MPMEngine (src/engine/MPMEngine.h:72-80). */
typedef struct nb_mpm_knobs {
  float restitution;     /* carried, unread (as in the reference) */
  float wall_friction;   /* carried, unread */
  int32_t num_substeps;  /* substeps per nb_mpm_step; dt_sub = dt/max(1,n) */
  uint32_t plastic_model;/* used only when no material table is set (never, as in the reference) */
  float q_max;
  float M_friction;
  float q_cohesion;
  float flip_ratio;      /* 0 = PIC, < -0.5 = APIC, > 0.01 = FLIP blend */
  /* 0 (default): the reference's SVD convention.  Its F^T F Jacobi leaves the
   * singular vectors in the rows of the GLSL mat3 while the Hencky stress and
   * the return-map recomposition use them as columns (shaders/mpm_common.glsl:
   * 250-293, shaders/mpm_g2p.comp:35,67), i.e. tau = U^T k U and
   * F <- U^T exp(eps) V with U, V the true singular-vector matrices.
   * 1: a proper polar SVD (U, V rotations, singular vectors as columns:
   * tau = U k U^T, F <- U exp(eps) V^T); an extension with no counterpart in
   * the reference. */
  uint32_t corrected_svd;
} nb_mpm_knobs;

typedef struct nb_mpm nb_mpm; /* opaque engine handle */

/* FP64 SVD fallback for near-degenerate F in the Drucker-Prager family
 * (Drucker-Prager, granular powder, viscoplastic mud).  No counterpart in the
 * reference, whose shaders are FP32 throughout.  G2P takes every SVD in FP32.
 * FP32 rounds F^T F by about 2^-22 s0^2, so for singular values
 * s0 >= s1 >= |s2| the vectors of a pair with s_i^2 - s_j^2 <= gap_rel * s0^2
 * are fixed only up to a rotation in their plane (of angle about
 * 2^-22 s0^2 / (s_i^2 - s_j^2), at most 1), and U = F v / sigma is set by
 * rounding when s2^2 <= zero_rel * s0^2.  The particle's return map (SVD,
 * projection and recomposition) and its Hencky stress are computed again in
 * FP64, with the same algorithm, and rounded to FP32 once, when s2 is that
 * small, or when such a rotation of a close pair would move the FP32 F' or
 * tau by more than 1e-5 of its scale (max |F'|; max(max |tau|,
 * 2 mu + 3 lambda)), the accuracy the FP32 path has anyway where the
 * singular values are well separated.  The reference convention's F <- U^T exp(eps) V and
 * tau = U^T k U change under the rotation by its angle times their
 * commutator with the rotation's generator; the corrected convention's by
 * the angle times the difference of the pair's values.  An F within rounding
 * of a multiple of the identity (material at rest or in free fall) has every
 * pair close but a result that commutes with every rotation, and stays on
 * the FP32 path.  The stress SVD of a recomposed F is tested the same way.
 * Particles that do not take the FP64 path give the same results bit for bit
 * with the fallback on or off.
 * Defaults: enabled, gap_rel 1e-4, zero_rel 1e-6, count_stats 0. */
typedef struct nb_mpm_svd_fallback {
  uint32_t enabled;
  float gap_rel;
  float zero_rel;
  uint32_t count_stats; /* 1: G2P counts the lanes and waves that take it */
} nb_mpm_svd_fallback;
/* counters of nb_mpm_svd_fallback_stats */
#define NB_SVD_FB_STATS 5
#define NB_SVD_FB_DP_LANES 0    /* Drucker-Prager-family particles updated     */
#define NB_SVD_FB_PROJ_LANES 1  /* of them, return maps taken in FP64          */
#define NB_SVD_FB_STRESS_LANES 2/* of them, stresses taken in FP64             */
#define NB_SVD_FB_FB_WAVES 3    /* wavefronts with at least one FP64 lane      */
#define NB_SVD_FB_DP_WAVES 4    /* wavefronts with at least one such particle  */

/* Thread-local description of the last failed call on this thread. */
const char* nb_last_error(void);
int nb_mpm_abi_version(void);

/* Fill cfg with MPMConfig's member defaults (src/engine/MPMEngine.h:25-39). */
void nb_mpm_config_default(nb_mpm_config* cfg);
/* Fill knobs with MPMEngine's member defaults (src/engine/MPMEngine.h:72-80). */
void nb_mpm_knobs_default(nb_mpm_knobs* knobs);

/* MPMEngine::init (src/engine/MPMEngine.cpp:34-156).  hip_stream: a hipStream_t
 * to enqueue on, or NULL for an engine-owned stream of default priority. */
int nb_mpm_create(const nb_mpm_config* cfg, void* hip_stream, nb_mpm** out);
/* MPMEngine::cleanup (src/engine/MPMEngine.cpp:160-163). */
int nb_mpm_destroy(nb_mpm* h);

/* ---- Batches of scenes ----
 * No counterpart in the reference, whose MPMEngine holds one scene.  For
 * callers that advance many independent scenes (parameter sweeps, training
 * data, differentiable fits): one handle holds them all and every substep
 * advances every scene in the same eleven launches.
 *
 * Each scene has its own domain [0, domain_size] and cell size (its grid is
 * ceil(domain_size / cell_size) cells, Domain.h's rule) and owns
 * particle_count consecutive particle rows, scenes in table order.  Particle
 * positions are in the scene's own coordinates.  All the particle, material,
 * force, knob, step and reverse-mode entry points work on a batch over the
 * concatenated rows; material ids index one shared table of up to
 * NB_MAX_BATCH_MATERIALS materials; forces, knobs and analytic colliders apply
 * to every scene (colliders in each scene's coordinates).  Scenes never
 * interact.  A batch starts with every particle at the origin, at rest,
 * F = I and material 0: set the state with nb_mpm_set_field.  Not available
 * on a batch (NB_ERR_UNSUPPORTED): the dense SDF, the pass-level entry points
 * and their exports, and changing the particle count. */
#define NB_MAX_BATCH_MATERIALS 1024
typedef struct nb_mpm_scene {
  float domain_size[3];
  float cell_size;
  uint32_t particle_count;
  uint32_t reserved[3];
} nb_mpm_scene;
int nb_mpm_create_batch(const nb_mpm_scene* scenes, uint32_t n_scenes, void* hip_stream, nb_mpm** out);
/* ---- Launch shapes, the autotuner and graph replay ----
 * No counterpart in the reference, whose dispatch sizes are fixed in its
 * shaders.  tile_blocks: blocks of the per-tile kernels (bin, P2G, grid; a
 * block walks tiles in strides); g2p_block: threads per G2P block (64, 128
 * or 256); p2g_item: particles one P2G thread takes from a cell before the
 * cell is split into more work items (4..32).  tile_blocks and g2p_block
 * never change results; p2g_item changes the order of P2G's sums, so results
 * move within rounding.  ms_per_substep: what nb_mpm_autotune measured for
 * the choice (0 when set by hand).  A new handle starts at tile_blocks 2048,
 * g2p_block 128, p2g_item 12. */
typedef struct nb_mpm_tuning {
  uint32_t tile_blocks;
  uint32_t g2p_block;
  uint32_t p2g_item;
  float ms_per_substep;
} nb_mpm_tuning;
int nb_mpm_get_tuning(const nb_mpm* h, nb_mpm_tuning* t);
int nb_mpm_set_tuning(nb_mpm* h, const nb_mpm_tuning* t);
/* Time candidate shapes on the engine's current state (trial_substeps
 * substeps of dt_sub each, the particle state restored after every trial)
 * and keep the fastest: a coordinate search over tile_blocks, g2p_block
 * and, with NB_TUNE_ROUNDING in flags, p2g_item.  The state is left as it
 * was; the choice is returned in *chosen (may be NULL). */
#define NB_TUNE_ROUNDING 1u
int nb_mpm_autotune(nb_mpm* h, float dt_sub, uint32_t trial_substeps, uint32_t flags, nb_mpm_tuning* chosen);
/* nb_mpm_autotune behind a persisted cache: the file at cache_path holds one
 * JSON object per line, {"key": "...", "tile_blocks": .., "g2p_block": ..,
 * "p2g_item": .., "ms_per_substep": ..}.  The key names the workload's shape:
 * the GPU (architecture and compute units), the live particle count and the
 * occupied tile count (each to the nearest power of two), the scene count,
 * the transfer mode, the material models present and the flags.  A cached
 * key sets the tuning without timing anything (*hit = 1); otherwise the
 * tuner runs and its choice is appended to the file (*hit = 0).  A missing
 * file is an empty cache; lines that do not parse are skipped. */
int nb_mpm_autotune_cached(nb_mpm* h, float dt_sub, uint32_t trial_substeps, uint32_t flags, const char* cache_path,
                           nb_mpm_tuning* chosen, int* hit);
/* n substeps as replays of one captured hipGraph of the substep's launches
 * (captured on first use, again whenever dt_sub, the particle count, the
 * material/force/collider counts, the knobs or the tuning change).  Same
 * results as nb_mpm_substeps, bit for bit; for short substeps where launch
 * overhead matters (batches of small scenes). */
int nb_mpm_graph_substeps(nb_mpm* h, float dt_sub, uint32_t n);

/* Scenes of a handle (0 for a plain engine). */
int nb_mpm_scene_count(const nb_mpm* h, uint32_t* n);
/* Grid resolution, first particle row and particle count of one scene. */
int nb_mpm_scene_info(const nb_mpm* h, uint32_t scene, uint32_t res[3], uint32_t* first_particle,
                      uint32_t* particle_count);

/* MPMEngine::step (src/engine/MPMEngine.cpp:404-474): forces, then
 * knobs.num_substeps substeps of dt/num_substeps.  Asynchronous on the stream. */
int nb_mpm_step(nb_mpm* h, float dt);
/* One substep of size dt_sub (the body of the reference's substep loop). */
int nb_mpm_substep(nb_mpm* h, float dt_sub);
/* Enqueue n substeps of size dt_sub. */
int nb_mpm_substeps(nb_mpm* h, float dt_sub, uint32_t n);
/* Wait for all work enqueued on the engine's stream. */
int nb_mpm_sync(nb_mpm* h);
/* Enqueue the engine's later work on hip_stream (a hipStream_t), or on an
 * engine-owned stream of default priority when NULL: the counterpart of the
 * command buffer MPMEngine::step records into (src/engine/MPMEngine.cpp:404).
 * Waits first for what the previous stream holds.  The caller keeps its
 * stream alive while the engine uses it. */
int nb_mpm_set_stream(nb_mpm* h, void* hip_stream);

int nb_mpm_set_knobs(nb_mpm* h, const nb_mpm_knobs* knobs);
int nb_mpm_get_knobs(const nb_mpm* h, nb_mpm_knobs* knobs);
/* The FP64 SVD fallback (nb_mpm_svd_fallback above); an extension with no
 * counterpart in the reference.  gap_rel and zero_rel must lie in [0, 1). */
int nb_mpm_set_svd_fallback(nb_mpm* h, const nb_mpm_svd_fallback* f);
int nb_mpm_get_svd_fallback(const nb_mpm* h, nb_mpm_svd_fallback* f);
/* The fallback's counters since the last reset (NB_SVD_FB_STATS words, see
 * NB_SVD_FB_*), counted while count_stats is 1; reset != 0 zeroes them
 * afterwards.  Synchronises the engine's stream. */
int nb_mpm_svd_fallback_stats(nb_mpm* h, uint64_t* out, int reset);
/* on != 0: nb_mpm_substeps runs the grid update inside G2P (one kernel per
 * occupied tile's particles, the tile's nodes in LDS) for every substep of a
 * call but the last; results are the same bit for bit.  Off by default: on
 * MI355X the fused kernel measures slower than the separate grid update and
 * G2P.  An extension with no counterpart in the reference. */
int nb_mpm_set_fused_g2p(nb_mpm* h, int on);
int nb_mpm_get_config(const nb_mpm* h, nb_mpm_config* cfg);

/* Domain helpers — src/core/Domain.h:14-39 and MPMConfig::totalCells. */
int nb_mpm_grid_res(const nb_mpm* h, uint32_t res[3]);
int nb_mpm_total_cells(const nb_mpm* h, uint32_t* total_cells);
/* MPMEngine::liveParticleCount. */
int nb_mpm_live_particle_count(const nb_mpm* h, uint32_t* n);

/* MPMEngine::setMaterials (src/engine/MPMEngine.cpp:239-243): n == 0 is a no-op,
 * n > NB_MAX_MATERIALS is NB_ERR_RANGE (NB_MAX_BATCH_MATERIALS on a batch). */
int nb_mpm_set_materials(nb_mpm* h, const nb_material* mats, uint32_t n);
/* MPMEngine::setParticleMaterialIds (src/engine/MPMEngine.cpp:245-256): overwrites
 * v[0..n) with (0, 0, 0, bits(id)). */
int nb_mpm_set_particle_material_ids(nb_mpm* h, const uint32_t* ids, uint32_t n);
/* MPMEngine::setColliders / clearAnalyticColliders (src/engine/MPMEngine.cpp:260-267). */
int nb_mpm_set_colliders(nb_mpm* h, const nb_collider* cols, uint32_t n);
int nb_mpm_clear_analytic_colliders(nb_mpm* h);
/* MPMEngine::setColliderSDF (src/engine/MPMEngine.cpp:204-209): a dense float
 * SDF indexed by Morton cell id; n must equal the total cell count. */
int nb_mpm_set_collider_sdf(nb_mpm* h, const float* morton_sdf, uint64_t n);
/* buildMeshSDF (src/core/MeshSDF.h:85-126) computed on the GPU and installed
 * as the collider SDF, as the reference's examples do with setColliderSDF
 * (examples/pyro_basic.cpp:123-134): tris holds n_tris triangles of 12 floats
 * (normal, v0, v1, v2; loadBinarySTL's MeshTriangle order), in the engine's
 * world coordinates.  Cells are signed by the normal of their closest
 * triangle; cube cells beyond the grid hold 1e9.
 * exact = 0 reproduces closestPtOnTriangle (MeshSDF.h:48-74), which has no
 * Voronoi region for the edge v1-v2: a cell nearest that edge gets its
 * distance to the triangle's plane, and near a mesh's convex edges that can
 * be too small and of the wrong sign.  exact = 1 adds the region (Ericson,
 * Real-Time Collision Detection, 5.1.5) and gives the true distance.  No
 * counterpart in the reference. */
int nb_mpm_set_collider_mesh(nb_mpm* h, const float* tris, uint32_t n_tris, int exact);
/* The collider SDF in force (totalCells floats, Morton order). */
int nb_mpm_get_collider_sdf(nb_mpm* h, float* out);
/* MPMEngine::setColliderSphere / clearCollider (src/engine/MPMEngine.cpp:180-202). */
int nb_mpm_set_collider_sphere(nb_mpm* h, float radius, float cx, float cy, float cz);
int nb_mpm_clear_collider(nb_mpm* h);
/* EngineBase::setForces / clearForces (src/engine/EngineBase.cpp:22-40).
 * Gravity, constant wind, turbulence (curl of fBm noise), noise (fBm along a
 * direction) and z-clamp (a no-op on a grid, as in the reference); the force
 * records are read by a compiled device switch (the reference regenerates
 * its grid shader instead, src/core/ForceShaderCompiler.cpp:139-166).  The
 * caller advances the records' seeds between steps (Force::advanceTime). */
int nb_mpm_set_forces(nb_mpm* h, const nb_force* forces, uint32_t n);

/* MPMEngine::appendParticles (src/engine/MPMEngine.cpp:211-235): host arrays of
 * n float4 (pos = (x, Vp), vel = (v, bits(id))); F = I, B = 0, tau = 0; silently
 * clamped to capacity.  *appended (may be NULL) receives the count written. */
int nb_mpm_append_particles(nb_mpm* h, const float* pos4, const float* vel4, uint32_t n, uint32_t* appended);
/* MPMEngine::resetParticles (src/engine/MPMEngine.cpp:281-287): no live
 * particles; every emitter's step count and the emitter RNG (mt19937, seed
 * 12345) start again. */
int nb_mpm_reset_particles(nb_mpm* h);

/* ---- Emitters ----
 * MPMEngine::addEmitter / clearEmitters / emitFromEmitters
 * (src/engine/MPMEngine.cpp:271-352) with the shapes of src/core/Emitter.h:
 * 39-224 in one record.  nb_mpm_step emits first, as the reference's step
 * does; nb_mpm_emit runs that part alone.  Particles are sampled on the host
 * with std::mt19937 (seed 12345, MPMEngine.h:165) and
 * std::uniform_real_distribution<float>, shape by shape as Emitter::sample,
 * then velocities as Emitter::sample_velocity; where the reference draws the
 * three components inside one constructor call (an order C++ leaves to the
 * compiler) they are drawn x, y, z.  Volume per particle:
 * MPMConfig::particleVolume (domain / (nx, ny, nz)). */
#define NB_EMITTER_AABB 0u
#define NB_EMITTER_SPHERE 1u
#define NB_EMITTER_ELLIPSE 2u
typedef struct nb_emitter {
  uint32_t shape;              /* NB_EMITTER_* */
  float center[3];
  float size[3];               /* AABB: edge lengths; SPHERE: radius in size[0];
                                  ELLIPSE: semi-axes along x and y in size[0], size[1] (z = center z) */
  float vel[3];                /* initial velocity (its length along the normal with emit_along_normal) */
  float center_vel[3];         /* the emitter moves by center_vel dt per emitting step */
  int32_t particles_per_step;
  uint32_t particle_type;      /* material id */
  uint32_t emit_along_normal;  /* 1: along the shape's outward normal */
  float velocity_randomness;   /* > 0: add a uniform [-r, r]^3 perturbation */
  int32_t step_count;          /* 0: every step, -1: the first step only, n > 0: n steps */
  uint32_t reserved[2];
} nb_emitter;
int nb_mpm_add_emitter(nb_mpm* h, const nb_emitter* e);
int nb_mpm_clear_emitters(nb_mpm* h);
int nb_mpm_emit(nb_mpm* h, float dt);
/* Emitter i as it stands (its center moves); *count the number of emitters. */
int nb_mpm_get_emitter(const nb_mpm* h, uint32_t i, nb_emitter* e, uint32_t* count);

/* Bulk state access in the reference's float4 layout (host pointers,
 * synchronous).  Rows [first, first+count) of field NB_FIELD_*. */
int nb_mpm_get_field(nb_mpm* h, int field, float* out4, uint32_t first, uint32_t count);
int nb_mpm_set_field(nb_mpm* h, int field, const float* in4, uint32_t first, uint32_t count);
/* Device pointer of a field (what getPositionBuffer/getVelocityBuffer return). */
int nb_mpm_field_device_ptr(nb_mpm* h, int field, void** dptr);
/* Set the live particle count directly (<= capacity); used with set_field. */
int nb_mpm_set_live_particle_count(nb_mpm* h, uint32_t n);

/* ---- pass-level access, the counterpart of tests/helpers/MPMHarness.h:56-59 ----
 * These run one stage of the substep on the current state so tests can isolate
 * it, and export the intermediate buffers in the reference's dense formats. */

/* Binning (K2-K7: zero_cells, mpm_hash_count, hash_scan_*, mpm_hash_sort). */
int nb_mpm_run_bin(nb_mpm* h);
/* After nb_mpm_run_bin: keys[N] (Morton cell id per particle), sorted[N]
 * (particle indices, stable by index within a cell), cell_count[NC] and
 * cell_end[NC] (END offsets, the reference's cellOffset after mpm_hash_sort).
 * Any pointer may be NULL. */
int nb_mpm_get_bins(nb_mpm* h, uint32_t* keys, uint32_t* sorted, uint32_t* cell_count, uint32_t* cell_end);
/* P2G (K1+K8) on the current bins; requires nb_mpm_run_bin first. */
int nb_mpm_run_p2g(nb_mpm* h, float dt_sub);
/* Grid update + dense-SDF BC (K9+K10) on the current grid. */
int nb_mpm_run_grid_update(nb_mpm* h, float dt_sub);
/* G2P (K11) on the current grid. */
int nb_mpm_run_g2p(nb_mpm* h, float dt_sub);
/* Dense grid export in the reference format: mom4 = 2*NC float4 ([0,NC) momentum
 * or velocity, [NC,2NC) saved pre-force velocity), mass = NC floats.  Nodes no
 * particle reaches read as zero. */
int nb_mpm_get_grid(nb_mpm* h, float* mom4, float* mass);
/* Dense grid import (MPMHarness::uploadGrid, tests/helpers/MPMHarness.cpp:182-188):
 * mom4 = NC float4 (the [0,NC) half), mass = NC floats, old4 = NC float4 of
 * saved pre-force velocity (the [NC,2NC) half) or NULL for zeros.  Makes the
 * whole grid visible to the next nb_mpm_run_grid_update / nb_mpm_run_g2p. */
int nb_mpm_set_grid(nb_mpm* h, const float* mom4, const float* mass, const float* old4);

/* ---- measurement ----
 * Run n substeps of dt_sub, recording HIP events around every kernel launch on
 * the engine's stream.  kernel_ms[i] receives the summed duration of kernel i,
 * names are written (comma separated) into names (capacity names_cap). */
int nb_mpm_profile_substeps(nb_mpm* h, float dt_sub, uint32_t n, float* kernel_ms, uint32_t kernel_cap,
                            uint32_t* n_kernels, char* names, uint32_t names_cap);
/* Launches of each kernel in the last nb_mpm_profile_substeps (same order as
 * its names): a call's substeps but the last run the grid update inside G2P
 * (g2p_grid), the last one runs grid_update and g2p. */
int nb_mpm_profile_calls(const nb_mpm* h, uint32_t* calls, uint32_t cap);
/* Calibration for HBM byte counters: copies `bytes` (rounded down to 16) from
 * one scratch buffer to another with a plain float4 copy kernel (k_copy),
 * `repeats` times, on the engine's stream; the buffers are allocated and freed
 * inside the call.  Diagnostic only; no counterpart in the reference. */
int nb_mpm_calibration_copy(nb_mpm* h, uint64_t bytes, uint32_t repeats);
/* Count of grid nodes with mass >= 1e-10 in the current grid (G_act). */
int nb_mpm_active_node_count(nb_mpm* h, uint64_t* g_act);
/* Number of occupied 8^3 cell tiles after the last bin. */
int nb_mpm_active_tile_count(nb_mpm* h, uint32_t* tiles);

/* ---- Reverse mode (adjoint) ----
 * The reference has no differentiation, so these entry points have no
 * counterpart in it.  Semantics: the substep
 * with corrected_svd = 1 (the reference's transposed-SVD stress is not a
 * differentiable function of F), PIC or APIC, elastic / Von Mises /
 * Drucker-Prager (granular powder) materials, forces, walls and the dense
 * SDF boundary; FLIP, analytic colliders, fluid and mud return
 * NB_ERR_UNSUPPORTED.  The J reset, the SDF push-out and the world clamp are
 * treated as constants where they act.
 *
 * Use: nb_mpm_adjoint_begin(h, n, segment) -> nb_mpm_forward_recorded(h, dt)
 * runs n substeps from the current state, keeping a checkpoint every `segment`
 * substeps -> fill the adjoint of the final state, dL/d(state), with
 * nb_mpm_set_adjoint (same eight float4 arrays and packing as the state:
 * dL/dx in P.xyz, dL/dv in v.xyz, dL/dF columns in F0..2.xyz, dL/dtau in the
 * w lanes, dL/dB columns in B0..2.xyz) -> nb_mpm_backward replaces it with
 * dL/d(state at the start of the run), recomputing the inner states of every
 * segment but the last from its checkpoint (segment = 1 or segment = n keeps
 * all n states and recomputes nothing), and returns dL/d(mu, lambda,
 * M_friction, q_cohesion) of every material (material_count x 4 floats) in
 * d_params.  The engine is left in the state the recorded run ended in (its
 * grid is not kept: nb_mpm_get_grid needs another substep first).
 * Memory: per particle of capacity (ceil(n/segment) + segment - 1) x 132 B
 * (records and bin keys of the kept states) + segment x 8 B (each substep's
 * sorted order) + 480 B; per grid node segment x 40 B (each substep's grid
 * and cell counts, kept so that the backward does not bin again or redo P2G
 * and the grid update) + 32 B. */
int nb_mpm_adjoint_begin(nb_mpm* h, uint32_t n_substeps, uint32_t segment);
/* nb_mpm_adjoint_begin with `tapes` segments' inner states kept at once
 * (nb_mpm_adjoint_begin keeps one): the recorded run leaves the last `tapes`
 * segments in place and the backward recomputes only the others, each from
 * its checkpoint.  tapes = 0 takes as many as fit in the device memory free
 * at the call (4 GiB and a tenth of the rest left over); tapes >= the number
 * of segments recomputes nothing.  Gradients are the same bit for bit
 * whatever `tapes` is.  Memory: the segment-dependent terms above times
 * `tapes`. */
int nb_mpm_adjoint_begin_tapes(nb_mpm* h, uint32_t n_substeps, uint32_t segment, uint32_t tapes);
/* The number of tapes the current nb_mpm_adjoint_begin* call reserved. */
int nb_mpm_adjoint_tapes(const nb_mpm* h, uint32_t* tapes);
int nb_mpm_forward_recorded(nb_mpm* h, float dt_sub);
int nb_mpm_set_adjoint(nb_mpm* h, int field, const float* in4, uint32_t first, uint32_t count);
int nb_mpm_get_adjoint(nb_mpm* h, int field, float* out4, uint32_t first, uint32_t count);
int nb_mpm_adjoint_device_ptr(nb_mpm* h, int field, void** dptr);
int nb_mpm_backward(nb_mpm* h, float* d_params, uint32_t params_cap);
int nb_mpm_adjoint_end(nb_mpm* h);

#ifdef __cplusplus
}
#endif

#endif /* NEBULA_MPM_H */
