// nb_engine.hpp — internal state of the MI355X MPM engine and the launcher
// interface between the host translation unit (mpm_engine.hip) and the kernel
// translation unit (mpm_kernels.hip).
//
// HBM layout (one GPU):
//   particles  8 x float4 x capacity, the reference's SoA (P, v, F0-2, B0-2),
//              which the ABI exposes, and the working copy the kernels use:
//              one 128-byte record per particle (the same 32 floats,
//              reordered), converted at the API boundary
//   bins       keys[N], rank[N], sorted[N]; cell_count/cell_end dense over the
//              Morton cube (only occupied 8^3 tiles are ever touched)
//   tiles      the Morton cube cut into 8^3-cell tiles (tile id = key >> 9):
//              per-tile flags, active-tile list, dilated node-tile list
//   partials   per active tile, the 10^3 node box its particles reach
//              (float4 = momentum xyz, mass), written without atomics
//   grid       grid_node[NC] float4 = (velocity or momentum, mass),
//              grid_old[NC] float4 = pre-force velocity (FLIP)
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nebula_mpm.h"

namespace nb {

constexpr int kTileBits = 9;          // 8^3 cells per tile
constexpr int kTileCells = 512;
constexpr int kBox = 10;              // node box edge of a tile (8 + 2 halo)
constexpr int kBoxNodes = 1000;

// Working particle record, 8 float4 = 128 bytes (one cache line), holding the
// same 32 floats as the reference's eight arrays P, v, F0-2, B0-2 (tau packed
// in their w lanes) in an order that puts everything P2G reads in the first
// 96 bytes:
//   R_P  = P                        (x, y, z, Vp)
//   R_V  = v                        (vx, vy, vz, material id bits)
//   R_B0 = B0, R_B1 = B1, R_B2 = B2 (affine columns; tau_xy, tau_xz, tau_yz)
//   R_T  = (F0.w, F1.w, F2.w, F0.x) (tau_xx, tau_yy, tau_zz, F col 0 x)
//   R_FA = (F0.y, F0.z, F1.x, F1.y)
//   R_FB = (F1.z, F2.x, F2.y, F2.z)
constexpr int kRec = 8;
enum : int { R_P = 0, R_V = 1, R_B0 = 2, R_B1 = 3, R_B2 = 4, R_T = 5, R_FA = 6, R_FB = 7 };
struct FieldPtrs {
  float4* f[NB_NUM_FIELDS];
};

// counters[] slots (device)
enum : int { CNT_ACTIVE = 0, CNT_NODE_TILES = 1, CNT_ALL_TILES = 2, CNT_UNITS = 3, CNT_SLOTS = 4 };
// particles per work unit of the fused grid-update + G2P kernel (a tile's
// particles split into units of at most this many)
constexpr uint32_t kG2PUnit = 1024;

// Geometry of one scene's grid.  A plain engine holds one scene, described by
// SubParams' own fields (key_base 0).  A batch holds a table of scenes, each
// with its own Morton cube placed at key_base (a multiple of 512) in the
// shared key space, so that each scene's tiles are consecutive tile ids;
// positions stay in the scene's own coordinates.
struct SceneGeo {
  float h, inv_h;      // cell size and its reciprocal
  float wmin[3];
  float wmax[3];
  int32_t grid_res[3];
  uint32_t cube_res;
  uint32_t tile_dim;   // tiles per axis of the scene's cube
  uint32_t key_base;   // first cell key of the scene
  uint32_t h_pow2;     // 1: h is a power of two, so (x - wmin) / h == (x - wmin) * inv_h exactly
  uint32_t pad;
};
static_assert(sizeof(SceneGeo) == 64, "one scene record per 64 bytes");

// 1 when h is a normal power of two whose reciprocal is a normal float: then
// x / h and x * (1 / h) are the same IEEE number for every x (both are exact
// scalings by a power of two, or overflow alike)
inline uint32_t cell_size_is_pow2(float h) {
  if (!(h >= 1.17549435e-38f && h <= 8.50705917e+37f)) return 0u;  // [2^-126, 2^126]
  int e = 0;
  return std::frexp(h, &e) == 0.5f ? 1u : 0u;
}

// Substep constants, passed by value to every kernel (the role MPMSimPC plays
// for the reference's shaders, src/core/MPMSimPC.h:14-73).
struct SubParams {
  float dt;
  float h;            // cell size
  uint32_t h_pow2;    // 1: h is a power of two (see SceneGeo)
  float wmin[3];
  float wmax[3];
  int32_t grid_res[3];
  uint32_t cube_res;
  uint32_t tile_dim;  // tiles per axis = max(cube_res, 8) / 8
  uint32_t n;         // live particles
  float flip_ratio;
  uint32_t material_count;
  uint32_t collider_count;
  uint32_t force_count;
  uint32_t write_old;  // write the pre-force velocity (FLIP, or pass-level export)
  uint32_t corrected_svd;  // 1: singular vectors as columns (see nb_mpm_knobs)
  uint32_t noise_forces;   // turbulence or noise records among the forces
  // FP64 SVD of near-degenerate F for the Drucker-Prager family (see
  // nb_mpm_svd_fallback): on/off, the two relative thresholds, and the
  // counters G2P adds to when non-null (NB_SVD_FB_STATS 64-bit words).
  // G2P queues the particles the fallback takes over (fb_queue, slot
  // lengths in fb_ctl); k_g2p_fb then replaces their F and tau.
  uint32_t svd_fp64;
  float svd_gap_rel;
  float svd_zero_rel;
  unsigned long long* svd_stats;
  float4* fb_queue;
  uint32_t* fb_ctl;
  uint32_t fb_region;  // entries per queue slot: ceil(n / kFbSlots)
  // batch (n_scenes > 0): the scene table, the scene of every tile and of
  // every particle; the single-scene fields above are then unused
  // launch shape (nb_mpm_tuning): blocks of the per-tile kernels, threads
  // per G2P block, particles per P2G work item
  uint32_t tile_blocks;
  uint32_t g2p_block;
  uint32_t p2g_item;
  uint32_t n_scenes;
  const SceneGeo* scenes;
  const uint32_t* tile_scene;
  const uint32_t* pscene;
};
// the FP64 fallback's queue: slots, float4 per entry, words between slot counters
constexpr uint32_t kFbSlots = 64;
constexpr uint32_t kFbEntry = 5;
constexpr uint32_t kFbCtlStride = 64;
// materials a batch's table may hold (a plain engine: NB_MAX_MATERIALS)
constexpr uint32_t kBatchMaxMaterials = NB_MAX_BATCH_MATERIALS;

struct DevState {
  float4* field[NB_NUM_FIELDS];  // reference layout, what the ABI exposes
  float4* rec;                   // capacity * kRec: working particle records
  uint32_t* keys;
  uint32_t* rank;
  uint32_t* sorted;
  uint32_t* cell_count;   // NT*512
  uint32_t* cell_end;     // NT*512
  uint8_t* tile_flag;     // NT: tile holds a particle (set by the key pass)
  uint8_t* tile_mark;     // NT: bit0 active, bit1 node tile (dilation)
  uint64_t* blk_cnt;      // NT/256: per 256 tiles, active (low) and node (high) counts
  int32_t* tile_act;      // NT: index in act_tiles or -1
  uint32_t* act_tiles;    // NT
  uint32_t* node_tiles;   // NT
  uint32_t* all_tiles;    // NT: 0..NT-1 (dense pass-level mode)
  uint32_t* tile_total;   // NT
  uint32_t* tile_start;   // NT
  uint32_t* tile_order;   // NT: the occupied tiles' indices, most particles first (P2G's order)
  uint32_t* counters;     // CNT_SLOTS
  uint2* units;           // NT + cap / kG2PUnit + 1: (active tile, chunk) work units of k_g2p_tile
  float4* fb_queue;       // (cap + kFbSlots) * kFbEntry: particles G2P leaves to k_g2p_fb (FP64 fallback)
  uint32_t* fb_ctl;       // kFbSlots * kFbCtlStride + 1: slot lengths (zero between substeps)
  float4* partial;        // cap_tiles * 1000
  float4* grid_node;      // NT*512
  float4* grid_old;       // NT*512
  float4* grid_pre;       // NT*512 or null: (momentum / mass, mass) before the update, kept for the adjoint
  float* sdf;             // NC (null when no SDF collider)
  nb_material* materials;
  nb_collider* colliders;
  nb_force* forces;
};

// names of the kernels, in the order the profiler indexes them
constexpr int kNumProfiledKernels = 13;
extern const char* const kKernelNames[kNumProfiledKernels];

// Optional per-kernel timing: HIP events recorded around each launch on the
// launch stream; record() waits for the end event and accumulates.
struct Prof {
  hipEvent_t ev[2 * kNumProfiledKernels];
  double ms[kNumProfiledKernels];
  uint32_t calls[kNumProfiledKernels];
  void record(int id) {
    (void)hipEventSynchronize(ev[2 * id + 1]);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ev[2 * id], ev[2 * id + 1]);
    ms[id] += t;
    calls[id]++;
  }
};

// ---- launchers (mpm_kernels.hip); pf may be null ----
// keys_valid: the previous G2P wrote every particle's key (else compute them)
void launch_bin(const DevState& d, const SubParams& p, uint32_t n_tiles, bool keys_valid, hipStream_t s, Prof* pf);
void launch_p2g(const DevState& d, const SubParams& p, uint32_t n_tiles, hipStream_t s, Prof* pf);
// Sum partials into grid_node and apply the grid update (+ dense-SDF BC).
// update=false only sums (pass-level P2G export).  tiles/count select the
// node-tile list (the dilated list, or all tiles in dense mode).
void launch_grid(const DevState& d, const SubParams& p, uint32_t n_tiles, bool from_partials, bool update,
                 const uint32_t* tiles, const uint32_t* count, hipStream_t s, Prof* pf);
// use_sorted: visit particles in bin order (else index order)
// rec_out: where the updated records go (nullptr: in place over d.rec)
// Grid update and G2P in one kernel (k_g2p_tile), from the partials P2G wrote
// and the work units the bin listed; the grid arrays are not written.
void launch_g2p_fused(const DevState& d, const SubParams& p, uint32_t n_tiles, hipStream_t s, Prof* pf);
void launch_g2p_fb(const DevState& d, const SubParams& p, float4* rec, hipStream_t s, Prof* pf);
void launch_g2p(const DevState& d, const SubParams& p, bool use_sorted, hipStream_t s, Prof* pf,
                float4* rec_out = nullptr);
void launch_sdf_sphere(float* sdf, const SubParams& p, float r, float cx, float cy, float cz, hipStream_t s);
// the Morton SDF of a triangle mesh (n x 12 floats: normal, v0, v1, v2) on a grid of res cells of size cs
void launch_mesh_sdf(const float* tris, uint32_t n_tris, const int32_t res[3], uint32_t cube, float cs, bool exact,
                     float* sdf, hipStream_t s);
void launch_fill_identity(uint32_t* a, uint32_t n, hipStream_t s);
void launch_init_particles(const DevState& d, uint32_t first, uint32_t count, hipStream_t s);
void launch_copy(float4* dst, const float4* src, uint64_t n4, hipStream_t s);
void launch_soa_to_aos(const DevState& d, uint32_t n, hipStream_t s);
void launch_aos_to_soa(const DevState& d, uint32_t n, hipStream_t s);
void launch_soa_to_aos_ptrs(const FieldPtrs& f, float4* rec, uint32_t n, hipStream_t s);
void launch_aos_to_soa_ptrs(const float4* rec, const FieldPtrs& f, uint32_t n, hipStream_t s);

// Buffers of the reverse-mode substep (mpm_adjoint.hip).
struct AdjState {
  float4* coef;         // capacity * 6: per-particle grid-adjoint coefficients
  float4* partial_adj;  // cap_tiles * kBoxNodes: per-tile grid-adjoint boxes
  float4* grid_adj;     // NT * 512: (dL/dp, dL/dm) per node
  // material-parameter partials of the G2P backward: per block of the G2P
  // backward and material, d/d(mu, lambda, M, q_c) summed over the substeps of
  // the backward so far (each block adds to its own row, substep after
  // substep); launch_adjoint_params sums the rows at the end
  float4* pg_acc;       // blocks * pg_materials
  uint32_t pg_materials;
  float4* pg_chunk;     // materials x ceil(blocks / 64): per-chunk sums
  float* pgrad;         // materials x 4: d/d(mu, lambda, M, q_c)
};
// One substep backwards on the recomputed forward of state d.rec (bins, P2G
// partials and grid velocities already in d): adj_in is the adjoint of the
// state after the substep, adj_out receives the adjoint of d.rec.
void launch_adjoint_substep(const DevState& d, const AdjState& a, const SubParams& p, uint32_t n_tiles,
                            const float4* adj_in, float4* adj_out, hipStream_t s);
// pgrad = the sum of pg_acc's rows, per material, in a fixed order
void launch_adjoint_params(const AdjState& a, uint32_t n_particles, uint32_t n_materials, hipStream_t s);

// Sets the thread-local message nb_last_error() returns; returns code.
int set_error(int code, const char* fmt, ...);

}  // namespace nb
