// mpm_kernels.hip — the MPM substep as hand-written gfx950 kernels.
//
// One substep (the reference's 11 dispatches, src/engine/MPMEngine.cpp:415-473)
// runs here as:
//   bin   k_bin_clear -> k_bin_keys -> k_bin_mark -> k_bin_compact ->
//         k_bin_tile_scan -> k_bin_tile_offsets -> k_bin_scatter -> k_bin_stabilize
//         (replaces zero_cells, mpm_hash_count, hash_scan_local/global,
//          hash_add_base, mpm_hash_sort; output is bit-identical to the stable
//          counting sort with END offsets of tests/helpers/MPMHarness.cpp:132-178,
//          with no group-count limit)
//   P2G   k_p2g: one block per occupied 8^3 tile; its occupied cells are
//         compacted onto threads, each thread walks its cell's particles (one
//         128-byte record each) and sums their 27 node contributions in
//         registers; the sums are merged into the tile's 10^3 LDS node box in 27
//         barrier-separated phases (every thread owns a distinct node in a
//         phase: no atomics, fixed order) and written as a per-tile partial
//         (replaces mpm_zero_grid + mpm_p2g.comp's node-centric gather)
//   grid  k_grid: one block per node tile; sums the <= 8 partials covering each
//         node in a fixed order, then velocity, cap, forces, walls, analytic
//         colliders and the dense-SDF BC in registers
//         (replaces mpm_grid_update.comp + mpm_nanovdb_bc.comp)
//   G2P   k_g2p: one thread per particle in bin order; 27-node gather, F update,
//         register SVD, return map, Hencky stress, guards, advection; also
//         writes the particle's cell key for the next substep's bin
//         (replaces mpm_g2p.comp)
// Particle state lives in an interleaved record per particle (the eight float4
// of the reference's layout side by side: 128 bytes, one cache line), so the
// gathers of P2G and G2P read whole lines; k_soa_to_aos / k_aos_to_soa move it
// to and from the reference's eight arrays at the API boundary.
// Every kernel is deterministic: identical inputs give bit-identical outputs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nb_device_math.hpp"
#include "nb_engine.hpp"
#include "nb_kernel_util.hpp"

namespace nb {

const char* const kKernelNames[kNumProfiledKernels] = {
    "bin_clear", "bin_keys", "bin_mark", "bin_compact", "bin_tile_scan",
    "bin_tile_offsets", "bin_scatter", "bin_stabilize", "p2g", "grid_update", "g2p", "g2p_grid", "g2p_fb"};

// ---------------------------------------------------------------------------
// BIN
// ---------------------------------------------------------------------------
// Zero the dense cell counters of the tiles the previous bin occupied; outside
// those tiles the counters are already zero.
__global__ __launch_bounds__(512) void k_bin_clear(uint32_t* __restrict__ cell_count,
                                                   const uint32_t* __restrict__ act_tiles,
                                                   const uint32_t* __restrict__ counters) {
  const uint32_t na = counters[CNT_ACTIVE];
  for (uint32_t a = blockIdx.x; a < na; a += gridDim.x) {
    cell_count[(size_t(act_tiles[a]) << kTileBits) + threadIdx.x] = 0u;
  }
}

// Key of every particle (computed from the record, or read when the previous
// G2P wrote it), its rank among equal keys (wave-aggregated atomics: one
// atomic per distinct key per wave), and the occupied-tile flags.
template <bool FROM_REC, bool BATCH>
__global__ __launch_bounds__(256) void k_bin_keys(SubParams p, const float4* __restrict__ rec,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ rank,
                                                  uint32_t* __restrict__ cell_count, uint8_t* __restrict__ tile_flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < p.n;
  uint32_t key = 0xFFFFFFFFu;
  if (valid) {
    if (FROM_REC) {
      const float4 x = rec[size_t(i) * kRec + R_P];
      key = cell_key(particle_geo<BATCH>(p, i), x.x, x.y, x.z);
      keys[i] = key;
    } else {
      key = keys[i];
    }
  }
  // group the wave's lanes by key (ballots only), then one atomic per group,
  // all groups in the same instruction
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t pending = __ballot(valid);
  int my_leader = 0;
  uint32_t my_pos = 0, my_cnt = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lkey = __shfl(key, leader);
    const uint64_t peers = __ballot(valid && key == lkey);
    if (valid && key == lkey) {
      my_leader = leader;
      my_pos = uint32_t(__popcll(peers & lt_mask));
    }
    if (int(lane) == leader) my_cnt = uint32_t(__popcll(peers));
    pending &= ~peers;
  }
  uint32_t base = 0;
  if (valid && my_pos == 0u) {
    base = atomicAdd(&cell_count[key], my_cnt);
    tile_flag[key >> kTileBits] = 1;
  }
  base = __shfl(base, my_leader);
  if (valid) rank[i] = base + my_pos;
}

// Per tile: bit0 = occupied, bit1 = some tile in its 3x3x3 neighbourhood is
// occupied (its nodes can receive mass).  Each block of 256 tiles also stores
// its counts of both kinds (occupied in the low, node tiles in the high half).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_bin_mark(SubParams p, const uint8_t* __restrict__ tile_flag,
                                                  uint8_t* __restrict__ tile_mark, uint64_t* __restrict__ blk_cnt,
                                                  uint32_t nt) {
  __shared__ uint64_t s_scan[256 / 64 + 1];
  const uint32_t T = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t m = 0;
  if (T < nt) {
    const SceneGeo g = tile_geo<BATCH>(p, T);  // neighbours stay inside the tile's scene
    int tx, ty, tz;
    tile_coords(g, T, tx, ty, tz);
    const int td = int(g.tile_dim);
    uint8_t node = 0;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int x = tx + dx, y = ty + dy, z = tz + dz;
          if (x < 0 || y < 0 || z < 0 || x >= td || y >= td || z >= td) continue;
          node |= tile_flag[tile_id(g, x, y, z)];
        }
    m = (tile_flag[T] ? 1u : 0u) | (node ? 2u : 0u);
    tile_mark[T] = uint8_t(m);
  }
  uint64_t tot;
  (void)block_exclusive_scan<256>(uint64_t(m & 1u) | (uint64_t(m >> 1) << 32), s_scan, tot);
  if (threadIdx.x == 0) blk_cnt[blockIdx.x] = tot;
}

// Ordered compaction of occupied tiles and node tiles, one block per 256
// tiles: the block's offsets are the sums of the counts of the blocks before
// it; resets the occupied flags for the next bin.
__global__ __launch_bounds__(256) void k_bin_compact(const uint8_t* __restrict__ tile_mark,
                                                     const uint64_t* __restrict__ blk_cnt,
                                                     uint8_t* __restrict__ tile_flag, int32_t* __restrict__ tile_act,
                                                     uint32_t* __restrict__ act_tiles,
                                                     uint32_t* __restrict__ node_tiles,
                                                     uint32_t* __restrict__ counters, uint32_t nt) {
  __shared__ uint64_t s_scan[256 / 64 + 1];
  uint64_t before = 0;
  for (uint32_t i = threadIdx.x; i < blockIdx.x; i += blockDim.x) before += blk_cnt[i];
  uint64_t off;
  (void)block_exclusive_scan<256>(before, s_scan, off);
  const uint32_t T = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t m = T < nt ? tile_mark[T] : 0u;
  uint64_t tot;
  const uint64_t pre = block_exclusive_scan<256>(uint64_t(m & 1u) | (uint64_t(m >> 1) << 32), s_scan, tot);
  const uint32_t oa = uint32_t(off) + uint32_t(pre), on = uint32_t(off >> 32) + uint32_t(pre >> 32);
  if (T < nt) {
    if (m & 1u) act_tiles[oa] = T;
    tile_act[T] = (m & 1u) ? int32_t(oa) : -1;
    if (m & 2u) node_tiles[on] = T;
    tile_flag[T] = 0;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    counters[CNT_ACTIVE] = uint32_t(off) + uint32_t(tot);
    counters[CNT_NODE_TILES] = uint32_t(off >> 32) + uint32_t(tot >> 32);
  }
}

// Per occupied tile: exclusive scan of its 512 cell counts (local cell start,
// kept in cell_end until the scatter) and the tile total.
__global__ __launch_bounds__(512) void k_bin_tile_scan(const uint32_t* __restrict__ cell_count,
                                                       uint32_t* __restrict__ cell_end,
                                                       const uint32_t* __restrict__ act_tiles,
                                                       uint32_t* __restrict__ tile_total,
                                                       const uint32_t* __restrict__ counters) {
  __shared__ uint32_t s_wave[512 / 64 + 1];
  const uint32_t na = counters[CNT_ACTIVE];
  for (uint32_t a = blockIdx.x; a < na; a += gridDim.x) {
    const size_t key = (size_t(act_tiles[a]) << kTileBits) + threadIdx.x;
    uint32_t c = cell_count[key];
    uint32_t tot;
    uint32_t ex = block_exclusive_scan<512>(c, s_wave, tot);
    cell_end[key] = ex;
    if (threadIdx.x == 0) tile_total[a] = tot;
  }
}

// Single block: exclusive scan of the tile totals in tile order, and the
// list of the fused G2P's work units (kG2PUnit particles of a tile each, tile
// by tile) with their count.  Particles (< 2^32) and units are scanned
// together as the low and high halves of one 64-bit sum.
__global__ __launch_bounds__(1024) void k_bin_tile_offsets(const uint32_t* __restrict__ tile_total,
                                                           uint32_t* __restrict__ tile_start,
                                                           uint32_t* __restrict__ tile_order,
                                                           uint32_t* __restrict__ counters,
                                                           uint2* __restrict__ units) {
  __shared__ uint64_t s_wave[1024 / 64 + 1], s_wave2[1024 / 64 + 1];
  const uint32_t na = counters[CNT_ACTIVE];
  const uint32_t chunk = (na + 1023u) / 1024u;
  const uint32_t lo = min(na, threadIdx.x * chunk), hi = min(na, lo + chunk);
  auto units_of = [](uint32_t t) { return (t + kG2PUnit - 1u) / kG2PUnit; };
  uint64_t s = 0;
  for (uint32_t a = lo; a < hi; a++) {
    const uint32_t t = tile_total[a];
    s += (uint64_t(units_of(t)) << 32) + t;
  }
  uint64_t tot;
  const uint64_t pre = block_exclusive_scan<1024>(s, s_wave, tot);
  uint32_t o = uint32_t(pre), uo = uint32_t(pre >> 32);
  for (uint32_t a = lo; a < hi; a++) {
    const uint32_t t = tile_total[a];
    tile_start[a] = o;
    o += t;
    for (uint32_t c = 0, nu = units_of(t); c < nu; c++) units[uo++] = make_uint2(a, c);
  }
  if (threadIdx.x == 0) counters[CNT_UNITS] = uint32_t(tot >> 32);
  // P2G's order of the tiles: four classes by particle count, largest first,
  // tile order within a class (a stable counting sort; the class counts are
  // packed 16 bits each into one 64-bit block scan).  A block's work grows
  // with its tile's particles, so the long tiles start first and the short
  // ones fill the end of the launch.  Which block takes which tile does not
  // change any result.
  auto cls = [](uint32_t t) { return t >= 2048u ? 0u : (t >= 1024u ? 1u : (t >= 512u ? 2u : 3u)); };
  uint64_t cnt = 0;
  if (na < 65536u)
    for (uint32_t a = lo; a < hi; a++) cnt += uint64_t(1) << (16u * cls(tile_total[a]));
  uint64_t ctot;
  const uint64_t cpre = block_exclusive_scan<1024>(cnt, s_wave2, ctot);
  if (na >= 65536u) {
    for (uint32_t a = lo; a < hi; a++) tile_order[a] = a;
  } else {
    uint32_t base[4];
    uint32_t acc = 0;
#pragma unroll
    for (uint32_t c = 0; c < 4u; c++) {
      base[c] = acc + (uint32_t(cpre >> (16u * c)) & 0xFFFFu);
      acc += uint32_t(ctot >> (16u * c)) & 0xFFFFu;
    }
    for (uint32_t a = lo; a < hi; a++) {
      const uint32_t c = cls(tile_total[a]);
      const uint32_t pos = c == 0u ? base[0]++ : (c == 1u ? base[1]++ : (c == 2u ? base[2]++ : base[3]++));
      tile_order[pos] = a;
    }
  }
}

// Place every particle at tile start + cell start + its rank.
__global__ __launch_bounds__(256) void k_bin_scatter(uint32_t n, const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ rank,
                                                     const uint32_t* __restrict__ cell_end,
                                                     const int32_t* __restrict__ tile_act,
                                                     const uint32_t* __restrict__ tile_start,
                                                     uint32_t* __restrict__ sorted) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t key = keys[i];
  const int32_t a = tile_act[key >> kTileBits];
  const uint32_t slot = tile_start[a] + cell_end[key] + rank[i];
  if (slot < n) sorted[slot] = i;  // always true when the counters started at zero
}

// Per cell: order its particles by index (the ranks of different waves arrive
// in atomic order), then store the END offset, the reference's cellOffset.
// Cells of up to 32 particles are sorted in registers by their thread:
// independent loads, an odd-even transposition network as wide as the wave's
// largest cell (8, 16 or 32) that stops once a round pair swaps nothing in the
// wave, write-back only if the order changed.  Larger
// cells (the SDF push-out piles particles up) are listed and then sorted by
// the whole block: staged in LDS, each element's place is the number of
// smaller elements (indices are distinct).  Cells beyond the LDS stage fall
// back to an in-place insertion sort by their thread.
constexpr int kRegSort = 32;
constexpr uint32_t kBlockSort = 2048;

template <int M>
__device__ __forceinline__ void cell_reg_sort(uint32_t* __restrict__ sorted, uint32_t start, uint32_t m) {
  uint32_t v[M];
#pragma unroll
  for (int i = 0; i < M; i++) v[i] = uint32_t(i) < m ? sorted[start + i] : 0xFFFFFFFFu;
  bool changed = false;
#pragma unroll
  for (int round = 0; round < M; round += 2) {
    bool swapped = false;
#pragma unroll
    for (int half = 0; half < 2; half++) {
#pragma unroll
      for (int i = half; i + 1 < M; i += 2) {
        const uint32_t lo = min(v[i], v[i + 1]), hi = max(v[i], v[i + 1]);
        swapped |= lo != v[i];
        v[i] = lo;
        v[i + 1] = hi;
      }
    }
    changed |= swapped;
    if (!__any(swapped)) break;
  }
  if (changed) {
#pragma unroll
    for (int i = 0; i < M; i++)
      if (uint32_t(i) < m) sorted[start + i] = v[i];
  }
}

__global__ __launch_bounds__(512) void k_bin_stabilize(uint32_t n, const uint32_t* __restrict__ cell_count,
                                                       uint32_t* __restrict__ cell_end,
                                                       const uint32_t* __restrict__ act_tiles,
                                                       const uint32_t* __restrict__ tile_start,
                                                       uint32_t* __restrict__ sorted,
                                                       const uint32_t* __restrict__ counters) {
  __shared__ uint32_t s_big_start[kTileCells], s_big_m[kTileCells];
  __shared__ uint32_t s_vals[kBlockSort];
  __shared__ uint32_t s_nbig;
  const uint32_t na = counters[CNT_ACTIVE];
  for (uint32_t a = blockIdx.x; a < na; a += gridDim.x) {
    if (threadIdx.x == 0) s_nbig = 0u;
    __syncthreads();
    const size_t key = (size_t(act_tiles[a]) << kTileBits) + threadIdx.x;
    const uint32_t start = tile_start[a] + cell_end[key];
    const uint32_t cnt = cell_count[key];
    const uint32_t stop = min(start + cnt, n);
    const uint32_t m = stop > start ? stop - start : 0u;
    const bool reg = m <= uint32_t(kRegSort);
    if (reg) {
      const uint32_t mr = m;
      if (__any(mr > 16u)) cell_reg_sort<32>(sorted, start, mr);
      else if (__any(mr > 8u)) cell_reg_sort<16>(sorted, start, mr);
      else cell_reg_sort<8>(sorted, start, mr);
    } else if (m <= kBlockSort) {
      const uint32_t slot = atomicAdd(&s_nbig, 1u);
      s_big_start[slot] = start;
      s_big_m[slot] = m;
    } else {
      for (uint32_t k = start + 1; k < stop; k++) {
        uint32_t v = sorted[k];
        uint32_t j = k;
        while (j > start && sorted[j - 1] > v) {
          sorted[j] = sorted[j - 1];
          j--;
        }
        sorted[j] = v;
      }
    }
    cell_end[key] = start + cnt;
    __syncthreads();
    const uint32_t nbig = s_nbig;
    for (uint32_t b = 0; b < nbig; b++) {
      const uint32_t bs = s_big_start[b], bm = s_big_m[b];
      for (uint32_t e = threadIdx.x; e < bm; e += blockDim.x) s_vals[e] = sorted[bs + e];
      __syncthreads();
      for (uint32_t e = threadIdx.x; e < bm; e += blockDim.x) {
        const uint32_t v = s_vals[e];
        uint32_t r = 0;
        for (uint32_t j = 0; j < bm; j++) r += s_vals[j] < v ? 1u : 0u;
        sorted[bs + r] = v;
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// P2G
// ---------------------------------------------------------------------------
// P2G: one block per occupied tile.  The tile's occupied cells are compacted
// (in Morton order) onto the block's threads as work items: a cell of up to
// p.p2g_item particles (tuned, 12 by default) is one item, a more crowded cell is split into up to
// kMaxChunks items of consecutive particles (so one crowded cell does not hold
// its wave for the whole walk).  A thread walks its item's particles in bin
// order, reading the first 96 bytes of each record, and sums their
// contributions to the 27 nodes the cell reaches in registers.  The sums are
// merged into the tile's 10^3 LDS node box chunk layer by chunk layer, each in
// 27 barrier-separated phases: within a phase the writing threads hold
// distinct cells, so every one adds to a distinct node, plain read-add-write
// is exact and the order of the sums is fixed.
constexpr int kP2GThreads = 256;
constexpr uint32_t kIdxStage = 4096;  // particle indices staged in LDS per pass

constexpr uint32_t kMaxChunks = 4;

__device__ __forceinline__ uint32_t p2g_chunks(uint32_t count, uint32_t item) {
  return count == 0u ? 0u : min((count + item - 1u) / item, kMaxChunks);
}
// item length groups, longest first: >= 12, 10-11, 8-9, 6-7, <= 5 particles
constexpr uint32_t kGroups = 5;
__device__ __forceinline__ uint32_t p2g_group(uint32_t count, uint32_t chunks) {
  if (chunks == 0u) return 0u;
  const uint32_t per = (count + chunks - 1u) / chunks;
  return per >= 12u ? 0u : (per >= 10u ? 1u : (per >= 8u ? 2u : (per >= 6u ? 3u : 4u)));
}

// One particle's node contributions, in the factored chains described at the
// walk below.  MASK: wx and wyz are set to zero at each node whose weight
// (wx wy) wz the reference skips (fma(a, 0, s) leaves s as it was); every
// other node gets the same bits as without the mask.
#define NB_P2G_CHAINS(MASK)                                                                           \
  {                                                                                                   \
    const float mp = q.mp, mvx = q.mvx, mvy = q.mvy, mvz = q.mvz;                                     \
    const float c00 = q.c00, c01 = q.c01, c02 = q.c02, c10 = q.c10, c11 = q.c11, c12 = q.c12;         \
    const float c20 = q.c20, c21 = q.c21, c22 = q.c22;                                                \
    const float t00 = q.t00, t11 = q.t11, t22 = q.t22, t01 = q.t01, t02 = q.t02, t12 = q.t12;         \
    const Axis3 &sx = q.sx, &sy = q.sy, &sz = q.sz;                                                   \
    float mwx[3], Qx[3], Qy[3], Qz[3], Rx[3], Ry[3], Rz[3], Sx[3], Sy[3], Sz[3];                      \
    _Pragma("unroll") for (int i = 0; i < 3; i++) {                                                   \
      mwx[i] = mp * sx.w[i];                                                                          \
      const float wdx = sx.w[i] * sx.dx[i];                                                           \
      Qx[i] = fmaf(-sx.g[i], t00, wdx * c00);                                                         \
      Qy[i] = fmaf(-sx.g[i], t01, wdx * c01);                                                         \
      Qz[i] = fmaf(-sx.g[i], t02, wdx * c02);                                                         \
      Rx[i] = fmaf(sy.w[i], fmaf(c10, sy.dx[i], mvx), -(sy.g[i] * t01));                              \
      Ry[i] = fmaf(sy.w[i], fmaf(c11, sy.dx[i], mvy), -(sy.g[i] * t11));                              \
      Rz[i] = fmaf(sy.w[i], fmaf(c12, sy.dx[i], mvz), -(sy.g[i] * t12));                              \
      const float wdz = sz.w[i] * sz.dx[i];                                                           \
      Sx[i] = fmaf(-sz.g[i], t02, wdz * c20);                                                         \
      Sy[i] = fmaf(-sz.g[i], t12, wdz * c21);                                                         \
      Sz[i] = fmaf(-sz.g[i], t22, wdz * c22);                                                         \
    }                                                                                                 \
    _Pragma("unroll") for (int kz = 0; kz < 3; kz++) {                                                \
      _Pragma("unroll") for (int jy = 0; jy < 3; jy++) {                                              \
        const float wyz = sy.w[jy] * sz.w[kz];                                                        \
        const float Px = fmaf(sz.w[kz], Rx[jy], sy.w[jy] * Sx[kz]);                                   \
        const float Py = fmaf(sz.w[kz], Ry[jy], sy.w[jy] * Sy[kz]);                                   \
        const float Pz = fmaf(sz.w[kz], Rz[jy], sy.w[jy] * Sz[kz]);                                   \
        _Pragma("unroll") for (int ix = 0; ix < 3; ix++) {                                            \
          const int o = (kz * 3 + jy) * 3 + ix;                                                       \
          float wx = sx.w[ix], wn = wyz;                                                              \
          if (MASK) {                                                                                 \
            const bool live = (sx.w[ix] * sy.w[jy]) * sz.w[kz] >= 1e-12f;                             \
            wx = live ? wx : 0.f;                                                                     \
            wn = live ? wn : 0.f;                                                                     \
          }                                                                                           \
          am[o] = fmaf(mwx[ix], wn, am[o]);                                                           \
          ax[o] = fmaf(wn, Qx[ix], fmaf(wx, Px, ax[o]));                                              \
          ay[o] = fmaf(wn, Qy[ix], fmaf(wx, Py, ay[o]));                                              \
          az[o] = fmaf(wn, Qz[ix], fmaf(wx, Pz, az[o]));                                              \
        }                                                                                             \
      }                                                                                               \
    }                                                                                                 \
  }

template <bool BATCH>
__global__ __launch_bounds__(kP2GThreads, 2) void k_p2g(SubParams p, const float4* __restrict__ rec,
                                                     const nb_material* __restrict__ mats,
                                                     const uint32_t* __restrict__ cell_count,
                                                     const uint32_t* __restrict__ cell_end,
                                                     const uint32_t* __restrict__ sorted,
                                                     const uint32_t* __restrict__ act_tiles,
                                                     const uint32_t* __restrict__ tile_order,
                                                     const uint32_t* __restrict__ tile_start,
                                                     const uint32_t* __restrict__ tile_total,
                                                     const uint32_t* __restrict__ counters,
                                                     float4* __restrict__ partial) {
  __shared__ float s_m[kBoxNodes], s_x[kBoxNodes], s_y[kBoxNodes], s_z[kBoxNodes];
  __shared__ uint16_t s_item[kTileCells * kMaxChunks];  // cell | chunk << 9
  __shared__ uint32_t s_layers;
  __shared__ uint32_t s_cnt[kTileCells], s_end[kTileCells];
  __shared__ uint32_t s_idx[kIdxStage];
  __shared__ uint64_t s_scan64[kP2GThreads / 64 + 1];
  __shared__ float s_rho0[BATCH ? kBatchMaxMaterials : NB_MAX_MATERIALS];
  for (uint32_t i = threadIdx.x; i < p.material_count; i += kP2GThreads) s_rho0[i] = mats[i].rho0;
  const uint32_t na = counters[CNT_ACTIVE];
  const bool apic = p.flip_ratio < -0.5f;
  const uint32_t mlast = p.material_count - 1u;

  for (uint32_t a0 = blockIdx.x; a0 < na; a0 += gridDim.x) {
    // the tile of place a0 in the largest-first order (its own index if the
    // order were ever out of range)
    const uint32_t ao = tile_order[a0], a = ao < na ? ao : a0;
    const uint32_t T = act_tiles[a];
    const SceneGeo g = tile_geo<BATCH>(p, T);
    const float D_inv = 4.f / (g.h * g.h);
    const float inv_h = g.inv_h;
    int tx, ty, tz;
    tile_coords(g, T, tx, ty, tz);
    const size_t tbase = size_t(T) << kTileBits;
    for (int e = threadIdx.x; e < kBoxNodes; e += kP2GThreads) {
      s_m[e] = 0.f; s_x[e] = 0.f; s_y[e] = 0.f; s_z[e] = 0.f;
    }
    // Work items of the occupied cells (thread t looks at cells 2t, 2t+1),
    // grouped by particle count (largest first, cell order within a group) so
    // that the lanes of a wave walk similar lengths: a stable counting sort
    // over kGroups groups, whose per-group counts (< 4096) are packed 12 bits
    // each into one 64-bit block scan.
    if (threadIdx.x == 0) s_layers = 0u;
    const uint32_t ca = 2u * threadIdx.x, cb = ca + 1u;
    const uint32_t na_ = cell_count[tbase + ca], nb_ = cell_count[tbase + cb];
    const uint32_t ea = cell_end[tbase + ca], eb = cell_end[tbase + cb];
    // the tile's particles are one contiguous range of the bin order: their
    // indices are staged in LDS with coalesced loads (a tile beyond the stage
    // reads the bin order directly), loaded here so that they are in flight
    // during the scan below and stored to LDS after it
    const uint32_t lo = tile_start[a], tile_n = tile_total[a];
    const bool staged = tile_n <= kIdxStage;
    uint32_t stg[kIdxStage / kP2GThreads];
#pragma unroll
    for (uint32_t r = 0; r < kIdxStage / kP2GThreads; r++) {
      const uint32_t e = threadIdx.x + r * kP2GThreads;
      stg[r] = (staged && e < tile_n) ? sorted[lo + e] : 0u;
    }
    s_cnt[ca] = na_;
    s_cnt[cb] = nb_;
    s_end[ca] = ea;
    s_end[cb] = eb;
    const uint32_t qa = p2g_chunks(na_, p.p2g_item), qb = p2g_chunks(nb_, p.p2g_item);
    const uint32_t ga = p2g_group(na_, qa), gb = p2g_group(nb_, qb);
    const uint64_t mine = (uint64_t(qa) << (12 * ga)) + (uint64_t(qb) << (12 * gb));
    uint64_t tot64;
    const uint64_t pre = block_exclusive_scan<kP2GThreads>(mine, s_scan64, tot64);
    if (max(qa, qb) > 1u) atomicMax(&s_layers, max(qa, qb) - 1u);
    uint32_t gbase[kGroups];
    {
      uint32_t acc = 0;
#pragma unroll
      for (int g = 0; g < int(kGroups); g++) {
        gbase[g] = acc;
        acc += uint32_t(tot64 >> (12 * g)) & 4095u;
      }
    }
    uint32_t pa = 0, pb = 0;
#pragma unroll
    for (int g = 0; g < int(kGroups); g++) {
      if (ga == uint32_t(g)) pa = gbase[g] + (uint32_t(pre >> (12 * g)) & 4095u);
      if (gb == uint32_t(g)) pb = gbase[g] + (uint32_t(pre >> (12 * g)) & 4095u) + (ga == gb ? qa : 0u);
    }
    for (uint32_t q = 0; q < qa; q++) s_item[pa + q] = uint16_t(ca | (q << 9));
    for (uint32_t q = 0; q < qb; q++) s_item[pb + q] = uint16_t(cb | (q << 9));
#pragma unroll
    for (uint32_t r = 0; r < kIdxStage / kP2GThreads; r++) {
      const uint32_t e = threadIdx.x + r * kP2GThreads;
      if (staged && e < tile_n) s_idx[e] = stg[r];
    }
    __syncthreads();
    uint32_t n_items = 0;
#pragma unroll
    for (int g = 0; g < int(kGroups); g++) n_items += uint32_t(tot64 >> (12 * g)) & 4095u;
    const uint32_t n_layers = s_layers + 1u;
    auto idx_of = [&](uint32_t k) -> uint32_t { return staged ? s_idx[k - lo] : sorted[k]; };

    for (uint32_t base = 0; base < n_items; base += kP2GThreads) {
      const uint32_t j = base + threadIdx.x;
      const bool have = j < n_items;
      const uint32_t it = have ? uint32_t(s_item[j]) : 0u;
      const uint32_t c = it & 511u, chunk = it >> 9;
      const int lx = int(morton_compact(c)), ly = int(morton_compact(c >> 1)), lz = int(morton_compact(c >> 2));
      const int cx = tx * 8 + lx, cy = ty * 8 + ly, cz = tz * 8 + lz;
      uint32_t kbeg = 0, kend = 0;
      if (have) {
        const uint32_t cnt = s_cnt[c], end = s_end[c];
        const uint32_t per = (cnt + p2g_chunks(cnt, p.p2g_item) - 1u) / p2g_chunks(cnt, p.p2g_item);
        kbeg = end - cnt + chunk * per;
        kend = min(kbeg + per, end);
      }
      float am[27], ax[27], ay[27], az[27];
#pragma unroll
      for (int o = 0; o < 27; o++) { am[o] = 0.f; ax[o] = 0.f; ay[o] = 0.f; az[o] = 0.f; }

      // Per particle (shaders/mpm_p2g.comp:40-90): mass m = rho0 Vp, affine
      // C = m (4/h^2) B (APIC only), dt Vp tau from the w lanes
      // (shaders/mpm_common.glsl:104-113), and the stencil of its cell.
      struct Src {
        Axis3 sx, sy, sz;
        float mp, mvx, mvy, mvz;
        float c00, c01, c02, c10, c11, c12, c20, c21, c22;
        float t00, t11, t22, t01, t02, t12;
      };
      // (the stencil's quotient (x - wmin)/h is taken as a product with 1/h:
      // the cell is the bin's, only the weights use it)
      // (a particle of the cell takes the fixed-piece form; one the bin
      // placed there from outside the grid, the general one).  Returns
      // whether the particle may have a node the reference skips, w =
      // (wx wy) wz < 1e-12 (shaders/mpm_p2g.comp:60-61): the smallest positive
      // per-axis weights bound every w from below (1.0001e-12 covers the
      // rounding of the product order).  In the fixed-piece form the middle
      // weight is >= 0.5 and the first one > 0, so an axis's smallest
      // positive weight is the smaller outer one (the last counting when > 0).
      auto stencil = [&](const float4& pp, Src& q) -> bool {
        const float fx = (pp.x - g.wmin[0]) * inv_h, fy = (pp.y - g.wmin[1]) * inv_h;
        const float fz = (pp.z - g.wmin[2]) * inv_h;
        bool ok = axis_stencil_cell(fx, pp.x, g.wmin[0], g.h, inv_h, cx, q.sx);
        ok &= axis_stencil_cell(fy, pp.y, g.wmin[1], g.h, inv_h, cy, q.sy);
        ok &= axis_stencil_cell(fz, pp.z, g.wmin[2], g.h, inv_h, cz, q.sz);
        if (ok) {
          auto mpos = [](const Axis3& a) { return a.w[2] > 0.f ? fminf(a.w[0], a.w[2]) : a.w[0]; };
          return (mpos(q.sx) * mpos(q.sy)) * mpos(q.sz) < 1.0001e-12f;
        }
        axis_stencil_f(fx, pp.x, g.wmin[0], g.h, inv_h, cx - 1, q.sx);
        axis_stencil_f(fy, pp.y, g.wmin[1], g.h, inv_h, cy - 1, q.sy);
        axis_stencil_f(fz, pp.z, g.wmin[2], g.h, inv_h, cz - 1, q.sz);
        return (min_positive3(q.sx.w) * min_positive3(q.sy.w)) * min_positive3(q.sz.w) < 1.0001e-12f;
      };
      // keep = 0 scales the particle's contribution to exactly zero
      auto setup = [&](const float4& pp, const float4& vv, const float4& b0, const float4& b1, const float4& b2,
                       const float4& tt, float keep, Src& q) {
        uint32_t mid = __float_as_uint(vv.w);
        mid = mid > mlast ? mlast : mid;
        q.mp = s_rho0[mid] * pp.w * keep;
        const float sc = apic ? q.mp * D_inv : 0.f;
        q.c00 = sc * b0.x; q.c01 = sc * b0.y; q.c02 = sc * b0.z;
        q.c10 = sc * b1.x; q.c11 = sc * b1.y; q.c12 = sc * b1.z;
        q.c20 = sc * b2.x; q.c21 = sc * b2.y; q.c22 = sc * b2.z;
        const float st = p.dt * pp.w * keep;
        q.t00 = st * tt.x; q.t11 = st * tt.y; q.t22 = st * tt.z;
        q.t01 = st * b0.w; q.t02 = st * b1.w; q.t12 = st * b2.w;
        q.mvx = q.mp * vv.x; q.mvy = q.mp * vv.y; q.mvz = q.mp * vv.z;
      };

      // Fast loop.  Node contributions factor as
      //   m w = (m wx) wyz,  mom = wx P(y,z) + wyz Q(x)
      // with P = wyz (m v + C (0, dy, dz)) - tau (0, gy wz, wy gz) and
      // Q = wx dx C e_x - gx tau e_x: 7 FMAs per node (explicit FMA chains;
      // strict evaluation of acc + (a b + c d) costs a multiply, an FMA and an
      // add).  P itself splits into per-axis terms, P = wz R(y) + wy S(z) with
      // R = wy (m v + dy C e_y) - gy tau e_y and S = wz dz C e_z - gz tau e_z,
      // so each of the 9 (y, z) pairs costs a multiply and an FMA per
      // component.  A particle with a sub-cutoff node is scaled to zero here and
      // added by the exact loop below, which visits only those particles:
      // the first two by their offsets from kbeg (16 bits each in tiny_at,
      // 0xFFFF = none), all of the item's particles when there are more (or
      // an offset does not fit).  The next particle's record is loaded while
      // this one is scattered; the walk's last iteration loads the first such
      // particle's record instead (pref), so the exact loop rarely waits.
      uint32_t tiny_at = 0xFFFFFFFFu;
      bool tiny_all = false, pref = false;
      float4 npp, nvv, nb0, nb1, nb2, ntt;
      auto load_rec = [&](uint32_t q) {
        const float4* r = rec + size_t(q) * kRec;
        npp = r[R_P]; nvv = r[R_V];
        nb0 = r[R_B0]; nb1 = r[R_B1]; nb2 = r[R_B2];
        ntt = r[R_T];
      };
      if (kbeg < kend) load_rec(idx_of(kbeg));
      for (uint32_t k = kbeg; k < kend; k++) {
        const float4 pp = npp, vv = nvv, b0 = nb0, b1 = nb1, b2 = nb2, tt = ntt;
        if (k + 1 < kend) {
          load_rec(idx_of(k + 1));
        } else if ((tiny_at & 0xFFFFu) != 0xFFFFu) {
          load_rec(idx_of(kbeg + (tiny_at & 0xFFFFu)));
          pref = true;
        }
        Src q;
        const bool tiny = stencil(pp, q);
        if (tiny) {
          const uint32_t off = k - kbeg;
          if (off >= 0xFFFFu || (tiny_at >> 16) != 0xFFFFu) {
            tiny_all = true;
          } else if ((tiny_at & 0xFFFFu) == 0xFFFFu) {
            tiny_at = 0xFFFF0000u | off;
          } else {
            tiny_at = (tiny_at & 0xFFFFu) | (off << 16);
          }
        }
        setup(pp, vv, b0, b1, b2, tt, tiny ? 0.f : 1.f, q);
        NB_P2G_CHAINS(false)
      }
      // Exact loop: the particles skipped above, the same chains with the
      // cutoff's node mask, in bin order (the order of the adds is the one a
      // walk over the whole item would give).
      const uint32_t n_exact = tiny_all ? kend - kbeg
                                        : ((tiny_at & 0xFFFFu) != 0xFFFFu ? 1u : 0u) + ((tiny_at >> 16) != 0xFFFFu ? 1u : 0u);
      if (__any(n_exact != 0u)) {
        for (uint32_t s = 0; s < n_exact; s++) {
          const uint32_t k = kbeg + (tiny_all ? s : (s == 0u ? (tiny_at & 0xFFFFu) : (tiny_at >> 16)));
          if (!(pref && !tiny_all && s == 0u)) load_rec(idx_of(k));
          Src q;
          const float4 pp = npp;
          if (!stencil(pp, q)) continue;
          setup(pp, nvv, nb0, nb1, nb2, ntt, 1.f, q);
          NB_P2G_CHAINS(true)
        }
      }
      for (uint32_t layer = 0; layer < n_layers; layer++) {
        const bool mine = have && chunk == layer;
#pragma unroll
        for (int kz = 0; kz < 3; kz++)
#pragma unroll
          for (int jy = 0; jy < 3; jy++)
#pragma unroll
            for (int ix = 0; ix < 3; ix++) {
              const int o = (kz * 3 + jy) * 3 + ix;
              const int e = ((lz + kz) * kBox + (ly + jy)) * kBox + (lx + ix);
              __syncthreads();
              if (mine) {
                s_m[e] += am[o];
                s_x[e] += ax[o];
                s_y[e] += ay[o];
                s_z[e] += az[o];
              }
            }
      }
      __syncthreads();
    }
    float4* out = partial + size_t(a) * kBoxNodes;
    for (int e = threadIdx.x; e < kBoxNodes; e += kP2GThreads) out[e] = make_float4(s_x[e], s_y[e], s_z[e], s_m[e]);
    __syncthreads();
  }
}

#undef NB_P2G_CHAINS

// ---------------------------------------------------------------------------
// GRID UPDATE
// ---------------------------------------------------------------------------
// Mass and momentum of node (local (lx, ly, lz) in tile (tx, ty, tz) of scene
// g): the partial boxes of the occupied tiles whose 10^3 box contains it,
// summed in a fixed order.  False when no occupied tile's box contains it
// (then it lies in no particle's stencil).
__device__ __forceinline__ bool sum_partials(const SceneGeo& g, int tx, int ty, int tz, int lx, int ly, int lz,
                                             const int32_t* __restrict__ tile_act,
                                             const float4* __restrict__ partial, float& mx, float& my, float& mz,
                                             float& mass) {
  const int td = int(g.tile_dim);
  mx = my = mz = mass = 0.f;
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
        const int32_t a = tile_act[tile_id(g, sx, sy, sz)];
        if (a < 0) continue;
        const int bx = lx - 8 * dx + 1, by = ly - 8 * dy + 1, bz = lz - 8 * dz + 1;
        const float4 q = partial[size_t(a) * kBoxNodes + (bz * kBox + by) * kBox + bx];
        mx += q.x; my += q.y; mz += q.z; mass += q.w;
        covered = true;
      }
    }
  }
  return covered;
}

// The update of a node with mass >= 1e-10 at grid coordinates (nx, ny, nz),
// in registers and in this order: the 15 m/s cap, the saved (pre-force)
// velocity, forces, walls, analytic colliders, the dense-SDF BC
// (shaders/mpm_grid_update.comp:30-110, shaders/mpm_nanovdb_bc.comp:23-54).
// v enters as momentum / mass; o receives the saved velocity.
template <bool NOISE>
__device__ __forceinline__ void node_update(const SubParams& p, const SceneGeo& g, int nx, int ny, int nz,
                                            float& vx, float& vy, float& vz, float& ox, float& oy, float& oz,
                                            const float* __restrict__ sdf, const nb_collider* __restrict__ cols,
                                            const nb_force* __restrict__ forces) {
  {
    const float sp = len3(vx, vy, vz);
    const float vmax = 15.f;
    if (sp > vmax) {
      const float k = vmax / sp;
      vx *= k; vy *= k; vz *= k;
    }
  }
  ox = vx; oy = vy; oz = vz;
  const float px = (float(nx) + 0.5f) * g.h + g.wmin[0];
  const float py = (float(ny) + 0.5f) * g.h + g.wmin[1];
  const float pz = (float(nz) + 0.5f) * g.h + g.wmin[2];
  apply_forces<NOISE>(vx, vy, vz, px, py, pz, forces, p.force_count, p.dt);
  // separating walls on the two outer node layers (shaders/mpm_grid_update.comp:54-60)
  const int Nx = g.grid_res[0] - 1, Ny = g.grid_res[1] - 1, Nz = g.grid_res[2] - 1;
  if (nx <= 1 && vx < 0.f) vx = 0.f;
  if (nx >= Nx - 1 && vx > 0.f) vx = 0.f;
  if (ny <= 1 && vy < 0.f) vy = 0.f;
  if (ny >= Ny - 1 && vy > 0.f) vy = 0.f;
  if (nz <= 1 && vz < 0.f) vz = 0.f;
  if (nz >= Nz - 1 && vz > 0.f) vz = 0.f;
  for (uint32_t ci = 0; ci < p.collider_count; ci++) {
    const nb_collider c = cols[ci];
    float cnx, cny, cnz;
    const float sd = collider_sdf(c, px, py, pz, cnx, cny, cnz);
    if (sd <= 0.f) collider_bc(vx, vy, vz, cnx, cny, cnz, c);
  }
  // dense-SDF BC (shaders/mpm_nanovdb_bc.comp:23-54)
  if (sdf) {
    const float sd = sample_sdf(sdf, g, nx, ny, nz);
    if (sd < 0.f) {
      vx = vy = vz = 0.f;
    } else if (sd < g.h * 1.5f) {
      const float gx = sample_sdf(sdf, g, nx + 1, ny, nz) - sample_sdf(sdf, g, nx - 1, ny, nz);
      const float gy = sample_sdf(sdf, g, nx, ny + 1, nz) - sample_sdf(sdf, g, nx, ny - 1, nz);
      const float gz = sample_sdf(sdf, g, nx, ny, nz + 1) - sample_sdf(sdf, g, nx, ny, nz - 1);
      const float gl = len3(gx, gy, gz);
      if (gl > 1e-6f) {
        const float ux = gx / gl, uy = gy / gl, uz = gz / gl;
        const float vn = vx * ux + vy * uy + vz * uz;
        if (vn < 0.f) {
          vx -= vn * ux; vy -= vn * uy; vz -= vn * uz;
        }
      }
    }
  }
}

template <bool FROM_PARTIALS, bool UPDATE, bool BATCH, bool NOISE>
__global__ __launch_bounds__(512) void k_grid(SubParams p, const float4* __restrict__ partial,
                                              const int32_t* __restrict__ tile_act, const uint32_t* __restrict__ tiles,
                                              const uint32_t* __restrict__ count, float4* __restrict__ grid_node,
                                              float4* __restrict__ grid_old, float4* __restrict__ grid_pre,
                                              const float* __restrict__ sdf, const nb_collider* __restrict__ cols,
                                              const nb_force* __restrict__ forces) {
  const uint32_t nb_tiles = *count;
  // thread j holds the tile's j-th node in Morton order (contiguous in HBM)
  const int lx = int(morton_compact(threadIdx.x)), ly = int(morton_compact(threadIdx.x >> 1)),
            lz = int(morton_compact(threadIdx.x >> 2));
  for (uint32_t b = blockIdx.x; b < nb_tiles; b += gridDim.x) {
    const uint32_t N = tiles[b];
    const SceneGeo g = tile_geo<BATCH>(p, N);
    int tx, ty, tz;
    tile_coords(g, N, tx, ty, tz);
    const int nx = tx * 8 + lx, ny = ty * 8 + ly, nz = tz * 8 + lz;
    const size_t nid = (size_t(N) << kTileBits) | threadIdx.x;
    float mx, my, mz, mass;
    if (FROM_PARTIALS) {
      // A node outside every occupied tile's box receives nothing and lies in
      // no particle's stencil: it is not written (readers treat it as empty).
      if (!sum_partials(g, tx, ty, tz, lx, ly, lz, tile_act, partial, mx, my, mz, mass)) continue;
    } else {
      const float4 g4 = grid_node[nid];
      mx = g4.x; my = g4.y; mz = g4.z; mass = g4.w;
    }
    if (grid_pre)
      grid_pre[nid] = mass < 1e-10f ? make_float4(0.f, 0.f, 0.f, mass) : make_float4(mx / mass, my / mass, mz / mass, mass);
    if (!UPDATE || mass < 1e-10f) {
      // massless nodes keep their momentum (shaders/mpm_grid_update.comp:109-110)
      grid_node[nid] = make_float4(mx, my, mz, mass);
      continue;
    }
    float vx = mx / mass, vy = my / mass, vz = mz / mass, ox, oy, oz;
    node_update<NOISE>(p, g, nx, ny, nz, vx, vy, vz, ox, oy, oz, sdf, cols, forces);
    if (p.write_old) grid_old[nid] = make_float4(ox, oy, oz, 0.f);
    grid_node[nid] = make_float4(vx, vy, vz, mass);
  }
}

// ---------------------------------------------------------------------------
// G2P
// ---------------------------------------------------------------------------
// What the 27-node gather gives a particle (shaders/mpm_g2p.comp:80-113): the
// PIC velocity, FLIP's saved-velocity interpolant, the velocity gradient
// vg.m[col = gradient axis][row = component] and APIC's B.
struct G2PGather {
  float vpx, vpy, vpz, vox, voy, voz;
  m3 vg, Bn;
};

// The gather.  node(ix, iy, iz, v, vo) returns whether the stencil node
// (base + (ix, iy, iz)) takes part (in the grid, mass >= 1e-10) and its
// velocity in v (and FLIP's saved velocity in vo), both zero when it does not.
// The 27 nodes are contracted axis by axis (x, then y, then z) in three rolled
// z slabs, so only one slab's node loads are live.  When some node of the
// particle has 0 < w < 1e-12 (tiny), the reference's weight cut
// (shaders/mpm_g2p.comp:93-101) changes the sums, and they are taken node by
// node with the cut instead; each lane decides alone, so each particle's
// result depends on its own data only.
template <bool FLIP, class NodeFn>
__device__ __forceinline__ void g2p_gather(const Axis3 (&ax3)[3], bool tiny, NodeFn node, G2PGather& o) {
  float vpx = 0.f, vpy = 0.f, vpz = 0.f, vox = 0.f, voy = 0.f, voz = 0.f;
  m3 vg = m3_zero(), Bn = m3_zero();
  const Axis3 &X = ax3[0], &Y = ax3[1];
  // the slab's z stencil values sit in position 0 and rotate at the end of
  // each iteration
  float zw[3] = {ax3[2].w[0], ax3[2].w[1], ax3[2].w[2]};
  float zg[3] = {ax3[2].g[0], ax3[2].g[1], ax3[2].g[2]};
  float zd[3] = {ax3[2].dx[0], ax3[2].dx[1], ax3[2].dx[2]};
  if (!tiny) {
    float wdx[3], wdy[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      wdx[i] = X.w[i] * X.dx[i];
      wdy[i] = Y.w[i] * Y.dx[i];
    }
#pragma unroll 1
    for (int iz = 0; iz < 3; iz++) {
      // sums start from their first term (no zero-initialised accumulators)
      float Y00[3], Y01[3], Y02[3], Y1[3], Y2[3], YO[3];
#pragma unroll
      for (int iy = 0; iy < 3; iy++) {
        float gv[3][3], go[3][3];
#pragma unroll
        for (int ix = 0; ix < 3; ix++) (void)node(ix, iy, iz, gv[ix], go[ix]);
        float X0[3], X1[3], X2[3], XO[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          X0[c] = X.w[0] * gv[0][c] + X.w[1] * gv[1][c] + X.w[2] * gv[2][c];
          X1[c] = X.g[0] * gv[0][c] + X.g[1] * gv[1][c] + X.g[2] * gv[2][c];
          X2[c] = wdx[0] * gv[0][c] + wdx[1] * gv[1][c] + wdx[2] * gv[2][c];
          if (FLIP) XO[c] = X.w[0] * go[0][c] + X.w[1] * go[1][c] + X.w[2] * go[2][c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
          if (iy == 0) {
            Y00[c] = Y.w[0] * X0[c];
            Y01[c] = Y.g[0] * X0[c];
            Y02[c] = wdy[0] * X0[c];
            Y1[c] = Y.w[0] * X1[c];
            Y2[c] = Y.w[0] * X2[c];
            if (FLIP) YO[c] = Y.w[0] * XO[c];
          } else {
            Y00[c] += Y.w[iy] * X0[c];
            Y01[c] += Y.g[iy] * X0[c];
            Y02[c] += wdy[iy] * X0[c];
            Y1[c] += Y.w[iy] * X1[c];
            Y2[c] += Y.w[iy] * X2[c];
            if (FLIP) YO[c] += Y.w[iy] * XO[c];
          }
        }
      }
      const float wz = zw[0], gz = zg[0], wdz = zw[0] * zd[0];
      float* vp3[3] = {&vpx, &vpy, &vpz};
      float* vo3[3] = {&vox, &voy, &voz};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        *vp3[c] += wz * Y00[c];
        if (FLIP) *vo3[c] += wz * YO[c];
        // vg = sum outer(v_i, grad w_i): vg.m[col = grad axis][row = velocity component]
        vg.m[0][c] += wz * Y1[c];
        vg.m[1][c] += wz * Y01[c];
        vg.m[2][c] += gz * Y00[c];
        // Bn = sum outer(w_i v_i, x_i - x_p)
        Bn.m[0][c] += wz * Y2[c];
        Bn.m[1][c] += wz * Y02[c];
        Bn.m[2][c] += wdz * Y00[c];
      }
      zw[0] = zw[1]; zw[1] = zw[2];
      zg[0] = zg[1]; zg[1] = zg[2];
      zd[0] = zd[1]; zd[1] = zd[2];
    }
  } else {
#pragma unroll 1
    for (int iz = 0; iz < 3; iz++) {
      const float wz = zw[0], gz = zg[0], dxz = zd[0];
#pragma unroll
      for (int iy = 0; iy < 3; iy++) {
#pragma unroll
        for (int ix = 0; ix < 3; ix++) {
          float gvv[3], gov[3];
          const bool in = node(ix, iy, iz, gvv, gov);
          const float wx = X.w[ix], wy = Y.w[iy];
          const float w0 = wx * wy * wz;
          const bool live = in && w0 >= 1e-12f;
          const float gx = live ? gvv[0] : 0.f, gy = live ? gvv[1] : 0.f, gzv = live ? gvv[2] : 0.f;
          const float w = live ? w0 : 0.f;
          const float dwx = live ? X.g[ix] * wy * wz : 0.f;
          const float dwy = live ? wx * Y.g[iy] * wz : 0.f;
          const float dwz = live ? wx * wy * gz : 0.f;
          const float dxx = X.dx[ix], dxy = Y.dx[iy];
          vpx += w * gx; vpy += w * gy; vpz += w * gzv;
          // outerProduct(vi, dw)[j][i] = vi[i]*dw[j]
          vg.m[0][0] += gx * dwx; vg.m[0][1] += gy * dwx; vg.m[0][2] += gzv * dwx;
          vg.m[1][0] += gx * dwy; vg.m[1][1] += gy * dwy; vg.m[1][2] += gzv * dwy;
          vg.m[2][0] += gx * dwz; vg.m[2][1] += gy * dwz; vg.m[2][2] += gzv * dwz;
          const float wgx = w * gx, wgy = w * gy, wgz = w * gzv;
          Bn.m[0][0] += wgx * dxx; Bn.m[0][1] += wgy * dxx; Bn.m[0][2] += wgz * dxx;
          Bn.m[1][0] += wgx * dxy; Bn.m[1][1] += wgy * dxy; Bn.m[1][2] += wgz * dxy;
          Bn.m[2][0] += wgx * dxz; Bn.m[2][1] += wgy * dxz; Bn.m[2][2] += wgz * dxz;
          if (FLIP) {
            vox += w * (live ? gov[0] : 0.f);
            voy += w * (live ? gov[1] : 0.f);
            voz += w * (live ? gov[2] : 0.f);
          }
        }
      }
      zw[0] = zw[1]; zw[1] = zw[2];
      zg[0] = zg[1]; zg[1] = zg[2];
      zd[0] = zd[1]; zd[1] = zd[2];
    }
  }
  o.vpx = vpx; o.vpy = vpy; o.vpz = vpz;
  o.vox = vox; o.voy = voy; o.voz = voz;
  o.vg = vg;
  o.Bn = Bn;
}

// Stencil of particle position x: per axis the first stencil node base =
// floor((x - wmin)/h) - 1 (the quotient clamped to a range that keeps every
// integer below in bounds without changing which nodes are valid) and the
// weights; tiny: some node has 0 < w < 1e-12 (the smallest positive
// per-axis weights bound every w from below).
__device__ __forceinline__ bool g2p_stencil(const SceneGeo& g, float xp, float yp, float zp, int (&base)[3],
                                            Axis3 (&ax3)[3]) {
  const float xs[3] = {xp, yp, zp};
  float fq[3];
  bool ok = true;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    fq[a] = cell_coord(g, xs[a], a);
    const float f = fminf(fmaxf(fq[a], -4.f), float(g.grid_res[a]) + 4.f);
    base[a] = int(floorf(f)) - 1;
    ok &= axis_stencil_cell(fq[a], xs[a], g.wmin[a], g.h, g.inv_h, base[a] + 1, ax3[a]);
  }
  if (!ok) {  // the first cell layer, or outside the grid
#pragma unroll
    for (int a = 0; a < 3; a++) axis_stencil_f(fq[a], xs[a], g.wmin[a], g.h, g.inv_h, base[a], ax3[a]);
  }
  return (min_positive3(ax3[0].w) * min_positive3(ax3[1].w)) * min_positive3(ax3[2].w) < 1e-12f;
}

// The fallback's counters (NB_SVD_FB_*): one set of atomics per wave, from
// ballots over the lanes that reach the call together.
__device__ __forceinline__ void fb_count_stats(unsigned long long* st, bool dp_lane, bool fb_proj, bool fb_stress) {
  const uint64_t b_dp = __ballot(dp_lane), b_pr = __ballot(fb_proj), b_st = __ballot(fb_stress);
  if ((threadIdx.x & 63u) == uint32_t(__ffsll((unsigned long long)__ballot(true)) - 1)) {
    if (b_dp) {
      atomicAdd(&st[NB_SVD_FB_DP_LANES], (unsigned long long)__popcll(b_dp));
      atomicAdd(&st[NB_SVD_FB_DP_WAVES], 1ull);
    }
    if (b_pr) atomicAdd(&st[NB_SVD_FB_PROJ_LANES], (unsigned long long)__popcll(b_pr));
    if (b_st) atomicAdd(&st[NB_SVD_FB_STRESS_LANES], (unsigned long long)__popcll(b_st));
    if (b_pr | b_st) atomicAdd(&st[NB_SVD_FB_FB_WAVES], 1ull);
  }
}

// The FP64 fallback's computation, from an FP32 trial F: the return map
// (SVD, projection, recomposition) and the Hencky stress of its result, in
// FP64 and rounded once.  All of it stays FP64: the stress SVD of a
// recomposed F' with a close pair is as sensitive to F' as the first SVD is
// to F, so an F' rounded to FP32 (or built with FP32 exp) would decide its
// singular vectors again.
struct FTau {
  m3 F, tau;
};
__device__ __forceinline__ FTau plastic_fp64(const m3& F, bool dp, float mu_f, float lam_f, float q_max, float M,
                                             float q_c, bool cs) {
  const double mu = mu_f, lam = lam_f;
  const PlasticOutT<double> pd =
      project_plastic(m3_convert<double>(F), dp, mu, lam, double(q_max), double(M), double(q_c), cs);
  FTau r;
  r.F = m3_round(pd.F);
  r.tau = m3_round((cs || !pd.recomposed) ? hencky_from_log(pd.U, pd.eps[0], pd.eps[1], pd.eps[2], mu, lam)
                                          : hencky_stress(pd.F, mu, lam, cs));
  return r;
}

// The FP64 fallback's queue: kFbSlots slots of p.fb_region = ceil(n /
// kFbSlots) entries; a particle at place k of the G2P order goes to slot
// k / fb_region, so a slot never overflows and its atomics come from a
// fraction of the waves.  An entry is kFbEntry float4: q, mode, then F_old
// and the velocity gradient, 9 floats each in m[0][0], m[0][1], ..., m[2][2]
// order (m[col][row] storage).  mode 1: return map and stress from the trial
// F = (I + dt grad v) F_old; 2: stress of the record's F; 0: dropped.  Slot
// counters sit kFbCtlStride words apart, the k_g2p_fb block count after them.
constexpr uint32_t kFbNone = 0xFFFFFFFFu;

// queue the particle if it takes the fallback (all lanes of the wave call
// this together); returns its entry or kFbNone
__device__ __forceinline__ uint32_t queue_fb(const SubParams& p, uint32_t k, uint32_t q, const float4* r,
                                             const m3& vg, bool fb_proj, bool fb_stress) {
  const bool fb = fb_proj || fb_stress;
  uint64_t pending = __ballot(fb);
  if (!pending) return kFbNone;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t slot = k / p.fb_region;
  uint32_t at = kFbNone;
  // the wave's queued lanes grouped by slot (one group unless the wave
  // straddles a slot boundary); one atomic per group
  while (pending) {
    const uint32_t leader = uint32_t(__ffsll((unsigned long long)pending) - 1);
    const uint32_t ls = uint32_t(__shfl(int(slot), int(leader)));
    const uint64_t peers = __ballot(fb && slot == ls);
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(&p.fb_ctl[ls * kFbCtlStride], uint32_t(__popcll(peers)));
    base = uint32_t(__shfl(int(base), int(leader)));
    if (fb && slot == ls) at = ls * p.fb_region + base + uint32_t(__popcll(peers & ((1ull << lane) - 1ull)));
    pending &= ~peers;
  }
  if (fb) {
    float4* e = p.fb_queue + size_t(at) * kFbEntry;
    if (fb_proj) {
      // F_old is still the record's (the record is written after this)
      const float4 rt = r[R_T], fa = r[R_FA], fb4 = r[R_FB];
      e[0] = make_float4(__uint_as_float(q), __uint_as_float(1u), rt.w, fa.x);
      e[1] = make_float4(fa.y, fa.z, fa.w, fb4.x);
      e[2] = make_float4(fb4.y, fb4.z, fb4.w, vg.m[0][0]);
      e[3] = make_float4(vg.m[0][1], vg.m[0][2], vg.m[1][0], vg.m[1][1]);
      e[4] = make_float4(vg.m[1][2], vg.m[2][0], vg.m[2][1], vg.m[2][2]);
    } else {
      e[0] = make_float4(__uint_as_float(q), __uint_as_float(2u), 0.f, 0.f);
    }
  }
  return at;
}

// Everything of G2P after the gather (shaders/mpm_g2p.comp:115-272): F
// update, constitutive update (fluid; elastic; Von Mises / Drucker-Prager
// family return map with the FP64 fallback for near-degenerate F), mud
// viscosity, Jacobian reset, PIC/APIC/FLIP velocity, advection, SDF push-out,
// world clamp; writes the record and the next substep's cell key.
// A particle whose update the FP64 fallback takes over finishes here in FP32
// with its stress set to the mud viscosity term alone (zero for the other
// models) and is queued for k_g2p_fb (queue_fb), which replaces F and tau
// with the FP64 results; if the Jacobian reset or the SDF push-out then sets
// F = I, tau = 0 (both decided on the FP32 F), its entry is cancelled.  The
// FP64 code thus stays out of the kernels that run every particle (a call
// site alone costs such a kernel registers and time even when no lane takes
// it).  k: the particle's place in the G2P order, which picks its queue slot.
template <bool FLIP, bool BATCH>
__device__ __forceinline__ void g2p_particle(const SubParams& p, uint32_t k, uint32_t q, const float4* r, float4* ro,
                                             uint32_t* __restrict__ keys, const nb_material& mat,
                                             const G2PGather& gth, const float4 pp, const float4 vel,
                                             const float* __restrict__ sdf) {
  const float xp = pp.x, yp = pp.y, zp = pp.z, Vp = pp.w;
  constexpr bool flip = FLIP;  // flip_ratio > 0.01
  float vpx = gth.vpx, vpy = gth.vpy, vpz = gth.vpz;
  m3 Bn = gth.Bn;
  const float4 rt = r[R_T], fa = r[R_FA], fb = r[R_FB];
  m3 F;
  F.m[0][0] = rt.w; F.m[0][1] = fa.x; F.m[0][2] = fa.y;
  F.m[1][0] = fa.z; F.m[1][1] = fa.w; F.m[1][2] = fb.x;
  F.m[2][0] = fb.y; F.m[2][1] = fb.z; F.m[2][2] = fb.w;
  m3 L;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int rr = 0; rr < 3; rr++) L.m[c][rr] = (c == rr ? 1.f : 0.f) + p.dt * gth.vg.m[c][rr];
  m3 Fn = m3_mul(L, F);

  m3 tau;
  const bool cs = p.corrected_svd != 0u;
  bool dp_lane = false, fb_proj = false, fb_stress = false;  // for the fallback's counters
  if (mat.model == NB_MAT_FLUID) {
    float J = fmaxf(m3_det(Fn), 1e-4f);
    const float cr = powf(J, 1.f / 3.f);
    Fn = m3_diag(cr, cr, cr);
    const float pr = mat.bulkK * (J - 1.f);
    tau = m3_diag(pr, pr, pr);
  } else {
    // elastic: Hencky stress of F; Von Mises / Drucker-Prager (also granular
    // powder and viscoplastic mud): return map, then Hencky stress of the
    // projected F.  In corrected mode the return map's SVD is the true SVD of
    // the projected F, so its factors give the stress without a second SVD.
    // In either mode a Von Mises particle inside the yield surface keeps F
    // unchanged, so the second SVD would repeat the first on the same input.
    const bool dp = mat.model == NB_MAT_DRUCKER_PRAGER || mat.model == NB_MAT_GRANULAR_POWDER ||
                    mat.model == NB_MAT_VISCOPLASTIC_MUD;
    const bool plastic = dp || mat.model == NB_MAT_VON_MISES;
    const bool fb_on = dp && p.svd_fp64 != 0u;
    dp_lane = dp;
    bool have_tau = false;
    if (plastic) {
      const PlasticOut po = project_plastic(Fn, dp, mat.mu, mat.lambda, mat.q_max, mat.M_friction, mat.q_cohesion, cs);
      Fn = po.F;
      if (cs || !po.recomposed) {
        tau = hencky_from_log(po.U, po.eps[0], po.eps[1], po.eps[2], mat.mu, mat.lambda);
        have_tau = true;
      }
      if (fb_on && fb_rounding_decides(po.sig, p.svd_gap_rel, p.svd_zero_rel, cs, po.recomposed, Fn, po.eps,
                                       have_tau, tau, po.eps, mat.mu, mat.lambda)) {
        // rounding decides the trial F's singular vectors and the result
        // depends on them: k_g2p_fb redoes the return map and the stress in
        // FP64 from the same FP32 trial F
        have_tau = true;
        fb_proj = true;
      }
    }
    if (!have_tau) {
      // Hencky stress of F (hencky_stress written out, so that a particle that
      // does not take the fallback runs the same instructions whether it is on
      // or off); the stress SVD of a recomposed Drucker-Prager F is tested on
      // its own singular values
      m3 U, V;
      float sig[3];
      svd3(Fn, U, sig, V, cs);
      const float e[3] = {log_strain(sig[0]), log_strain(sig[1]), log_strain(sig[2])};
      tau = hencky_from_log(U, e[0], e[1], e[2], mat.mu, mat.lambda);
      if (fb_on && fb_rounding_decides(sig, p.svd_gap_rel, p.svd_zero_rel, cs, false, Fn, e, true, tau, e, mat.mu,
                                       mat.lambda))
        fb_stress = true;
    }
    if (fb_proj || fb_stress) tau = m3_zero();  // k_g2p_fb adds the FP64 stress
    if (mat.model == NB_MAT_VISCOPLASTIC_MUD) {
      m3 D;
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int rr = 0; rr < 3; rr++) D.m[c][rr] = 0.5f * (gth.vg.m[c][rr] + gth.vg.m[rr][c]);
      const float trD = D.m[0][0] + D.m[1][1] + D.m[2][2];
      const float k2 = 2.f * mat.viscosity;
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int rr = 0; rr < 3; rr++) tau.m[c][rr] += k2 * (D.m[c][rr] - (c == rr ? trD / 3.f : 0.f));
    }
  }
  // every lane of the wave is here
  if (p.svd_stats) fb_count_stats(p.svd_stats, dp_lane, fb_proj, fb_stress);
  const uint32_t fb_at = p.svd_fp64 ? queue_fb(p, k, q, r, gth.vg, fb_proj, fb_stress) : kFbNone;

  bool jreset = false;
  {
    const float Jd = fabsf(m3_det(Fn));
    if (Jd < 0.15f || Jd > 10.f) {
      Fn = m3_ident();
      tau = m3_zero();
      jreset = true;
    }
  }
  if (p.flip_ratio > -0.5f) Bn = m3_zero();

  float vnx, vny, vnz;
  if (flip) {
    const float rf = p.flip_ratio;
    const float fx = vel.x + (vpx - gth.vox), fy = vel.y + (vpy - gth.voy), fz = vel.z + (vpz - gth.voz);
    vnx = vpx * (1.f - rf) + fx * rf;
    vny = vpy * (1.f - rf) + fy * rf;
    vnz = vpz * (1.f - rf) + fz * rf;
  } else {
    vnx = vpx; vny = vpy; vnz = vpz;
  }
  if (jreset) { vnx *= 0.1f; vny *= 0.1f; vnz *= 0.1f; }
  float xn = xp + p.dt * vnx, yn = yp + p.dt * vny, zn = zp + p.dt * vnz;

  // The geometry is read again for the tail rather than kept live through
  // the SVD (a plain engine re-reads its kernel arguments).
  const SceneGeo ge = particle_geo<BATCH>(p, q);
  // particle-level SDF push-out (shaders/mpm_g2p.comp:220-252)
  bool pushed = false;
  if (sdf) {
    const int gx = ge.grid_res[0] - 1, gy = ge.grid_res[1] - 1, gz = ge.grid_res[2] - 1;
    float fx = fminf(fmaxf(cell_coord(ge, xn, 0), -4.f), float(ge.grid_res[0]) + 4.f);
    float fy = fminf(fmaxf(cell_coord(ge, yn, 1), -4.f), float(ge.grid_res[1]) + 4.f);
    float fz = fminf(fmaxf(cell_coord(ge, zn, 2), -4.f), float(ge.grid_res[2]) + 4.f);
    const int ix = min(max(int(floorf(fx)), 0), gx);
    const int iy = min(max(int(floorf(fy)), 0), gy);
    const int iz = min(max(int(floorf(fz)), 0), gz);
    const float psdf = sdf[morton3(ix, iy, iz)];
    if (psdf < 0.f) {
      const float gnx = sdf[morton3(min(ix + 1, gx), iy, iz)] - sdf[morton3(max(ix - 1, 0), iy, iz)];
      const float gny = sdf[morton3(ix, min(iy + 1, gy), iz)] - sdf[morton3(ix, max(iy - 1, 0), iz)];
      const float gnz = sdf[morton3(ix, iy, min(iz + 1, gz))] - sdf[morton3(ix, iy, max(iz - 1, 0))];
      const float gl = len3(gnx, gny, gnz);
      if (gl > 1e-6f) {
        const float ux = gnx / gl, uy = gny / gl, uz = gnz / gl;
        xn -= psdf * ux; yn -= psdf * uy; zn -= psdf * uz;
        const float vn = vnx * ux + vny * uy + vnz * uz;
        if (vn < 0.f) { vnx -= vn * ux; vny -= vn * uy; vnz -= vn * uz; }
      } else {
        xn = xp; yn = yp; zn = zp;
        vnx = vny = vnz = 0.f;
      }
      Fn = m3_ident();
      tau = m3_zero();
      pushed = true;
    }
  }
  // world clamp at a 1.5h margin (shaders/mpm_g2p.comp:255-263)
  {
    const float mg = ge.h * 1.5f;
    const float lox = ge.wmin[0] + mg, loy = ge.wmin[1] + mg, loz = ge.wmin[2] + mg;
    const float hix = ge.wmax[0] - mg, hiy = ge.wmax[1] - mg, hiz = ge.wmax[2] - mg;
    if (xn < lox) { xn = lox; vnx = fmaxf(vnx, 0.f); }
    if (xn > hix) { xn = hix; vnx = fminf(vnx, 0.f); }
    if (yn < loy) { yn = loy; vny = fmaxf(vny, 0.f); }
    if (yn > hiy) { yn = hiy; vny = fminf(vny, 0.f); }
    if (zn < loz) { zn = loz; vnz = fmaxf(vnz, 0.f); }
    if (zn > hiz) { zn = hiz; vnz = fminf(vnz, 0.f); }
  }
  ro[R_P] = make_float4(xn, yn, zn, Vp);
  ro[R_V] = make_float4(vnx, vny, vnz, vel.w);
  ro[R_B0] = make_float4(Bn.m[0][0], Bn.m[0][1], Bn.m[0][2], tau.m[0][1]);
  ro[R_B1] = make_float4(Bn.m[1][0], Bn.m[1][1], Bn.m[1][2], tau.m[0][2]);
  ro[R_B2] = make_float4(Bn.m[2][0], Bn.m[2][1], Bn.m[2][2], tau.m[1][2]);
  ro[R_T] = make_float4(tau.m[0][0], tau.m[1][1], tau.m[2][2], Fn.m[0][0]);
  ro[R_FA] = make_float4(Fn.m[0][1], Fn.m[0][2], Fn.m[1][0], Fn.m[1][1]);
  ro[R_FB] = make_float4(Fn.m[1][2], Fn.m[2][0], Fn.m[2][1], Fn.m[2][2]);
  keys[q] = cell_key(ge, xn, yn, zn);  // the next substep's bin key
  // F and tau were set to I and 0: the fallback's entry is dropped
  if (fb_at != kFbNone && (jreset || pushed)) reinterpret_cast<uint32_t*>(p.fb_queue + size_t(fb_at) * kFbEntry)[1] = 0u;
}

// the material of a particle: clamp to materialCount-1 (shaders/mpm_g2p.comp:128-130)
__device__ __forceinline__ uint32_t particle_material(const SubParams& p, const float4 vel) {
  const uint32_t mid = __float_as_uint(vel.w);
  return mid > p.material_count - 1u ? p.material_count - 1u : mid;
}

// G2P of particle q (place k of the G2P order) reading the updated grid from HBM.
template <bool FLIP, bool BATCH>
__device__ __forceinline__ void g2p_one(const SubParams& p, uint32_t k, uint32_t q, const float4* rec_in, float4* rec_out,
                                        uint32_t* __restrict__ keys, const nb_material* __restrict__ mats,
                                        const nb_material* s_mats, const float4* __restrict__ grid_node,
                                        const float4* __restrict__ grid_old, const float* __restrict__ sdf) {
  const SceneGeo g = particle_geo<BATCH>(p, q);
  const float4* r = rec_in + size_t(q) * kRec;
  float4* ro = rec_out + size_t(q) * kRec;
  const float4 pp = r[R_P];
  int base[3];
  Axis3 ax3[3];
  const bool tiny = g2p_stencil(g, pp.x, pp.y, pp.z, base, ax3);
  // out-of-grid nodes read a clamped (valid) address and are dropped together
  // with the massless ones
  auto node = [&](int ix, int iy, int iz, float (&v)[3], float (&vo)[3]) -> bool {
    const int nx = base[0] + ix, ny = base[1] + iy, nz = base[2] + iz;
    const uint32_t cx = uint32_t(min(max(nx, 0), g.grid_res[0] - 1));
    const uint32_t cy = uint32_t(min(max(ny, 0), g.grid_res[1] - 1));
    const uint32_t cz = uint32_t(min(max(nz, 0), g.grid_res[2] - 1));
    const uint32_t id = node_id(g, cx, cy, cz);
    const float4 gn = grid_node[id];
    // bitwise, not short-circuit: the load then has a use on every path and
    // stays one unconditional 16-byte load, issued with the slab's others
    const bool live = (uint32_t(nx) < uint32_t(g.grid_res[0])) & (uint32_t(ny) < uint32_t(g.grid_res[1])) &
                      (uint32_t(nz) < uint32_t(g.grid_res[2])) & (gn.w >= 1e-10f);
    v[0] = live ? gn.x : 0.f;
    v[1] = live ? gn.y : 0.f;
    v[2] = live ? gn.z : 0.f;
    if (FLIP) {
      const float4 o = grid_old[id];
      vo[0] = live ? o.x : 0.f;
      vo[1] = live ? o.y : 0.f;
      vo[2] = live ? o.z : 0.f;
    }
    return live;
  };
  G2PGather gth;
  g2p_gather<FLIP>(ax3, tiny, node, gth);
  const float4 vel = r[R_V];
  const uint32_t mid = particle_material(p, vel);
  const nb_material mat = BATCH ? mats[mid] : s_mats[mid];
  g2p_particle<FLIP, BATCH>(p, k, q, r, ro, keys, mat, gth, pp, vel, sdf);
}

// a plain engine's material table is staged in LDS; a batch's (up to
// kBatchMaxMaterials) is read through the caches
template <bool BATCH>
__device__ __forceinline__ void stage_materials(const SubParams& p, const nb_material* __restrict__ mats,
                                                nb_material* s_mats) {
  if constexpr (!BATCH) {
    float* dst = reinterpret_cast<float*>(s_mats);
    const float* src = reinterpret_cast<const float*>(mats);
    const uint32_t nf = p.material_count * uint32_t(sizeof(nb_material) / 4);
    for (uint32_t e = threadIdx.x; e < nf; e += blockDim.x) dst[e] = src[e];
    __syncthreads();
  }
}

// G2P reading the updated grid from HBM, one thread per particle in bin order.
// rec_out may equal rec_in: every record slot written is data-dependent on the
// load of the same slot (or not read at all), so updating in place is safe.
// 4 waves per SIMD (<= 128 VGPRs): the gather's latency needs the occupancy
template <bool FLIP, bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_g2p(SubParams p,
                                                                                   const float4* rec_in,
                                                                                   float4* rec_out,
                                             uint32_t* __restrict__ keys,
                                             const nb_material* __restrict__ mats,
                                             const uint32_t* __restrict__ sorted,
                                             const float4* __restrict__ grid_node,
                                             const float4* __restrict__ grid_old, const float* __restrict__ sdf) {
  __shared__ nb_material s_mats[BATCH ? 1 : NB_MAX_MATERIALS];
  stage_materials<BATCH>(p, mats, s_mats);
  const uint32_t k = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (k >= p.n) return;
  const uint32_t q = sorted ? sorted[k] : k;
  g2p_one<FLIP, BATCH>(p, k, q, rec_in, rec_out, keys, mats, s_mats, grid_node, grid_old, sdf);
}

// The FP64 fallback's second half: for every entry G2P queued (queue_fb),
// the return map and Hencky stress in FP64 from the trial F (mode 1), or the
// Hencky stress of the record's F (mode 2), rounded once; F is replaced and
// the stress added to the record's (the mud viscosity term, or zero).
// Every block reads the slots' lengths and the grid strides over their
// concatenation; k_g2p_fb_reset then zeroes the lengths for the next G2P.
// (A count of finished blocks, the last one zeroing, cost tens of ns per
// block: device-scope atomics on one address serialise across the XCDs.)
template <bool BATCH>
__global__ __launch_bounds__(256) void k_g2p_fb(SubParams p, float4* rec, const nb_material* __restrict__ mats) {
  __shared__ nb_material s_mats[BATCH ? 1 : NB_MAX_MATERIALS];
  stage_materials<BATCH>(p, mats, s_mats);
  // the slots' lengths and their exclusive prefix (wave 0), so that the
  // grid's threads share the queue's entries evenly whatever the slots hold
  __shared__ uint32_t s_pre[kFbSlots + 1];
  if (threadIdx.x < 64u) {
    const uint32_t len = threadIdx.x < kFbSlots ? p.fb_ctl[threadIdx.x * kFbCtlStride] : 0u;
    uint32_t inc = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = uint32_t(__shfl_up(int(inc), d));
      if (threadIdx.x >= uint32_t(d)) inc += o;
    }
    if (threadIdx.x < kFbSlots) s_pre[threadIdx.x] = inc - len;
    if (threadIdx.x == kFbSlots - 1u) s_pre[kFbSlots] = inc;
  }
  __syncthreads();
  const uint32_t total = s_pre[kFbSlots];
  const bool cs = p.corrected_svd != 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    // the slot holding entry i of the concatenated queue: the last s with s_pre[s] <= i
    uint32_t slot = 0u;
#pragma unroll
    for (uint32_t step = kFbSlots / 2u; step > 0u; step >>= 1)
      if (s_pre[slot + step] <= i) slot += step;
    const float4* e = p.fb_queue + (size_t(slot) * p.fb_region + (i - s_pre[slot])) * kFbEntry;
    const float4 e0 = e[0];
    const uint32_t q = __float_as_uint(e0.x), mode = __float_as_uint(e0.y);
    if (mode == 0u) continue;
    float4* ro = rec + size_t(q) * kRec;
    const uint32_t mid = particle_material(p, ro[R_V]);
    const nb_material mat = BATCH ? mats[mid] : s_mats[mid];
    const bool dp = true;  // only the Drucker-Prager family is queued
    const float4 rt = ro[R_T];
    float* b0 = reinterpret_cast<float*>(ro + R_B0);
    float* b1 = reinterpret_cast<float*>(ro + R_B1);
    float* b2 = reinterpret_cast<float*>(ro + R_B2);
    m3 F, tau;
    if (mode == 1u) {
      const float4 e1 = e[1], e2 = e[2], e3 = e[3], e4 = e[4];
      m3 Fo, L;
      Fo.m[0][0] = e0.z; Fo.m[0][1] = e0.w; Fo.m[0][2] = e1.x;
      Fo.m[1][0] = e1.y; Fo.m[1][1] = e1.z; Fo.m[1][2] = e1.w;
      Fo.m[2][0] = e2.x; Fo.m[2][1] = e2.y; Fo.m[2][2] = e2.z;
      const float vg[9] = {e2.w, e3.x, e3.y, e3.z, e3.w, e4.x, e4.y, e4.z, e4.w};
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int rr = 0; rr < 3; rr++) L.m[c][rr] = (c == rr ? 1.f : 0.f) + p.dt * vg[c * 3 + rr];
      const FTau ft = plastic_fp64(m3_mul(L, Fo), dp, mat.mu, mat.lambda, mat.q_max, mat.M_friction,
                                   mat.q_cohesion, cs);
      F = ft.F;
      tau = ft.tau;
    } else {
      const float4 fa = ro[R_FA], fb = ro[R_FB];
      F.m[0][0] = rt.w; F.m[0][1] = fa.x; F.m[0][2] = fa.y;
      F.m[1][0] = fa.z; F.m[1][1] = fa.w; F.m[1][2] = fb.x;
      F.m[2][0] = fb.y; F.m[2][1] = fb.z; F.m[2][2] = fb.w;
      tau = m3_round(hencky_stress(m3_convert<double>(F), double(mat.mu), double(mat.lambda), cs));
    }
    // the record's stress: the viscosity term G2P left there
    ro[R_T] = make_float4(tau.m[0][0] + rt.x, tau.m[1][1] + rt.y, tau.m[2][2] + rt.z, F.m[0][0]);
    b0[3] = tau.m[0][1] + b0[3];
    b1[3] = tau.m[0][2] + b1[3];
    b2[3] = tau.m[1][2] + b2[3];
    if (mode == 1u) {
      ro[R_FA] = make_float4(F.m[0][1], F.m[0][2], F.m[1][0], F.m[1][1]);
      ro[R_FB] = make_float4(F.m[1][2], F.m[2][0], F.m[2][1], F.m[2][2]);
    }
  }
}

__global__ __launch_bounds__(64) void k_g2p_fb_reset(uint32_t* __restrict__ fb_ctl) {
  if (threadIdx.x < kFbSlots) fb_ctl[threadIdx.x * kFbCtlStride] = 0u;
}

// Grid update and G2P fused, one block per work unit: a unit is up to
// kG2PUnit particles of one occupied tile (a tile's particles are one
// contiguous range of the bin order; k_bin_tile_offsets lists the units).
// The block builds the tile's 10^3 node box in LDS from the partials (the
// same sums and node update as k_grid, bit for bit), stores each node's
// G2P view (velocity, or zero for a node that does not take part, and FLIP's
// saved velocity), then runs G2P for the unit's particles gathering from
// LDS.  A particle whose stencil leaves the box (its position lies outside
// the grid) reads the box with a range check: the nodes outside it are
// outside the grid too.  Results equal k_grid + k_g2p bit for bit; the
// grid itself is not written (the last substep of a call runs unfused, so the
// exported grid stays that of the last substep).
template <bool FLIP, bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_g2p_tile(
    SubParams p, const float4* rec_in, float4* rec_out, uint32_t* __restrict__ keys,
    const nb_material* __restrict__ mats, const uint32_t* __restrict__ sorted, const uint2* __restrict__ units,
    const uint32_t* __restrict__ act_tiles, const uint32_t* __restrict__ tile_start,
    const uint32_t* __restrict__ tile_total, const uint32_t* __restrict__ counters,
    const int32_t* __restrict__ tile_act, const float4* __restrict__ partial, const float* __restrict__ sdf,
    const nb_collider* __restrict__ cols, const nb_force* __restrict__ forces) {
  __shared__ float4 s_v[kBoxNodes];                 // xyz: velocity (0 when not taking part), w: 1 / 0
  __shared__ float4 s_o[FLIP ? kBoxNodes : 1];      // FLIP's saved velocity
  __shared__ nb_material s_mats[BATCH ? 1 : NB_MAX_MATERIALS];
  if constexpr (!BATCH) {
    float* dst = reinterpret_cast<float*>(s_mats);
    const float* src = reinterpret_cast<const float*>(mats);
    const uint32_t nf = p.material_count * uint32_t(sizeof(nb_material) / 4);
    for (uint32_t e = threadIdx.x; e < nf; e += blockDim.x) dst[e] = src[e];
  }
  const uint32_t n_units = counters[CNT_UNITS];
  for (uint32_t u = blockIdx.x; u < n_units; u += gridDim.x) {
    const uint2 unit = units[u];
    const uint32_t a = unit.x, chunk = unit.y;
    const uint32_t T = act_tiles[a];
    const SceneGeo g = tile_geo<BATCH>(p, T);
    int tx, ty, tz;
    tile_coords(g, T, tx, ty, tz);
    __syncthreads();  // the previous unit's particles are done with the box
    for (uint32_t e = threadIdx.x; e < kBoxNodes; e += blockDim.x) {
      const int bx = int(e % kBox), by = int((e / kBox) % kBox), bz = int(e / (kBox * kBox));
      const int nx = tx * 8 - 1 + bx, ny = ty * 8 - 1 + by, nz = tz * 8 - 1 + bz;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f), o = make_float4(0.f, 0.f, 0.f, 0.f);
      if (nx >= 0 && nx < g.grid_res[0] && ny >= 0 && ny < g.grid_res[1] && nz >= 0 && nz < g.grid_res[2]) {
        // the node's own tile and its place in it, as k_grid sees the node
        const int ntx = nx >> 3, nty = ny >> 3, ntz = nz >> 3;
        float mx, my, mz, mass;
        (void)sum_partials(g, ntx, nty, ntz, nx & 7, ny & 7, nz & 7, tile_act, partial, mx, my, mz, mass);
        if (mass >= 1e-10f) {
          float vx = mx / mass, vy = my / mass, vz = mz / mass, ox, oy, oz;
          node_update<false>(p, g, nx, ny, nz, vx, vy, vz, ox, oy, oz, sdf, cols, forces);
          v = make_float4(vx, vy, vz, 1.f);
          o = make_float4(ox, oy, oz, 0.f);
        }
      }
      s_v[e] = v;
      if (FLIP) s_o[e] = o;
    }
    __syncthreads();
    const uint32_t first = tile_start[a] + chunk * kG2PUnit;
    const uint32_t last = min(tile_start[a] + tile_total[a], first + kG2PUnit);
    // box node of grid node n along each axis: n - (8 t - 1)
    const int ox0 = tx * 8 - 1, oy0 = ty * 8 - 1, oz0 = tz * 8 - 1;
    for (uint32_t k = first + threadIdx.x; k < last; k += blockDim.x) {
      const uint32_t q = sorted[k];
      const float4* r = rec_in + size_t(q) * kRec;
      float4* ro = rec_out + size_t(q) * kRec;
      const float4 pp = r[R_P];
      int base[3];
      Axis3 ax3[3];
      const bool tiny = g2p_stencil(g, pp.x, pp.y, pp.z, base, ax3);
      const int bx0 = base[0] - ox0, by0 = base[1] - oy0, bz0 = base[2] - oz0;
      const bool inbox = bx0 >= 0 && bx0 <= kBox - 3 && by0 >= 0 && by0 <= kBox - 3 && bz0 >= 0 && bz0 <= kBox - 3;
      const int e0 = (bz0 * kBox + by0) * kBox + bx0;
      G2PGather gth;
      if (inbox) {
        auto node = [&](int ix, int iy, int iz, float (&v)[3], float (&vo)[3]) -> bool {
          const int e = e0 + (iz * kBox + iy) * kBox + ix;
          const float4 n4 = s_v[e];
          v[0] = n4.x; v[1] = n4.y; v[2] = n4.z;
          if (FLIP) {
            const float4 o4 = s_o[e];
            vo[0] = o4.x; vo[1] = o4.y; vo[2] = o4.z;
          }
          return n4.w != 0.f;
        };
        g2p_gather<FLIP>(ax3, tiny, node, gth);
      } else {
        auto node = [&](int ix, int iy, int iz, float (&v)[3], float (&vo)[3]) -> bool {
          const int bx = bx0 + ix, by = by0 + iy, bz = bz0 + iz;
          const bool in = bx >= 0 && bx < kBox && by >= 0 && by < kBox && bz >= 0 && bz < kBox;
          const int e = in ? (bz * kBox + by) * kBox + bx : 0;
          const float4 n4 = s_v[e];
          const bool live = in && n4.w != 0.f;
          v[0] = live ? n4.x : 0.f; v[1] = live ? n4.y : 0.f; v[2] = live ? n4.z : 0.f;
          if (FLIP) {
            const float4 o4 = s_o[e];
            vo[0] = live ? o4.x : 0.f; vo[1] = live ? o4.y : 0.f; vo[2] = live ? o4.z : 0.f;
          }
          return live;
        };
        g2p_gather<FLIP>(ax3, tiny, node, gth);
      }
      const float4 vel = r[R_V];
      const uint32_t mid = particle_material(p, vel);
      const nb_material mat = BATCH ? mats[mid] : s_mats[mid];
      g2p_particle<FLIP, BATCH>(p, k, q, r, ro, keys, mat, gth, pp, vel, sdf);
    }
  }
}

// ---------------------------------------------------------------------------
// utilities
// ---------------------------------------------------------------------------
__global__ void k_sdf_sphere(float* __restrict__ sdf, SubParams p, float r, float cx, float cy, float cz) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nc = p.cube_res * p.cube_res * p.cube_res;
  if (c >= nc) return;
  const int x = int(morton_compact(c)), y = int(morton_compact(c >> 1)), z = int(morton_compact(c >> 2));
  if (x >= p.grid_res[0] || y >= p.grid_res[1] || z >= p.grid_res[2]) {
    sdf[c] = 0.f;  // cells outside gridRes stay unwritten (zero) as in the reference's host loop
    return;
  }
  // src/engine/MPMEngine.cpp:186-193 (cell centres from the origin)
  const float wx = (float(x) + 0.5f) * p.h, wy = (float(y) + 0.5f) * p.h, wz = (float(z) + 0.5f) * p.h;
  const float d = sqrtf((wx - cx) * (wx - cx) + (wy - cy) * (wy - cy) + (wz - cz) * (wz - cz));
  sdf[c] = d - r;
}

__global__ void k_fill_identity(uint32_t* __restrict__ a, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

// F = I, B = 0, tau = 0 for rows [first, first+count) (appendParticles)
__global__ void k_init_FB(float4* F0, float4* F1, float4* F2, float4* B0, float4* B1, float4* B2, uint32_t first,
                          uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t q = first + i;
  F0[q] = make_float4(1.f, 0.f, 0.f, 0.f);
  F1[q] = make_float4(0.f, 1.f, 0.f, 0.f);
  F2[q] = make_float4(0.f, 0.f, 1.f, 0.f);
  B0[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  B1[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  B2[q] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// reference layout (eight float4 arrays) <-> one 128-byte record per particle
__global__ __launch_bounds__(256) void k_soa_to_aos(FieldPtrs f, float4* __restrict__ rec, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 F0 = f.f[NB_FIELD_F0][i], F1 = f.f[NB_FIELD_F1][i], F2 = f.f[NB_FIELD_F2][i];
  float4* r = rec + size_t(i) * kRec;
  r[R_P] = f.f[NB_FIELD_P][i];
  r[R_V] = f.f[NB_FIELD_V][i];
  r[R_B0] = f.f[NB_FIELD_B0][i];
  r[R_B1] = f.f[NB_FIELD_B1][i];
  r[R_B2] = f.f[NB_FIELD_B2][i];
  r[R_T] = make_float4(F0.w, F1.w, F2.w, F0.x);
  r[R_FA] = make_float4(F0.y, F0.z, F1.x, F1.y);
  r[R_FB] = make_float4(F1.z, F2.x, F2.y, F2.z);
}
__global__ __launch_bounds__(256) void k_aos_to_soa(const float4* __restrict__ rec, FieldPtrs f, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* r = rec + size_t(i) * kRec;
  const float4 t = r[R_T], a = r[R_FA], b = r[R_FB];
  f.f[NB_FIELD_P][i] = r[R_P];
  f.f[NB_FIELD_V][i] = r[R_V];
  f.f[NB_FIELD_B0][i] = r[R_B0];
  f.f[NB_FIELD_B1][i] = r[R_B1];
  f.f[NB_FIELD_B2][i] = r[R_B2];
  f.f[NB_FIELD_F0][i] = make_float4(t.w, a.x, a.y, t.x);
  f.f[NB_FIELD_F1][i] = make_float4(a.z, a.w, b.x, t.y);
  f.f[NB_FIELD_F2][i] = make_float4(b.y, b.z, b.w, t.z);
}

// plain streaming copy (HBM byte-counter calibration)
__global__ __launch_bounds__(256) void k_copy(float4* __restrict__ dst, const float4* __restrict__ src, uint64_t n4) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
// blocks for a tile-stride kernel: never more than tiles, at most the tuned cap
static inline uint32_t tile_grid(uint32_t n_tiles, uint32_t cap) {
  const uint32_t c = cap ? cap : 2048u;
  return n_tiles < c ? (n_tiles ? n_tiles : 1u) : c;
}

struct ProfScope {
  Prof* pf;
  int id;
  hipStream_t s;
  ProfScope(Prof* pf_, int id_, hipStream_t s_) : pf(pf_), id(id_), s(s_) {
    if (pf) (void)hipEventRecord(pf->ev[2 * id], s);
  }
  ~ProfScope() {
    if (pf) {
      (void)hipEventRecord(pf->ev[2 * id + 1], s);
      pf->record(id);
    }
  }
};

void launch_bin(const DevState& d, const SubParams& p, uint32_t n_tiles, bool keys_valid, hipStream_t s, Prof* pf) {
  const uint32_t tg = tile_grid(n_tiles, p.tile_blocks);
  const bool batch = p.n_scenes > 0;
  { ProfScope ps(pf, 0, s); k_bin_clear<<<tg, 512, 0, s>>>(d.cell_count, d.act_tiles, d.counters); }
  {
    ProfScope ps(pf, 1, s);
    const uint32_t nb = cdiv(p.n, 256);
    if (keys_valid)
      k_bin_keys<false, false><<<nb, 256, 0, s>>>(p, d.rec, d.keys, d.rank, d.cell_count, d.tile_flag);
    else if (batch)
      k_bin_keys<true, true><<<nb, 256, 0, s>>>(p, d.rec, d.keys, d.rank, d.cell_count, d.tile_flag);
    else
      k_bin_keys<true, false><<<nb, 256, 0, s>>>(p, d.rec, d.keys, d.rank, d.cell_count, d.tile_flag);
  }
  {
    ProfScope ps(pf, 2, s);
    if (batch)
      k_bin_mark<true><<<cdiv(n_tiles, 256), 256, 0, s>>>(p, d.tile_flag, d.tile_mark, d.blk_cnt, n_tiles);
    else
      k_bin_mark<false><<<cdiv(n_tiles, 256), 256, 0, s>>>(p, d.tile_flag, d.tile_mark, d.blk_cnt, n_tiles);
  }
  {
    ProfScope ps(pf, 3, s);
    k_bin_compact<<<cdiv(n_tiles, 256), 256, 0, s>>>(d.tile_mark, d.blk_cnt, d.tile_flag, d.tile_act, d.act_tiles,
                                                     d.node_tiles, d.counters, n_tiles);
  }
  { ProfScope ps(pf, 4, s); k_bin_tile_scan<<<tg, 512, 0, s>>>(d.cell_count, d.cell_end, d.act_tiles, d.tile_total, d.counters); }
  { ProfScope ps(pf, 5, s); k_bin_tile_offsets<<<1, 1024, 0, s>>>(d.tile_total, d.tile_start, d.tile_order, d.counters, d.units); }
  { ProfScope ps(pf, 6, s); k_bin_scatter<<<cdiv(p.n, 256), 256, 0, s>>>(p.n, d.keys, d.rank, d.cell_end, d.tile_act, d.tile_start, d.sorted); }
  { ProfScope ps(pf, 7, s); k_bin_stabilize<<<tg, 512, 0, s>>>(p.n, d.cell_count, d.cell_end, d.act_tiles, d.tile_start, d.sorted, d.counters); }
}

void launch_p2g(const DevState& d, const SubParams& p, uint32_t n_tiles, hipStream_t s, Prof* pf) {
  ProfScope ps(pf, 8, s);
  if (p.n_scenes > 0)
    k_p2g<true><<<tile_grid(n_tiles, p.tile_blocks), kP2GThreads, 0, s>>>(p, d.rec, d.materials, d.cell_count, d.cell_end, d.sorted,
                                                           d.act_tiles, d.tile_order, d.tile_start, d.tile_total,
                                                           d.counters, d.partial);
  else
    k_p2g<false><<<tile_grid(n_tiles, p.tile_blocks), kP2GThreads, 0, s>>>(p, d.rec, d.materials, d.cell_count, d.cell_end,
                                                            d.sorted, d.act_tiles, d.tile_order, d.tile_start,
                                                            d.tile_total, d.counters, d.partial);
}

void launch_grid(const DevState& d, const SubParams& p, uint32_t n_tiles, bool from_partials, bool update,
                 const uint32_t* tiles, const uint32_t* count, hipStream_t s, Prof* pf) {
  ProfScope ps(pf, 9, s);
  const uint32_t g = tile_grid(n_tiles, p.tile_blocks);
#define NB_GRID_LAUNCH(FP, UP, BT, NZ)                                                                       \
  k_grid<FP, UP, BT, NZ><<<g, 512, 0, s>>>(p, d.partial, d.tile_act, tiles, count, d.grid_node, d.grid_old, \
                                           d.grid_pre, d.sdf, d.colliders, d.forces)
  const bool nz = p.noise_forces != 0;
  if (p.n_scenes > 0) {
    if (from_partials && update) {
      if (nz) NB_GRID_LAUNCH(true, true, true, true); else NB_GRID_LAUNCH(true, true, true, false);
    } else if (from_partials) {
      NB_GRID_LAUNCH(true, false, true, false);
    } else {
      if (nz) NB_GRID_LAUNCH(false, true, true, true); else NB_GRID_LAUNCH(false, true, true, false);
    }
  } else {
    if (from_partials && update) {
      if (nz) NB_GRID_LAUNCH(true, true, false, true); else NB_GRID_LAUNCH(true, true, false, false);
    } else if (from_partials) {
      NB_GRID_LAUNCH(true, false, false, false);
    } else {
      if (nz) NB_GRID_LAUNCH(false, true, false, true); else NB_GRID_LAUNCH(false, true, false, false);
    }
  }
#undef NB_GRID_LAUNCH
}

// The fused grid update + G2P (k_g2p_tile) over the work units the bin listed
// (not with turbulence or noise forces: their noise code would halve the
// kernel's occupancy; those substeps run k_grid + k_g2p).
void launch_g2p_fused(const DevState& d, const SubParams& p, uint32_t n_tiles, hipStream_t s, Prof* pf) {
  // units <= occupied tiles + n / kG2PUnit; blocks stride over them
  const uint32_t max_units = n_tiles + p.n / kG2PUnit + 1u;
  const uint32_t g = tile_grid(max_units, p.tile_blocks);
  const bool flip = p.flip_ratio > 0.01f, bt = p.n_scenes > 0;
#define NB_G2PT(FL, BT)                                                                                    \
  k_g2p_tile<FL, BT><<<g, 256, 0, s>>>(p, d.rec, d.rec, d.keys, d.materials, d.sorted, d.units, d.act_tiles, \
                                       d.tile_start, d.tile_total, d.counters, d.tile_act, d.partial, d.sdf, \
                                       d.colliders, d.forces)
  {
    ProfScope ps(pf, 11, s);
    if (bt) {
      if (flip) NB_G2PT(true, true); else NB_G2PT(false, true);
    } else {
      if (flip) NB_G2PT(true, false); else NB_G2PT(false, false);
    }
  }
#undef NB_G2PT
  launch_g2p_fb(d, p, d.rec, s, pf);
}

// The FP64 fallback's fix-up after a G2P (nothing to do when it is off).
void launch_g2p_fb(const DevState& d, const SubParams& p, float4* rec, hipStream_t s, Prof* pf) {
  if (!p.svd_fp64) return;
  ProfScope ps(pf, 12, s);
  // a fixed grid, at most one thread per particle, that strides over the
  // queue: 4096 blocks give every entry its own thread up to a queue of a
  // million (a block past the queue's end only reads the 64 lengths)
  const uint32_t g = std::min(cdiv(p.n, 256u), 4096u);
  if (p.n_scenes > 0)
    k_g2p_fb<true><<<g, 256, 0, s>>>(p, rec, d.materials);
  else
    k_g2p_fb<false><<<g, 256, 0, s>>>(p, rec, d.materials);
  k_g2p_fb_reset<<<1, 64, 0, s>>>(p.fb_ctl);
}

void launch_g2p(const DevState& d, const SubParams& p, bool use_sorted, hipStream_t s, Prof* pf, float4* rec_out) {
  const uint32_t* srt = use_sorted ? d.sorted : nullptr;
  float4* out = rec_out ? rec_out : d.rec;
  const bool flip = p.flip_ratio > 0.01f, bt = p.n_scenes > 0;
  {
    ProfScope ps(pf, 10, s);
#define NB_G2P_LAUNCH(FL, BT) \
  k_g2p<FL, BT><<<cdiv(p.n, bs), bs, 0, s>>>(p, d.rec, out, d.keys, d.materials, srt, d.grid_node, d.grid_old, d.sdf)
    const uint32_t bs = (p.g2p_block == 64u || p.g2p_block == 128u) ? p.g2p_block : 256u;  // launch bound 256
    if (bt) {
      if (flip) NB_G2P_LAUNCH(true, true);
      else NB_G2P_LAUNCH(false, true);
    } else {
      if (flip) NB_G2P_LAUNCH(true, false);
      else NB_G2P_LAUNCH(false, false);
    }
#undef NB_G2P_LAUNCH
  }
  launch_g2p_fb(d, p, out, s, pf);
}

// ---------------------------------------------------------------------------
// mesh SDF (src/core/MeshSDF.h:85-126): per cell of the grid, the distance to
// the closest point of any triangle (Ericson's closest-point regions,
// MeshSDF.h:42-75), signed by that triangle's normal (the first triangle at
// the minimum distance); cube cells outside the grid hold 1e9.  One thread
// per cell, triangles staged through LDS; every product and sum is rounded
// as written (no FMA contraction), as on the host.
// ---------------------------------------------------------------------------
constexpr uint32_t kMeshChunk = 256;         // triangles staged in LDS at a time
constexpr uint32_t kMeshLaunchTris = 2048;   // triangles per launch
template <bool EXACT>
__global__ __launch_bounds__(256) void k_mesh_sdf(const float* __restrict__ tris, uint32_t n_tris, int rx, int ry,
                                                  int rz, float cs, float* __restrict__ sdf) {
#pragma clang fp contract(off)
  __shared__ float s_t[kMeshChunk * 12];
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = c < uint32_t(rx) * uint32_t(ry) * uint32_t(rz);
  const int ix = int(c % uint32_t(rx)), iy = int((c / uint32_t(rx)) % uint32_t(ry)), iz = int(c / (uint32_t(rx) * uint32_t(ry)));
  const float px = (float(ix) + 0.5f) * cs, py = (float(iy) + 0.5f) * cs, pz = (float(iz) + 0.5f) * cs;
  const uint32_t m = valid ? morton3(uint32_t(ix), uint32_t(iy), uint32_t(iz)) : 0u;
  // the minimum over the triangles of earlier launches: its magnitude and
  // whether it lies outside (1e9 and outside before the first)
  const float prev = valid ? sdf[m] : 1e9f;
  float best = fabsf(prev);
  bool outer = !signbit(prev);
  auto dot3 = [](float ax, float ay, float az, float bx, float by, float bz) { return (ax * bx + ay * by) + az * bz; };
  for (uint32_t t0 = 0; t0 < n_tris; t0 += kMeshChunk) {
    const uint32_t nt = min(kMeshChunk, n_tris - t0);
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < nt * 12u; e += blockDim.x) s_t[e] = tris[size_t(t0) * 12 + e];
    __syncthreads();
    if (!valid) continue;
    for (uint32_t j = 0; j < nt; j++) {
      const float* t = s_t + 12 * j;  // normal, v0, v1, v2
      const float ax = t[3], ay = t[4], az = t[5];
      const float abx = t[6] - ax, aby = t[7] - ay, abz = t[8] - az;
      const float acx = t[9] - ax, acy = t[10] - ay, acz = t[11] - az;
      const float apx = px - ax, apy = py - ay, apz = pz - az;
      float qx, qy, qz;  // closest point
      const float d1 = dot3(abx, aby, abz, apx, apy, apz), d2 = dot3(acx, acy, acz, apx, apy, apz);
      if (d1 <= 0.f && d2 <= 0.f) {
        qx = ax; qy = ay; qz = az;
      } else {
        const float bpx = px - t[6], bpy = py - t[7], bpz = pz - t[8];
        const float d3 = dot3(abx, aby, abz, bpx, bpy, bpz), d4 = dot3(acx, acy, acz, bpx, bpy, bpz);
        const float vc = d1 * d4 - d3 * d2;
        if (d3 >= 0.f && d4 <= d3) {
          qx = t[6]; qy = t[7]; qz = t[8];
        } else if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
          const float v = d1 / (d1 - d3);
          qx = ax + v * abx; qy = ay + v * aby; qz = az + v * abz;
        } else {
          const float cpx = px - t[9], cpy = py - t[10], cpz = pz - t[11];
          const float d5 = dot3(abx, aby, abz, cpx, cpy, cpz), d6 = dot3(acx, acy, acz, cpx, cpy, cpz);
          const float vb = d5 * d2 - d1 * d6;
          if (d6 >= 0.f && d5 <= d6) {
            qx = t[9]; qy = t[10]; qz = t[11];
          } else if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
            const float w = d2 / (d2 - d6);
            qx = ax + w * acx; qy = ay + w * acy; qz = az + w * acz;
          } else {
            const float va = d3 * d6 - d5 * d4;
            if (EXACT && va <= 0.f && d4 - d3 >= 0.f && d5 - d6 >= 0.f) {  // edge v1-v2
              const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
              qx = t[6] + w * (t[9] - t[6]); qy = t[7] + w * (t[10] - t[7]); qz = t[8] + w * (t[11] - t[8]);
            } else {
              const float sum = va + vb + vc;
              const float v = vb / sum, w = vc / sum;
              qx = ax + v * abx + w * acx; qy = ay + v * aby + w * acy; qz = az + v * abz + w * acz;
            }
          }
        }
      }
      const float dx = px - qx, dy = py - qy, dz = pz - qz;
      const float d = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
      if (d < best) {
        best = d;
        outer = dot3(dx, dy, dz, t[0], t[1], t[2]) >= 0.f;
      }
    }
  }
  if (valid) sdf[m] = outer ? best : -best;
}
__global__ void k_fill(float* __restrict__ a, uint32_t n, float v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

void launch_mesh_sdf(const float* tris, uint32_t n_tris, const int32_t res[3], uint32_t cube, float cs, bool exact,
                     float* sdf, hipStream_t s) {
  const uint32_t nc = cube * cube * cube, cells = uint32_t(res[0]) * uint32_t(res[1]) * uint32_t(res[2]);
  k_fill<<<cdiv(nc, 256), 256, 0, s>>>(sdf, nc, 1e9f);
  if (!cells) return;
  // one launch per chunk of triangles, so that no launch runs long on a large
  // mesh; each continues from the minimum the earlier ones left in sdf
  for (uint32_t t0 = 0; t0 < n_tris; t0 += kMeshLaunchTris) {
    const uint32_t nt = std::min(kMeshLaunchTris, n_tris - t0);
    const float* tr = tris + size_t(t0) * 12;
    if (exact)
      k_mesh_sdf<true><<<cdiv(cells, 256), 256, 0, s>>>(tr, nt, res[0], res[1], res[2], cs, sdf);
    else
      k_mesh_sdf<false><<<cdiv(cells, 256), 256, 0, s>>>(tr, nt, res[0], res[1], res[2], cs, sdf);
  }
}

void launch_sdf_sphere(float* sdf, const SubParams& p, float r, float cx, float cy, float cz, hipStream_t s) {
  const uint32_t nc = p.cube_res * p.cube_res * p.cube_res;
  k_sdf_sphere<<<cdiv(nc, 256), 256, 0, s>>>(sdf, p, r, cx, cy, cz);
}

void launch_fill_identity(uint32_t* a, uint32_t n, hipStream_t s) { k_fill_identity<<<cdiv(n, 256), 256, 0, s>>>(a, n); }

void launch_soa_to_aos(const DevState& d, uint32_t n, hipStream_t s) {
  if (!n) return;
  FieldPtrs f;
  for (int i = 0; i < NB_NUM_FIELDS; i++) f.f[i] = d.field[i];
  k_soa_to_aos<<<cdiv(n, 256), 256, 0, s>>>(f, d.rec, n);
}

void launch_aos_to_soa(const DevState& d, uint32_t n, hipStream_t s) {
  if (!n) return;
  FieldPtrs f;
  for (int i = 0; i < NB_NUM_FIELDS; i++) f.f[i] = d.field[i];
  k_aos_to_soa<<<cdiv(n, 256), 256, 0, s>>>(d.rec, f, n);
}

void launch_soa_to_aos_ptrs(const FieldPtrs& f, float4* rec, uint32_t n, hipStream_t s) {
  if (n) k_soa_to_aos<<<cdiv(n, 256), 256, 0, s>>>(f, rec, n);
}
void launch_aos_to_soa_ptrs(const float4* rec, const FieldPtrs& f, uint32_t n, hipStream_t s) {
  if (n) k_aos_to_soa<<<cdiv(n, 256), 256, 0, s>>>(rec, f, n);
}

void launch_copy(float4* dst, const float4* src, uint64_t n4, hipStream_t s) {
  k_copy<<<8192, 256, 0, s>>>(dst, src, n4);
}

void launch_init_particles(const DevState& d, uint32_t first, uint32_t count, hipStream_t s) {
  if (!count) return;
  k_init_FB<<<cdiv(count, 256), 256, 0, s>>>(d.field[2], d.field[3], d.field[4], d.field[5], d.field[6], d.field[7],
                                             first, count);
}

}  // namespace nb
