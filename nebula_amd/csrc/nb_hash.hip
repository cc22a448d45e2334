// nb_hash.hip — the spatial hash of the reference's position-based engines
// (PBF fluid, cloth self-collision) behind nb_hash_build (include/nebula_hash.h).
//
// The reference records six dispatches (zero_cells, hash_count,
// hash_scan_local, hash_scan_global, hash_add_base, hash_sort) over
// engine-owned buffers.  Here:
//   k_hash_clear      zero cell_count (and the scratch counter)
//   k_hash_count      cell of every particle, wave-aggregated atomic counts
//   k_hash_chunk_sums / k_hash_chunk_scan / k_hash_offsets
//                     exclusive prefix of the counts in two levels (4096 cells
//                     per block; no group-count limit) -> START offsets
//   k_hash_scatter    slot = atomicAdd(START of the cell) -> cell_offset ends
//                     as END offsets, as after hash_sort
//   k_hash_stabilize  every cell's slice of `sorted` put in ascending particle
//                     order: cells of <= 16 in registers by a sorting network,
//                     more crowded cells listed for k_hash_crowded
//   k_hash_crowded    one block per crowded cell: bitonic sort in LDS (up to
//                     16384 particles) or in place in global memory
// Per-particle work reads the 16-byte predP (and typeFlag); there is no
// per-particle scratch.  The result does not depend on the order of the
// atomics, so it is deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/nebula_hash.h"
#include "../../include/nebula_mpm.h"
#include "nb_device_math.hpp"
#include "nb_engine.hpp"
#include "nb_kernel_util.hpp"

namespace nbh {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = 16;                         // cells per thread in the scan
constexpr uint32_t kChunk = kThreads * kItems;          // cells per scan block
constexpr uint32_t kMaxChunks = NB_HASH_MAX_CELLS / kChunk;
constexpr uint32_t kSmall = 16;                         // cells up to this size: sorted by their thread
constexpr uint32_t kLdsSort = 16384;                    // crowded cells up to this size: sorted in LDS
constexpr uint32_t kCrowdThreads = 512;
constexpr uint32_t kCrowdCap = 65536;                   // listed crowded cells; beyond: a dense sweep
// scratch layout (uint32 words)
constexpr uint32_t kOffCrowdCount = kMaxChunks;         // after the chunk sums
constexpr uint32_t kOffCrowdList = kMaxChunks + 64;
constexpr uint32_t kScratchWords = kOffCrowdList + kCrowdCap;

struct Geo {
  float wmin[3];
  float h;
  float gmax[3];  // grid_res - 1
};

// cellId (shaders/common.glsl:100-105): IEEE quotient, clamp to
// [0, gridRes - 1] (NaN maps to 0: fmaxf returns the other operand), truncation
__device__ __forceinline__ uint32_t hash_cell(const Geo& g, const float4 p) {
  const float lx = fminf(fmaxf((p.x - g.wmin[0]) / g.h, 0.f), g.gmax[0]);
  const float ly = fminf(fmaxf((p.y - g.wmin[1]) / g.h, 0.f), g.gmax[1]);
  const float lz = fminf(fmaxf((p.z - g.wmin[2]) / g.h, 0.f), g.gmax[2]);
  return nb::morton3(uint32_t(lx), uint32_t(ly), uint32_t(lz));
}

// Adds one per valid lane to counters[key], one atomic per distinct key per
// wave; returns the value before the wave's add plus the lane's position among
// the lanes of its key (lane order).  Call from uniform control flow.
__device__ __forceinline__ uint32_t wave_grouped_add(uint32_t* __restrict__ counters, uint32_t key, bool valid) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt = lane ? (~0ull >> (64u - lane)) : 0ull;
  uint64_t pending = __ballot(valid);
  int my_leader = 0;
  uint32_t my_pos = 0, my_cnt = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lkey = __shfl(key, leader);
    const uint64_t peers = __ballot(valid && key == lkey);
    if (valid && key == lkey) {
      my_leader = leader;
      my_pos = uint32_t(__popcll(peers & lt));
    }
    if (int(lane) == leader) my_cnt = uint32_t(__popcll(peers));
    pending &= ~peers;
  }
  uint32_t base = 0;
  if (valid && my_pos == 0u) base = atomicAdd(&counters[key], my_cnt);
  base = __shfl(base, my_leader);
  return base + my_pos;
}

__global__ __launch_bounds__(kThreads) void k_hash_clear(uint32_t* __restrict__ cell_count, uint32_t nc,
                                                         uint32_t* __restrict__ scratch) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if ((reinterpret_cast<uintptr_t>(cell_count) & 15u) == 0u) {
    uint4* c4 = reinterpret_cast<uint4*>(cell_count);
    for (uint32_t i = t; i < nc / 4u; i += stride) c4[i] = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t i = (nc & ~3u) + t; i < nc; i += stride) cell_count[i] = 0u;
  } else {
    for (uint32_t i = t; i < nc; i += stride) cell_count[i] = 0u;
  }
  if (t == 0) scratch[kOffCrowdCount] = 0u;
}

__global__ __launch_bounds__(kThreads) void k_hash_count(Geo g, const float4* __restrict__ pred,
                                                         const uint32_t* __restrict__ type_flag, uint32_t n,
                                                         uint32_t* __restrict__ cell_count) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  bool valid = i < n;
  uint32_t key = 0;
  if (valid) {
    if (type_flag != nullptr && type_flag[i] == NB_HASH_TYPE_UNUSED) valid = false;
    else key = hash_cell(g, pred[i]);
  }
  (void)wave_grouped_add(cell_count, key, valid);
}

// the kItems counts of thread t in chunk b (zeros past nc)
__device__ __forceinline__ void load_items(const uint32_t* __restrict__ c, uint32_t nc, uint32_t first,
                                           uint32_t (&v)[kItems]) {
  if (first + kItems <= nc && (reinterpret_cast<uintptr_t>(c + first) & 15u) == 0u) {
    const uint4* c4 = reinterpret_cast<const uint4*>(c + first);
#pragma unroll
    for (uint32_t q = 0; q < kItems / 4u; q++) {
      const uint4 a = c4[q];
      v[4 * q + 0] = a.x; v[4 * q + 1] = a.y; v[4 * q + 2] = a.z; v[4 * q + 3] = a.w;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kItems; k++) v[k] = first + k < nc ? c[first + k] : 0u;
  }
}

__global__ __launch_bounds__(kThreads) void k_hash_chunk_sums(const uint32_t* __restrict__ cell_count, uint32_t nc,
                                                              uint32_t* __restrict__ scratch) {
  __shared__ uint32_t s_wave[kThreads / 64 + 1];
  uint32_t v[kItems];
  load_items(cell_count, nc, blockIdx.x * kChunk + threadIdx.x * kItems, v);
  uint32_t s = 0;
#pragma unroll
  for (uint32_t k = 0; k < kItems; k++) s += v[k];
  uint32_t tot;
  (void)nb::block_exclusive_scan<kThreads>(s, s_wave, tot);
  if (threadIdx.x == 0) scratch[blockIdx.x] = tot;
}

// exclusive scan of the chunk sums in place, one block of 1024 threads
__global__ __launch_bounds__(1024) void k_hash_chunk_scan(uint32_t* __restrict__ scratch, uint32_t nchunks) {
  __shared__ uint32_t s_wave[1024 / 64 + 1];
  const uint32_t per = (nchunks + 1023u) / 1024u;
  const uint32_t lo = min(threadIdx.x * per, nchunks), hi = min(lo + per, nchunks);
  uint32_t s = 0;
  for (uint32_t i = lo; i < hi; i++) s += scratch[i];
  uint32_t tot;
  uint32_t run = nb::block_exclusive_scan<1024>(s, s_wave, tot);
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t x = scratch[i];
    scratch[i] = run;
    run += x;
  }
}

// START offsets: the chunk's base plus the exclusive prefix inside the chunk
__global__ __launch_bounds__(kThreads) void k_hash_offsets(const uint32_t* __restrict__ cell_count, uint32_t nc,
                                                           const uint32_t* __restrict__ scratch,
                                                           uint32_t* __restrict__ cell_off) {
  __shared__ uint32_t s_wave[kThreads / 64 + 1];
  const uint32_t first = blockIdx.x * kChunk + threadIdx.x * kItems;
  uint32_t v[kItems];
  load_items(cell_count, nc, first, v);
  uint32_t s = 0;
#pragma unroll
  for (uint32_t k = 0; k < kItems; k++) s += v[k];
  uint32_t tot;
  uint32_t run = scratch[blockIdx.x] + nb::block_exclusive_scan<kThreads>(s, s_wave, tot);
#pragma unroll
  for (uint32_t k = 0; k < kItems; k++) {
    const uint32_t x = v[k];
    v[k] = run;
    run += x;
  }
  if (first + kItems <= nc && (reinterpret_cast<uintptr_t>(cell_off + first) & 15u) == 0u) {
    uint4* o4 = reinterpret_cast<uint4*>(cell_off + first);
#pragma unroll
    for (uint32_t q = 0; q < kItems / 4u; q++) o4[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kItems; k++)
      if (first + k < nc) cell_off[first + k] = v[k];
  }
}

__global__ __launch_bounds__(kThreads) void k_hash_scatter(Geo g, const float4* __restrict__ pred,
                                                           const uint32_t* __restrict__ type_flag, uint32_t n,
                                                           uint32_t* __restrict__ cell_off, uint32_t* __restrict__ sorted) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  bool valid = i < n;
  uint32_t key = 0;
  if (valid) {
    if (type_flag != nullptr && type_flag[i] == NB_HASH_TYPE_UNUSED) valid = false;
    else key = hash_cell(g, pred[i]);
  }
  const uint32_t slot = wave_grouped_add(cell_off, key, valid);
  if (valid) sorted[slot] = i;
}

// ascending sort of the first W of v by odd-even transposition (W even)
template <uint32_t W>
__device__ __forceinline__ void sort_net(uint32_t (&v)[kSmall]) {
#pragma unroll
  for (uint32_t r = 0; r < W; r++) {
#pragma unroll
    for (uint32_t k = r & 1u; k + 1 < W; k += 2) {
      const uint32_t a = v[k], b = v[k + 1];
      v[k] = min(a, b);
      v[k + 1] = max(a, b);
    }
  }
}

template <uint32_t W>
__device__ __forceinline__ void sort_small(uint32_t* __restrict__ seg, uint32_t cnt) {
  uint32_t v[kSmall];
#pragma unroll
  for (uint32_t k = 0; k < W; k++) v[k] = k < cnt ? seg[k] : 0xFFFFFFFFu;
  sort_net<W>(v);
#pragma unroll
  for (uint32_t k = 0; k < W; k++)
    if (k < cnt) seg[k] = v[k];
}

__global__ __launch_bounds__(kThreads) void k_hash_stabilize(const uint32_t* __restrict__ cell_count,
                                                             const uint32_t* __restrict__ cell_off, uint32_t nc,
                                                             uint32_t* __restrict__ sorted,
                                                             uint32_t* __restrict__ scratch) {
  const uint32_t c = blockIdx.x * kThreads + threadIdx.x;
  uint32_t cnt = 0, end = 0;
  if (c < nc) {
    cnt = cell_count[c];
    end = cell_off[c];
  }
  const bool small = cnt >= 2u && cnt <= kSmall;
  if (cnt > kSmall) {
    const uint32_t slot = atomicAdd(&scratch[kOffCrowdCount], 1u);
    if (slot < kCrowdCap) scratch[kOffCrowdList + slot] = c;
  }
  // network width: the widest small cell of the wave
  uint32_t wmax = small ? cnt : 0u;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) wmax = max(wmax, uint32_t(__shfl_xor(wmax, d)));
  if (!small) return;
  uint32_t* seg = sorted + (end - cnt);
  if (wmax <= 4u) sort_small<4>(seg, cnt);
  else if (wmax <= 8u) sort_small<8>(seg, cnt);
  else sort_small<16>(seg, cnt);
}

// Bitonic sort (ascending) of a[0, cnt) over the padded length P (a power of
// two >= cnt), every merge in its "flip" form so that all compare-exchanges put
// the smaller value first: partners at or past cnt hold +inf and are skipped.
// Block-cooperative; a is LDS or global memory of this block alone.
__device__ __forceinline__ void cas(uint32_t* a, uint32_t i, uint32_t l, uint32_t cnt) {
  if (l < cnt) {
    const uint32_t x = a[i], y = a[l];
    if (x > y) {
      a[i] = y;
      a[l] = x;
    }
  }
}
__device__ void block_bitonic(uint32_t* a, uint32_t cnt, uint32_t P) {
  const uint32_t half = P >> 1;
  for (uint32_t k = 2; k <= P; k <<= 1) {
    const uint32_t hk = k >> 1;
    for (uint32_t t = threadIdx.x; t < half; t += blockDim.x) {
      const uint32_t b = (t / hk) * k, o = t % hk;
      cas(a, b + o, b + k - 1u - o, cnt);
    }
    __syncthreads();
    for (uint32_t j = k >> 2; j >= 1u; j >>= 1) {
      for (uint32_t t = threadIdx.x; t < half; t += blockDim.x) {
        const uint32_t i = (t / j) * (2u * j) + (t % j);
        cas(a, i, i + j, cnt);
      }
      __syncthreads();
    }
  }
}

__device__ void sort_crowded(uint32_t* __restrict__ sorted, const uint32_t* __restrict__ cell_count,
                             const uint32_t* __restrict__ cell_off, uint32_t c, uint32_t* s_buf) {
  const uint32_t cnt = cell_count[c];
  uint32_t* seg = sorted + (cell_off[c] - cnt);
  uint32_t P = 1;
  while (P < cnt) P <<= 1;
  if (cnt <= kLdsSort) {
    for (uint32_t k = threadIdx.x; k < P; k += blockDim.x) s_buf[k] = k < cnt ? seg[k] : 0xFFFFFFFFu;
    __syncthreads();
    block_bitonic(s_buf, P, P);
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) seg[k] = s_buf[k];
  } else {
    block_bitonic(seg, cnt, P);
  }
  __syncthreads();
}

// The crowded cells listed by k_hash_stabilize, one per block; when the list
// overflowed, every block sweeps its share of all cells instead.
__global__ __launch_bounds__(kCrowdThreads) void k_hash_crowded(const uint32_t* __restrict__ cell_count,
                                                                const uint32_t* __restrict__ cell_off, uint32_t nc,
                                                                uint32_t* __restrict__ sorted,
                                                                const uint32_t* __restrict__ scratch) {
  __shared__ uint32_t s_buf[kLdsSort];
  __shared__ uint32_t s_list[kCrowdThreads];
  __shared__ uint32_t s_n;
  const uint32_t m = scratch[kOffCrowdCount];
  if (m <= kCrowdCap) {
    for (uint32_t e = blockIdx.x; e < m; e += gridDim.x) sort_crowded(sorted, cell_count, cell_off,
                                                                      scratch[kOffCrowdList + e], s_buf);
    return;
  }
  for (uint32_t base = blockIdx.x * kCrowdThreads; base < nc; base += gridDim.x * kCrowdThreads) {
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const uint32_t c = base + threadIdx.x;
    if (c < nc && cell_count[c] > kSmall) s_list[atomicAdd(&s_n, 1u)] = c;
    __syncthreads();
    const uint32_t k = s_n;
    for (uint32_t e = 0; e < k; e++) sort_crowded(sorted, cell_count, cell_off, s_list[e], s_buf);
    __syncthreads();
  }
}

}  // namespace nbh

namespace nb {

uint32_t hash_scratch_words() { return nbh::kScratchWords; }

// Checks the arguments of a hash build (no device work); NB_OK or an error.
int validate_hash_build(const nb_hash_grid* grid, const float* pred_pos4, uint32_t n, const uint32_t* cell_count,
                        const uint32_t* cell_offset, const uint32_t* sorted) {
  if (!grid) return set_error(NB_ERR_ARG, "nb_hash_build: null grid");
  uint32_t maxr = 1;
  for (int a = 0; a < 3; a++) {
    if (grid->grid_res[a] == 0u || grid->grid_res[a] > 1024u)
      return set_error(NB_ERR_ARG, "nb_hash_build: grid_res[%d] = %u is outside [1, 1024]", a, grid->grid_res[a]);
    maxr = grid->grid_res[a] > maxr ? grid->grid_res[a] : maxr;
  }
  uint32_t cube = 1;
  while (cube < maxr) cube <<= 1;
  const uint64_t nc64 = uint64_t(cube) * cube * cube;
  if (uint64_t(grid->hash_cells) != nc64)
    return set_error(NB_ERR_ARG, "nb_hash_build: hash_cells = %u, expected %llu (the Morton cube of grid_res)",
                     grid->hash_cells, (unsigned long long)nc64);
  if (!(grid->cell_size > 0.f)) return set_error(NB_ERR_ARG, "nb_hash_build: cell_size must be positive");
  if (!cell_count || !cell_offset || (n && (!pred_pos4 || !sorted)))
    return set_error(NB_ERR_ARG, "nb_hash_build: null buffer");
  if (reinterpret_cast<uintptr_t>(pred_pos4) & 15u)
    return set_error(NB_ERR_ARG, "nb_hash_build: pred_pos4 must be 16-byte aligned (float4)");
  return NB_OK;
}

// Enqueues a validated hash build with the given scratch (hash_scratch_words()
// uint32 of device memory).
int enqueue_hash_build(const nb_hash_grid* grid, const float* pred_pos4, const uint32_t* type_flag, uint32_t n,
                       uint32_t* cell_count, uint32_t* cell_offset, uint32_t* sorted, uint32_t* scratch,
                       hipStream_t s) {
  using namespace nbh;
  const uint32_t nc = grid->hash_cells;
  Geo g{};
  for (int a = 0; a < 3; a++) {
    g.wmin[a] = grid->world_min[a];
    g.gmax[a] = float(grid->grid_res[a]) - 1.f;
  }
  g.h = grid->cell_size;
  const float4* pred = reinterpret_cast<const float4*>(pred_pos4);
  const uint32_t nchunks = (nc + kChunk - 1u) / kChunk;
  const uint32_t pblocks = uint32_t((uint64_t(n) + kThreads - 1u) / kThreads);
  const uint32_t cblocks = (nc + kThreads - 1u) / kThreads;
  hipLaunchKernelGGL(k_hash_clear, dim3(std::min<uint32_t>((nc / 4u + kThreads - 1u) / kThreads + 1u, 8192u)),
                     dim3(kThreads), 0, s, cell_count, nc, scratch);
  if (n) hipLaunchKernelGGL(k_hash_count, dim3(pblocks), dim3(kThreads), 0, s, g, pred, type_flag, n, cell_count);
  hipLaunchKernelGGL(k_hash_chunk_sums, dim3(nchunks), dim3(kThreads), 0, s, cell_count, nc, scratch);
  hipLaunchKernelGGL(k_hash_chunk_scan, dim3(1), dim3(1024), 0, s, scratch, nchunks);
  hipLaunchKernelGGL(k_hash_offsets, dim3(nchunks), dim3(kThreads), 0, s, cell_count, nc, scratch, cell_offset);
  if (n) {
    hipLaunchKernelGGL(k_hash_scatter, dim3(pblocks), dim3(kThreads), 0, s, g, pred, type_flag, n, cell_offset, sorted);
    hipLaunchKernelGGL(k_hash_stabilize, dim3(cblocks), dim3(kThreads), 0, s, cell_count, cell_offset, nc, sorted,
                       scratch);
    hipLaunchKernelGGL(k_hash_crowded, dim3(1024), dim3(kCrowdThreads), 0, s, cell_count, cell_offset, nc, sorted,
                       scratch);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(NB_ERR_HIP, "nb_hash_build: launch failed: %s", hipGetErrorString(e));
  return NB_OK;
}

}  // namespace nb

extern "C" int nb_hash_build(const nb_hash_grid* grid, const float* pred_pos4, const uint32_t* type_flag, uint32_t n,
                             uint32_t* cell_count, uint32_t* cell_offset, uint32_t* sorted, void* hip_stream) {
  const int v = nb::validate_hash_build(grid, pred_pos4, n, cell_count, cell_offset, sorted);
  if (v != NB_OK) return v;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  void* scratch = nullptr;
  hipError_t e = hipMallocAsync(&scratch, size_t(nb::hash_scratch_words()) * 4u, s);
  if (e != hipSuccess) return nb::set_error(NB_ERR_HIP, "nb_hash_build: hipMallocAsync failed: %s", hipGetErrorString(e));
  const int rc = nb::enqueue_hash_build(grid, pred_pos4, type_flag, n, cell_count, cell_offset, sorted,
                                        static_cast<uint32_t*>(scratch), s);
  e = hipFreeAsync(scratch, s);
  if (rc != NB_OK) return rc;
  if (e != hipSuccess) return nb::set_error(NB_ERR_HIP, "nb_hash_build: hipFreeAsync failed: %s", hipGetErrorString(e));
  return NB_OK;
}

// ---- nb_hash_build_noalloc: the same build on a file-scope scratch array ----
// The array lives in every GPU's copy of this code object.  Host code never
// passes its name to a kernel (on the host the name is a stand-in, not a
// device address): the device address comes from hipGetSymbolAddress, once per
// device, with that device current, and is kept per device.
#include <mutex>

namespace nbh_noalloc {

static __device__ __align__(256) uint32_t scratch[nbh::kScratchWords];
static_assert(sizeof(uint32_t) * nbh::kScratchWords == NB_HASH_NOALLOC_SCRATCH_BYTES,
              "NB_HASH_NOALLOC_SCRATCH_BYTES must equal the scratch the build uses");

constexpr int kMaxDevices = 64;
static std::mutex g_mu;
static uint32_t* g_ptr[kMaxDevices] = {};

// The current device's address of `scratch` into *out (looked up on the
// device's first use); NB_OK or an error code with the message set.
static int device_scratch(uint32_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return nb::set_error(NB_ERR_HIP, "nb_hash_build_noalloc: hipGetDevice failed: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return nb::set_error(NB_ERR_RANGE, "nb_hash_build_noalloc: device %d is beyond the %d this library serves", dev,
                         kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nb::set_error(NB_ERR_HIP, "nb_hash_build_noalloc: hipGetSymbolAddress of the scratch array failed: %s",
                           hipGetErrorString(e));
    g_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_ptr[dev];
  return NB_OK;
}

}  // namespace nbh_noalloc

extern "C" int nb_hash_build_noalloc(const nb_hash_grid* grid, const float* pred_pos4, const uint32_t* type_flag,
                                     uint32_t n, uint32_t* cell_count, uint32_t* cell_offset, uint32_t* sorted,
                                     void* hip_stream) {
  // validate_hash_build refuses any hash_cells other than the Morton cube of
  // a grid of at most 1024 per axis, so hash_cells <= NB_HASH_MAX_CELLS, the
  // largest hash the array's chunk sums cover; checked again here because the
  // array's size depends on it.
  const int v = nb::validate_hash_build(grid, pred_pos4, n, cell_count, cell_offset, sorted);
  if (v != NB_OK) return v;
  if (grid->hash_cells > NB_HASH_MAX_CELLS)
    return nb::set_error(NB_ERR_ARG, "nb_hash_build_noalloc: hash_cells = %u exceeds the scratch array's %u",
                         grid->hash_cells, NB_HASH_MAX_CELLS);
  uint32_t* s = nullptr;
  const int rc = nbh_noalloc::device_scratch(&s);
  if (rc != NB_OK) return rc;
  return nb::enqueue_hash_build(grid, pred_pos4, type_flag, n, cell_count, cell_offset, sorted, s,
                                static_cast<hipStream_t>(hip_stream));
}
