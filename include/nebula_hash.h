/*
 * nebula_hash.h — C ABI of the spatial hash that the reference's
 * position-based engines build every step: the PBF fluid
 * (src/engine/FluidEngine.cpp:444-474) and the cloth self-collision pass
 * (src/engine/SimulationEngine.cpp:297-320).  Both record the same six
 * dispatches, zero_cells -> hash_count -> hash_scan_local ->
 * hash_scan_global -> hash_add_base -> hash_sort (shaders/zero_cells.comp,
 * hash_count.comp, hash_scan_local.comp, hash_scan_global.comp,
 * hash_add_base.comp, hash_sort.comp), over buffers the engine owns; their
 * consumers (pbf_density.comp:50-51, solve_density.comp:43-44,
 * coupling_cloth.comp:49-50, ...) read a cell's particles as
 * sorted[cellOffset[c] - cellCount[c] .. cellOffset[c]).
 *
 * Here the six dispatches are one call on caller-owned device buffers.
 * Results are the stable counting sort the dispatches intend:
 *   - a particle's cell is cellId(predP.xyz) (shaders/common.glsl:100-105):
 *     clamp((p - worldMin) / cellSize, 0, gridRes - 1), truncated, Morton-
 *     interleaved; the quotient is the IEEE one;
 *   - particles whose typeFlag is NB_HASH_TYPE_UNUSED are skipped, in the
 *     count and in the sort (hash_count.comp:15, hash_sort.comp:16);
 *   - cell_count[c] = number of particles in cell c, for every c < hash_cells;
 *   - cell_offset[c] = END offset: the exclusive prefix sum of cell_count
 *     plus cell_count[c] (what cellOffset holds after hash_sort);
 *   - sorted[0 .. m), m = cell_offset[hash_cells - 1], holds the particle
 *     indices cell by cell, ascending within a cell.
 * The reference's order within a cell comes from atomics and is not
 * reproducible; ascending index is the one order its CPU mirror defines
 * (tests/helpers/MPMHarness.cpp:155-171).  Its scan also gives wrong offsets
 * once hash_cells exceeds 1024 groups of 256 (hash_scan_global.comp:10-34);
 * this one has no such limit.  The reference sizes cellOffset
 * hash_cells + ceil(hash_cells / 256) and uses the tail as scratch: here
 * cell_offset needs hash_cells entries and nothing past them is touched.
 *
 * Status codes and nb_last_error() as in nebula_mpm.h.
 */
#ifndef NEBULA_HASH_H
#define NEBULA_HASH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* typeFlag of the slots the hash skips (TYPE_FLAG_UNUSED, hash_count.comp:15) */
#define NB_HASH_TYPE_UNUSED 6u
/* largest hash_cells: Morton keys interleave 10 bits per axis (cube 1024^3) */
#define NB_HASH_MAX_CELLS (1u << 30)

/* The hash geometry, SimPC's worldMin, cellSize, gridRes and hashCells
 * (shaders/common.glsl:14-52).  hash_cells must be cube^3 with cube the
 * smallest power of two >= max(grid_res) (src/core/Domain.h:26-37). */
typedef struct nb_hash_grid {
  float world_min[3];
  float cell_size;
  uint32_t grid_res[3];
  uint32_t hash_cells;
} nb_hash_grid;

/* Build the hash of n particles.  Device pointers: pred_pos4 (n float4,
 * predP; w unused), type_flag (n uint32, or NULL: no particle skipped),
 * cell_count and cell_offset (hash_cells uint32 each), sorted (n uint32).
 * Enqueued on hip_stream (a hipStream_t; NULL = the default stream) and
 * asynchronous.  Scratch: ceil(hash_cells / 4096) uint32 of per-block
 * sums, taken from the stream-ordered allocator (hipMallocAsync) and
 * returned to it in the same call.  n = 0 gives zero counts and offsets. */
int nb_hash_build(const nb_hash_grid* grid, const float* pred_pos4, const uint32_t* type_flag, uint32_t n,
                  uint32_t* cell_count, uint32_t* cell_offset, uint32_t* sorted, void* hip_stream);

/* Bytes of the per-GPU scratch array nb_hash_build_noalloc uses: the chunk
 * sums of the largest hash (NB_HASH_MAX_CELLS / 4096 words) plus the list of
 * crowded cells (65,600 words), 1.25 MiB. */
#define NB_HASH_NOALLOC_SCRATCH_BYTES 1310976u

/* nb_hash_build that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results, limits and errors as nb_hash_build; the scratch is a
 * file-scope device array of NB_HASH_NOALLOC_SCRATCH_BYTES that each GPU
 * holds from the first call on it until the process ends, instead of a
 * hipMallocAsync.  Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (on streams not ordered with each other) race; calls
 * on one stream, or on synchronised streams, do not.  A hash larger than the
 * array serves (hash_cells > NB_HASH_MAX_CELLS) is refused with NB_ERR_ARG. */
int nb_hash_build_noalloc(const nb_hash_grid* grid, const float* pred_pos4, const uint32_t* type_flag, uint32_t n,
                          uint32_t* cell_count, uint32_t* cell_offset, uint32_t* sorted, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* NEBULA_HASH_H */
