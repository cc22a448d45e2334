"""The spatial hash of the reference's position-based engines (PBF fluid,
cloth self-collision) on the GPU: nb_hash_build (include/nebula_hash.h).

The reference's engines own three hash buffers, ``cellCount``,
``cellOffset`` and ``sortedIdx`` (src/engine/FluidEngine.cpp:444-474,
src/engine/SimulationEngine.cpp:297-320), and rebuild them from ``predP``
and ``typeFlag`` every step.  ``SpatialHash`` owns the same three buffers on
the device and rebuilds them with one call; a particle's neighbours in cell
c are ``sorted[offset[c] - count[c] : offset[c]]``, ascending by index.

Geometry as in the reference's SimPC: ``world_min``, ``cell_size`` and
``grid_res`` (cells per axis); the hash spans the Morton cube of grid_res
(src/core/Domain.h:26-37).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import hip, mpm

TYPE_UNUSED = 6  # typeFlag of slots the hash skips (shaders/hash_count.comp:15)


class HashGrid(C.Structure):
    """nb_hash_grid: SimPC's worldMin, cellSize, gridRes and hashCells."""

    _fields_ = [("world_min", C.c_float * 3), ("cell_size", C.c_float), ("grid_res", C.c_uint32 * 3),
                ("hash_cells", C.c_uint32)]


_bound = False


def _lib():
    global _bound
    L = mpm.lib()
    if not _bound:
        vp, up = C.c_void_p, C.POINTER(C.c_uint32)
        L.nb_hash_build.argtypes = [C.POINTER(HashGrid), vp, vp, C.c_uint32, vp, vp, vp, vp]
        L.nb_hash_build_noalloc.argtypes = [C.POINTER(HashGrid), vp, vp, C.c_uint32, vp, vp, vp, vp]
        _bound = True
    return L


def make_grid(world_min: Sequence[float], cell_size: float, grid_res: Sequence[int]) -> HashGrid:
    g = HashGrid()
    g.world_min[:] = [float(x) for x in world_min]
    g.cell_size = float(cell_size)
    g.grid_res[:] = [int(x) for x in grid_res]
    g.hash_cells = mpm.hashCells(tuple(int(x) for x in grid_res))
    return g


def build(grid: HashGrid, pred_ptr: int, type_flag_ptr: Optional[int], n: int, cell_count_ptr: int,
          cell_offset_ptr: int, sorted_ptr: int, stream: Optional[int] = None, noalloc: bool = False):
    """nb_hash_build on device addresses (ints); asynchronous on ``stream``.
    noalloc=True calls nb_hash_build_noalloc, which takes its scratch from a
    per-GPU array instead of allocating it."""
    L = _lib()
    fn, name = (L.nb_hash_build_noalloc, "nb_hash_build_noalloc") if noalloc else (L.nb_hash_build, "nb_hash_build")
    rc = fn(C.byref(grid), C.c_void_p(pred_ptr), C.c_void_p(type_flag_ptr) if type_flag_ptr else None,
            int(n), C.c_void_p(cell_count_ptr), C.c_void_p(cell_offset_ptr),
            C.c_void_p(sorted_ptr) if sorted_ptr else None, C.c_void_p(stream) if stream else None)
    if rc != mpm.NB_OK:
        raise mpm.NebulaError(f"{name} error {rc}: {mpm.lib().nb_last_error().decode()}")


class SpatialHash:
    """The engines' cellCount / cellOffset / sortedIdx buffers on the device,
    for up to ``capacity`` particles."""

    def __init__(self, world_min: Sequence[float], cell_size: float, grid_res: Sequence[int], capacity: int):
        self.grid = make_grid(world_min, cell_size, grid_res)
        self.hash_cells = int(self.grid.hash_cells)
        self.capacity = int(capacity)
        self.cell_count = hip.DeviceArray(self.hash_cells, np.uint32)
        self.cell_offset = hip.DeviceArray(self.hash_cells, np.uint32)
        self.sorted = hip.DeviceArray(max(self.capacity, 1), np.uint32)
        self.n = 0

    def build(self, pred_ptr: int, type_flag_ptr: Optional[int], n: int, stream: Optional[int] = None):
        """Rebuild from n predicted positions (float4, device) and optional
        typeFlags (uint32, device)."""
        if n > self.capacity:
            raise ValueError(f"{n} particles exceed the hash's capacity {self.capacity}")
        build(self.grid, pred_ptr, type_flag_ptr, n, self.cell_count.ptr, self.cell_offset.ptr, self.sorted.ptr,
              stream)
        self.n = n

    def build_from_host(self, pred4: np.ndarray, type_flag: Optional[np.ndarray] = None):
        """Upload (n, 4) float32 positions (and typeFlags), build, synchronise."""
        p = hip.DeviceArray.from_host(np.ascontiguousarray(pred4, np.float32).reshape(-1, 4))
        t = hip.DeviceArray.from_host(np.ascontiguousarray(type_flag, np.uint32)) if type_flag is not None else None
        self.build(p.ptr, t.ptr if t is not None else None, p.shape[0])
        hip.synchronize()
        p.free()
        if t is not None:
            t.free()

    def result(self):
        """(cell_count, cell_offset, sorted[:m]) as numpy arrays; m is the number
        of particles hashed (cell_offset[-1])."""
        hip.synchronize()
        cnt = self.cell_count.download()
        off = self.cell_offset.download()
        srt = self.sorted.download()
        m = int(off[-1]) if off.size else 0
        return cnt, off, srt[:m]
