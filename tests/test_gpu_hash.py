"""The PBF / cloth spatial hash on the GPU (nb_hash_build) against the oracle's
stable counting sort, bit for bit: counts, END offsets and the sorted
particle indices, at sizes from empty to a 256^3 hash.

The cases reach every path of the build: cells sorted by their own thread
(<= 16 particles), crowded cells sorted in LDS (17 .. 16384) and in place in
global memory (> 16384), more crowded cells than the list holds (the dense
sweep), unaligned output buffers, positions outside the domain and NaN, and
unused slots (typeFlag 6).  The oracle itself is pinned against a brute
force in tests/test_hash.py."""
import os

import numpy as np
import pytest

import oracle
from nebula_amd import hip
from nebula_amd.hash import SpatialHash, build, make_grid
from test_hash import cases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hash_pbf.npz")


def _gpu(p, tf, wmin, h, gr):
    sh = SpatialHash(wmin, h, gr, max(len(p), 1))
    sh.build_from_host(p, tf)
    return sh.result()


def _check(name, p, tf, wmin, h, gr):
    print(f"  hash {name}: {len(p)} particles, grid {tuple(gr)}", flush=True)
    g = _gpu(p, tf, wmin, h, gr)
    o = oracle.hash_build(p, tf, wmin, h, gr)
    for a, b, what in zip(g, o, ("cell_count", "cell_offset", "sorted")):
        assert a.shape == b.shape and np.array_equal(a, b), f"{name}: {what} differs"
    return g


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_hash_matches_oracle_small_cases(gpu_lib, case):
    _check(*case)


def test_hash_matches_golden_fixture(gpu_lib):
    d = np.load(GOLDEN, allow_pickle=False)
    gr = tuple(int(x) for x in d["grid_res"])
    cnt, off, srt = _gpu(d["pred"], d["type_flag"], tuple(d["world_min"]), float(d["cell_size"]), gr)
    assert np.array_equal(cnt, d["cell_count"]) and np.array_equal(off, d["cell_end"])
    assert np.array_equal(srt, d["sorted"])


def test_hash_crowded_cells_lds_and_global_paths(gpu_lib):
    rng = np.random.default_rng(21)
    h = 1.0 / 32
    parts = []
    for size, centre in ((17, 0.1), (300, 0.2), (16384, 0.4), (16385, 0.5), (40000, 0.7)):
        q = np.full((size, 3), centre + 0.3 * h) + rng.uniform(0, 0.3 * h, (size, 3))
        parts.append(q)
    parts.append(rng.uniform(0, 1, (20000, 3)))
    p3 = np.concatenate(parts).astype(np.float32)
    p3 = p3[rng.permutation(len(p3))]
    p = np.zeros((len(p3), 4), np.float32)
    p[:, :3] = p3
    cnt, _, _ = _check("crowded", p, None, (0, 0, 0), h, (32, 32, 32))
    assert cnt.max() >= 40000 and np.sum(cnt > 16) >= 5


def test_hash_more_crowded_cells_than_listed(gpu_lib):
    """70,000 cells of 17-20 particles: past the 65,536-entry crowded list,
    the crowded pass sweeps every cell."""
    rng = np.random.default_rng(22)
    res = 64
    cells = rng.choice(res ** 3, 70000, replace=False)
    ijk = np.stack(np.unravel_index(cells, (res, res, res)), axis=1)
    k = rng.integers(17, 21, len(cells))
    centres = np.repeat(ijk, k, axis=0) + rng.uniform(0.05, 0.95, (int(k.sum()), 3))
    p = np.zeros((len(centres), 4), np.float32)
    p[:, :3] = centres / res
    p = p[rng.permutation(len(p))]
    cnt, _, _ = _check("many_crowded", p, None, (0, 0, 0), 1.0 / res, (res, res, res))
    assert np.sum(cnt > 16) == 70000


def test_hash_large_uniform_with_unused_slots(gpu_lib):
    rng = np.random.default_rng(23)
    n = 1_500_000
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.uniform(-0.05, 1.05, (n, 3))
    tf = rng.integers(0, 7, n).astype(np.uint32)
    p[tf == 6, :3] = 0.0  # unused slots sit at the origin, as in the PBF engine
    p[:5, 0] = np.nan
    _check("large_128", p, tf, (0, 0, 0), 1.0 / 128, (128, 128, 128))
    _check("large_256", p, None, (0, 0, 0), 1.0 / 256, (256, 256, 256))


def test_hash_unaligned_outputs_and_repeat_builds(gpu_lib):
    """Output buffers at 4-byte offsets (the scalar paths of the clear and the
    scan); two builds give the same bytes."""
    rng = np.random.default_rng(24)
    n = 50000
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.uniform(0, 1, (n, 3))
    gr = (40, 40, 40)
    grid = make_grid((0, 0, 0), 1.0 / 40, gr)
    nc = int(grid.hash_cells)
    dp = hip.DeviceArray.from_host(p)
    cnt = hip.DeviceArray(nc + 1, np.uint32)
    off = hip.DeviceArray(nc + 1, np.uint32)
    srt = hip.DeviceArray(n, np.uint32)
    outs = []
    for _ in range(2):
        build(grid, dp.ptr, None, n, cnt.ptr + 4, off.ptr + 4, srt.ptr)
        hip.synchronize()
        outs.append((cnt.download()[1:].copy(), off.download()[1:].copy(), srt.download().copy()))
    o = oracle.hash_build(p, None, (0, 0, 0), 1.0 / 40, gr)
    for got in outs:
        assert np.array_equal(got[0], o[0]) and np.array_equal(got[1], o[1]) and np.array_equal(got[2], o[2])


def test_hash_empty_input_zeroes_counts_and_offsets(gpu_lib):
    sh = SpatialHash((0, 0, 0), 0.1, (10, 10, 10), 16)
    sh.cell_count.upload(np.full(sh.hash_cells, 7, np.uint32))
    sh.build(0, None, 0)
    cnt, off, srt = sh.result()
    assert not cnt.any() and not off.any() and srt.size == 0
