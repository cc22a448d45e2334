"""The PBF / cloth spatial hash (CPU-only): the oracle's restatement against
an independent brute force, and nb_hash_build's argument checks.

The brute force computes each particle's cell from the definition of cellId
(shaders/common.glsl:100-105) in numpy float32 -- IEEE quotient, clamp to
[0, gridRes - 1], truncation, bit interleave -- and orders the particles
whose typeFlag is not 6 (shaders/hash_count.comp:15, hash_sort.comp:16) by
numpy's stable argsort; END offsets are the cumulative counts."""
import ctypes as C

import numpy as np
import pytest

import oracle
from nebula_amd import hash as nbhash
from nebula_amd import mpm


def _interleave(x, y, z):
    k = np.zeros_like(x, dtype=np.uint64)
    for b in range(10):
        k |= ((x >> b) & 1).astype(np.uint64) << np.uint64(3 * b)
        k |= ((y >> b) & 1).astype(np.uint64) << np.uint64(3 * b + 1)
        k |= ((z >> b) & 1).astype(np.uint64) << np.uint64(3 * b + 2)
    return k.astype(np.uint32)


def brute_force(pred4, type_flag, world_min, cell_size, grid_res):
    p = np.asarray(pred4, np.float32)[:, :3]
    q = (p - np.asarray(world_min, np.float32)) / np.float32(cell_size)
    hi = np.asarray(grid_res, np.float32) - np.float32(1.0)
    g = np.minimum(np.maximum(q, np.float32(0.0)), hi).astype(np.uint32)
    keys = _interleave(g[:, 0], g[:, 1], g[:, 2])
    keep = np.ones(len(p), bool) if type_flag is None else np.asarray(type_flag) != 6
    idx = np.nonzero(keep)[0].astype(np.uint32)
    order = idx[np.argsort(keys[idx], kind="stable")]
    nc = oracle.cube_res_of(grid_res) ** 3
    cnt = np.bincount(keys[idx], minlength=nc).astype(np.uint32)
    return cnt, np.cumsum(cnt, dtype=np.uint64).astype(np.uint32), order


def cases():
    rng = np.random.default_rng(11)
    out = []
    # PBF-like: particles spread over the domain and a margin outside it, some unused slots
    n = 5000
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.uniform(-0.2, 2.2, (n, 3))
    tf = rng.integers(0, 7, n).astype(np.uint32)
    out.append(("spread", p, tf, (0.0, 0.0, 0.0), 2.0 / 32, (32, 32, 32)))
    # non-cubic grid (gridRes below the Morton cube on two axes), offset world
    p = np.zeros((3000, 4), np.float32)
    p[:, :3] = rng.uniform(-1.0, 1.0, (3000, 3)) * np.array([1.0, 0.3, 0.5]) + np.array([0.5, -0.2, 4.0])
    out.append(("box", p, None, (-0.5, -0.5, 3.5), 0.05, (20, 6, 11)))
    # crowded cells: unused slots all at the origin (what typeFlag 6 is for), a dense clump
    p = np.zeros((4000, 4), np.float32)
    tf = np.zeros(4000, np.uint32)
    tf[:1500] = 6
    p[1500:2500, :3] = rng.uniform(0.30, 0.31, (1000, 3))
    p[2500:, :3] = rng.uniform(0.0, 1.0, (1500, 3))
    perm = rng.permutation(4000)
    out.append(("crowded", p[perm], tf[perm], (0.0, 0.0, 0.0), 1.0 / 16, (16, 16, 16)))
    # empty
    out.append(("empty", np.zeros((0, 4), np.float32), None, (0.0, 0.0, 0.0), 0.1, (8, 8, 8)))
    return out


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_oracle_hash_equals_brute_force(case):
    _, p, tf, wmin, h, gr = case
    cnt, end, srt = oracle.hash_build(p, tf, wmin, h, gr)
    bc, be, bs = brute_force(p, tf, wmin, h, gr)
    assert np.array_equal(cnt, bc)
    assert np.array_equal(end, be)
    assert np.array_equal(srt, bs)
    if len(p):
        # every cell's slice is its particles, ascending
        c = np.nonzero(cnt)[0][0]
        sl = srt[end[c] - cnt[c]:end[c]]
        assert np.all(np.diff(sl.astype(np.int64)) > 0)


def test_hash_skips_only_type_six():
    p = np.zeros((8, 4), np.float32)
    tf = np.arange(8, dtype=np.uint32)
    cnt, end, srt = oracle.hash_build(p, tf, (0, 0, 0), 1.0, (4, 4, 4))
    assert end[-1] == 7 and list(srt) == [0, 1, 2, 3, 4, 5, 7]


def _grid(gr=(16, 16, 16), h=1.0 / 16, nc=None):
    g = nbhash.make_grid((0.0, 0.0, 0.0), h, gr)
    if nc is not None:
        g.hash_cells = nc
    return g


@pytest.mark.parametrize("bad", ["hash_cells", "grid_res_zero", "grid_res_big", "cell_size", "null", "align"])
def test_hash_build_rejects_bad_arguments_before_any_device_work(bad):
    """Argument errors are reported with NB_ERR_ARG and a message, before the
    call touches a device (these run without one)."""
    L = nbhash._lib()
    g = _grid()
    buf = (C.c_uint32 * 64)()
    ptr = C.addressof(buf)
    pred, n, cnt = ptr, 4, ptr
    if bad == "hash_cells":
        g.hash_cells = 16 ** 3 + 1
    elif bad == "grid_res_zero":
        g.grid_res[1] = 0
    elif bad == "grid_res_big":
        g = _grid((2048, 4, 4), 0.01, 2048 ** 3 % (1 << 32))
    elif bad == "cell_size":
        g.cell_size = 0.0
    elif bad == "null":
        cnt = None
    elif bad == "align":
        pred = ptr + 4
    rc = L.nb_hash_build(C.byref(g), C.c_void_p(pred), None, n, C.c_void_p(cnt), C.c_void_p(ptr), C.c_void_p(ptr), None)
    assert rc == mpm.NB_ERR_ARG
    assert b"nb_hash_build" in L.nb_last_error()
