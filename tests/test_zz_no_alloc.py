"""nb_hash_build_noalloc (the hash build that allocates nothing) against
nb_hash_build on the same inputs, bit for bit, through the C ABI: two calls in
a row on the GPU's one scratch array, the second with a smaller input and a
smaller hash, both with crowded cells (the scratch's crowded-cell list) and
unused slots; and the argument checks, which refuse before any device work."""
import numpy as np
import pytest

from nebula_amd import hip, mpm
from nebula_amd.hash import build, make_grid

pytestmark = pytest.mark.gpu


def _scene(seed, n, res, crowded):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.uniform(-0.02, 1.02, (n, 3))
    # a few crowded cells: sizes for the register, LDS and global-memory sorts
    at = 0
    for k, size in enumerate(crowded):
        c = (k + 1.3) / (len(crowded) + 2)
        p[at:at + size, :3] = c + rng.uniform(0, 0.2 / res, (size, 3))
        at += size
    p = p[rng.permutation(n)]
    tf = rng.integers(0, 7, n).astype(np.uint32)
    return p, tf, res


def _run(p, tf, res, noalloc):
    n = len(p)
    grid = make_grid((0, 0, 0), 1.0 / res, (res, res, res))
    nc = int(grid.hash_cells)
    dp = hip.DeviceArray.from_host(p)
    dt = hip.DeviceArray.from_host(tf)
    cnt = hip.DeviceArray(nc, np.uint32)
    off = hip.DeviceArray(nc, np.uint32)
    srt = hip.DeviceArray(n, np.uint32)
    cnt.upload(np.full(nc, 0xABCD, np.uint32))  # stale values must not survive
    build(grid, dp.ptr, dt.ptr, n, cnt.ptr, off.ptr, srt.ptr, noalloc=noalloc)
    hip.synchronize()
    c, o, s = cnt.download(), off.download(), srt.download()
    m = int(o[-1])
    for a in (dp, dt, cnt, off, srt):
        a.free()
    return c, o, s[:m]


def test_noalloc_hash_equals_allocating_build_twice(gpu_lib):
    cases = [("large", _scene(31, 400_000, 64, (17, 300, 17000))),
             ("small", _scene(32, 30_000, 40, (20, 40)))]
    for name, (p, tf, res) in cases:
        print(f"  noalloc hash {name}: {len(p)} particles, grid {res}^3", flush=True)
        ref = _run(p, tf, res, noalloc=False)
        got = _run(p, tf, res, noalloc=True)
        for a, b, what in zip(got, ref, ("cell_count", "cell_offset", "sorted")):
            assert a.shape == b.shape and np.array_equal(a, b), f"{name}: {what} differs"
        # a crowded cell beyond the register sort (> 16), and every used slot hashed
        assert ref[0].max() > 16 and int(ref[1][-1]) == int(np.sum(tf != 6))


def test_noalloc_hash_refuses_bad_arguments(gpu_lib):
    grid = make_grid((0, 0, 0), 0.1, (10, 10, 10))
    grid.hash_cells = 1000  # not the Morton cube (4096)
    d = hip.DeviceArray(4096, np.uint32)
    with pytest.raises(mpm.NebulaError, match="nb_hash_build_noalloc.*hash_cells"):
        build(grid, d.ptr, None, 16, d.ptr, d.ptr, d.ptr, noalloc=True)
    d.free()
