"""Mesh collider SDF on the CPU: binary STL I/O, the rigid transform of
transformTriangles (src/core/MeshSDF.h:130-140), and the oracle's
restatement of buildMeshSDF (MeshSDF.h:85-126).

The exact variant is pinned by the closed-form SDF of a box,
|max(q, 0)| + min(max(q), 0) with q = |p - c| - half.  The reference's
variant differs from it only where closestPtOnTriangle lacks a region (the
edge v1-v2); a brute-force distance to one triangle pins that difference."""
import numpy as np

import oracle
from nebula_amd import mesh


def test_binary_stl_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    t = rng.standard_normal((37, 12)).astype(np.float32)
    path = tmp_path / "m.stl"
    mesh.save_binary_stl(path, t)
    assert path.stat().st_size == 84 + 50 * 37
    assert np.array_equal(mesh.load_binary_stl(path), t)


def test_transform_rotates_normals_and_vertices():
    t = np.zeros((1, 12), np.float32)
    t[0, 0:3] = (1, 0, 0)
    t[0, 3:6], t[0, 6:9], t[0, 9:12] = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    s = np.float32(np.sqrt(0.5))
    out = mesh.transform_triangles(t, (0.5, 0.0, -1.0), (s, 0.0, 0.0, s))  # 90 deg about z
    np.testing.assert_allclose(out[0, 0:3], (0, 1, 0), atol=1e-6)
    np.testing.assert_allclose(out[0, 3:6], (0.5, 1, -1), atol=1e-6)
    np.testing.assert_allclose(out[0, 6:9], (-0.5, 0, -1), atol=1e-6)
    np.testing.assert_allclose(out[0, 9:12], (0.5, 0, 0), atol=1e-6)


def _compact(v):
    """Every third bit of v, starting at bit 0 (the inverse of a Morton spread)."""
    v = v & 0x09249249
    v = (v | (v >> 2)) & 0x030C30C3
    v = (v | (v >> 4)) & 0x0300F00F
    v = (v | (v >> 8)) & 0x030000FF
    return (v | (v >> 16)) & 0x000003FF


def test_compact_inverts_the_oracle_morton_encode():
    for x, y, z in ((0, 0, 0), (5, 9, 3), (31, 0, 17), (255, 128, 1)):
        c = oracle.morton_encode(x, y, z)
        assert (_compact(c), _compact(c >> 1), _compact(c >> 2)) == (x, y, z)


def _box_sdf(p, lo, hi):
    c, half = (np.asarray(lo) + np.asarray(hi)) / 2, (np.asarray(hi) - np.asarray(lo)) / 2
    q = np.abs(p - c) - half
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0)


def test_oracle_mesh_sdf_of_a_box_is_the_box_sdf():
    lo, hi = (0.26, 0.31, 0.22), (0.71, 0.52, 0.63)
    res = (20, 18, 16)
    cs = 1.0 / 20
    sdf = oracle.mesh_sdf(mesh.box_mesh(lo, hi), res, cs, exact=True)
    cube = oracle.cube_res_of(res)
    idx = np.arange(cube ** 3)
    x, y, z = [_compact(idx >> k) for k in range(3)]
    inside = (x < res[0]) & (y < res[1]) & (z < res[2])
    assert np.all(sdf[~inside] == np.float32(1e9))
    p = (np.stack([x, y, z], axis=1)[inside].astype(np.float64) + 0.5) * cs
    want = _box_sdf(p, lo, hi)
    np.testing.assert_allclose(sdf[inside], want, atol=2e-6)
    assert np.all(np.sign(sdf[inside][np.abs(want) > 1e-4]) == np.sign(want[np.abs(want) > 1e-4]))


def test_edge_v1_v2_takes_the_plane_projection():
    """closestPtOnTriangle (MeshSDF.h:48-74) has the Voronoi regions of the
    three vertices and of the edges v0-v1 and v0-v2 but none for v1-v2.
    Against the true distance to one triangle, by brute force over a dense
    barycentric sampling, the oracle must agree everywhere except at cells
    whose plane projection lies beyond edge v1-v2, where it returns the
    distance to the plane."""
    a, b, c = np.array([0.3, 0.2, 0.4]), np.array([0.8, 0.3, 0.5]), np.array([0.4, 0.7, 0.45])
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    tri = np.concatenate([n, a, b, c]).astype(np.float32)[None]
    res, cs = (16, 16, 16), 1.0 / 16
    sdf = oracle.mesh_sdf(tri, res, cs)
    exact = oracle.mesh_sdf(tri, res, cs, exact=True)
    ix = np.arange(16)
    x, y, z = np.meshgrid(ix, ix, ix, indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    got = sdf[[oracle.morton_encode(int(i), int(j), int(k)) for i, j, k in zip(x, y, z)]]
    p = (np.stack([x, y, z], 1) + 0.5) * cs
    # dense barycentric sampling of the triangle (spacing 1/400 of an edge)
    m = 400
    u, v = np.meshgrid(np.arange(m + 1), np.arange(m + 1), indexing="ij")
    keep = u + v <= m
    u, v = u[keep] / m, v[keep] / m
    pts = a + u[:, None] * (b - a) + v[:, None] * (c - a)
    true = np.array([np.sqrt(((pts - q) ** 2).sum(1)).min() for q in p])
    plane = (p - a) @ n
    # barycentric weight of v0 of the plane projection: negative beyond edge v1-v2
    e1, e2, r = b - a, c - a, p - plane[:, None] * n - a
    d11, d12, d22 = e1 @ e1, e1 @ e2, e2 @ e2
    s_ = (r @ e1 * d22 - r @ e2 * d12) / (d11 * d22 - d12 ** 2)
    t_ = (r @ e2 * d11 - r @ e1 * d12) / (d11 * d22 - d12 ** 2)
    beyond_bc = 1 - s_ - t_ < 0
    tol = 2e-3  # the sampling's resolution
    off = np.abs(np.abs(got) - true) > tol
    assert off.sum() > 50, "the grid must reach the region beyond edge v1-v2"
    assert np.all(beyond_bc[off])
    np.testing.assert_allclose(np.abs(got[off]), np.abs(plane[off]), atol=1e-5)
    assert np.all(np.sign(got) == np.where(plane >= 0, 1, -1))
    # the exact variant is the brute-force distance everywhere, and equals the
    # reference's wherever the reference is right
    ex = np.array([exact[oracle.morton_encode(int(i), int(j), int(k))] for i, j, k in zip(x, y, z)])
    np.testing.assert_allclose(np.abs(ex), true, atol=tol)
    assert np.array_equal(ex[~beyond_bc], got[~beyond_bc])


def test_reference_variant_breaks_only_outside_a_box():
    """On a closed box the reference's variant is exact at every cell inside
    and wrong only at some cells outside near edges."""
    lo, hi = (0.26, 0.31, 0.22), (0.71, 0.52, 0.63)
    tris = mesh.box_mesh(lo, hi)
    ref = oracle.mesh_sdf(tris, (20, 18, 16), 1.0 / 20)
    ex = oracle.mesh_sdf(tris, (20, 18, 16), 1.0 / 20, exact=True)
    inside = (ex < 0) & (ex > -1e8)
    assert inside.sum() > 100
    assert np.array_equal(ref[inside], ex[inside])
    diff = ref != ex
    assert diff.any() and np.all(np.abs(ref[diff]) <= np.abs(ex[diff]) + 1e-6)
