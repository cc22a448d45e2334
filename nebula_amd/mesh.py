"""Triangle meshes for the mesh collider SDF (src/core/MeshSDF.h).

Triangles are (n, 12) float32 rows: normal, v0, v1, v2 — the order of the
reference's MeshTriangle and of a binary STL record.  ``load_binary_stl``
reads the format loadBinarySTL reads (MeshSDF.h:19-40: 80-byte header,
uint32 count, per triangle 12 floats and a uint16 attribute);
``transform_triangles`` is transformTriangles (MeshSDF.h:130-140): normals
rotated, vertices rotated and translated.  MPMEngine.setColliderMesh builds
the SDF on the GPU (nb_mpm_set_collider_mesh)."""
import numpy as np

_RECORD = np.dtype([("t", "<f4", (12,)), ("attr", "<u2")])


def load_binary_stl(path) -> np.ndarray:
    with open(path, "rb") as f:
        f.read(80)
        count = int(np.frombuffer(f.read(4), "<u4")[0])
        rec = np.frombuffer(f.read(count * _RECORD.itemsize), _RECORD, count)
    return np.ascontiguousarray(rec["t"], np.float32)


def save_binary_stl(path, tris: np.ndarray, header: bytes = b"nebula_amd mesh"):
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 12)
    rec = np.zeros(t.shape[0], _RECORD)
    rec["t"] = t
    with open(path, "wb") as f:
        f.write(header[:80].ljust(80, b"\0"))
        f.write(np.array([t.shape[0]], "<u4").tobytes())
        f.write(rec.tobytes())


def _quat_rotate(q, v):
    """Rotate float32 vectors v (..., 3) by the unit quaternion q = (w, x, y, z):
    v + 2 (w (u x v) + u x (u x v)), u = (x, y, z)."""
    w = np.float32(q[0])
    u = np.asarray(q[1:], np.float32)
    uv = np.cross(u, v).astype(np.float32)
    uuv = np.cross(u, uv).astype(np.float32)
    return (v + (uv * w + uuv) * np.float32(2)).astype(np.float32)


def transform_triangles(tris: np.ndarray, translation=(0.0, 0.0, 0.0), rotation=(1.0, 0.0, 0.0, 0.0)) -> np.ndarray:
    """Rigid motion of the triangles: rotation (w, x, y, z) then translation."""
    t = np.asarray(tris, np.float32).reshape(-1, 12)
    out = np.empty_like(t)
    out[:, 0:3] = _quat_rotate(rotation, t[:, 0:3])
    tr = np.asarray(translation, np.float32)
    for k in range(3):
        out[:, 3 + 3 * k:6 + 3 * k] = _quat_rotate(rotation, t[:, 3 + 3 * k:6 + 3 * k]) + tr
    return out


def box_mesh(lo, hi) -> np.ndarray:
    """The 12 outward-facing triangles of an axis-aligned box."""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    c = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)], np.float32)
    quads = [((0, 2, 6, 4), (-1, 0, 0)), ((1, 5, 7, 3), (1, 0, 0)), ((0, 4, 5, 1), (0, -1, 0)),
             ((2, 3, 7, 6), (0, 1, 0)), ((0, 1, 3, 2), (0, 0, -1)), ((4, 6, 7, 5), (0, 0, 1))]
    out = []
    for (a, b, cc, d), n in quads:
        nv = np.asarray(n, np.float32)
        for i, j, k in ((a, b, cc), (a, cc, d)):
            if np.dot(np.cross(c[j] - c[i], c[k] - c[i]), nv) < 0:  # counter-clockwise seen from outside
                j, k = k, j
            out.append(np.concatenate([nv, c[i], c[j], c[k]]))
    return np.asarray(out, np.float32)
