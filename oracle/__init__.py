"""CPU oracle of the MPM substep — TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product library
(``nebula_amd``) never does, and its kernels never route through it.

``mpm_oracle.hpp`` restates the reference's GLSL substep (shaders/mpm_*.comp,
mpm_common.glsl, collider_common.glsl, mpm_nanovdb_bc.comp) in C++, templated
on the arithmetic type: ``precision=32`` is the reference-semantics FP32 oracle,
``precision=64`` the FP64 twin.  ``make -C oracle`` builds
``oracle/build/liboracle_mpm.so``, which this module loads through ctypes.

Parity status: every function is pinned by a test in ``tests/test_oracle*.py``
(known-answer tests restated from the reference's tests/test_mpm.cpp and
tests/test_domain.cpp, closed forms, numpy LAPACK, brute force).  The
degenerate SVD branches (the ``sign(0)`` Jacobi stall and the ``sigma <= 1e-8``
unit vectors) and G2P's Jacobian reset are pinned by results worked by hand
from the GLSL text on inputs that reach them (``degenerate_svd_cases`` and
``jacobian_reset_state`` in tests/nbtest.py).  The analytic collider SDFs,
applyColliderBC and its use in the grid update, the dense-SDF node BC, G2P's
SDF push-out, the FLIP blend, the fluid and the viscoplastic-mud stress are
pinned by closed forms computed in numpy from the GLSL definitions
(tests/test_oracle_bc_kats.py).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle_mpm.so")

_lib = None


class OrcScene(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("cell_size", C.c_float),
        ("grid_res", C.c_uint32 * 3),
        ("world_min", C.c_float * 3),
        ("world_max", C.c_float * 3),
        ("flip_ratio", C.c_float),
        ("cube_res", C.c_uint32),
        ("material_count", C.c_uint32),
        ("collider_count", C.c_uint32),
        ("force_count", C.c_uint32),
        ("has_sdf", C.c_uint32),
        ("corrected_svd", C.c_uint32),
    ]


def build():
    """Compile the oracle with its Makefile (g++)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        L.orc_sim_create.restype = vp
        L.orc_sim_create.argtypes = [C.POINTER(OrcScene), vp, vp, vp, fp, C.c_int]
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sim_set_dt.argtypes = [vp, C.c_float]
        L.orc_sim_set_flip_ratio.argtypes = [vp, C.c_float]
        L.orc_sim_set_corrected_svd.argtypes = [vp, C.c_int]
        L.orc_sim_set_colliders.argtypes = [vp, vp, C.c_uint32]
        L.orc_sim_set_state.argtypes = [vp, fp, C.c_uint32]
        L.orc_sim_get_state.argtypes = [vp, fp]
        L.orc_sim_get_state_f64.argtypes = [vp, dp]
        L.orc_sim_set_state_f64.argtypes = [vp, dp, C.POINTER(C.c_uint32), C.c_uint32]
        for nm in ("orc_sim_bin", "orc_sim_p2g", "orc_sim_grid_update", "orc_sim_sdf_bc", "orc_sim_g2p"):
            getattr(L, nm).argtypes = [vp]
        L.orc_sim_substeps.argtypes = [vp, C.c_uint32]
        L.orc_sim_substeps.restype = C.c_double
        L.orc_sim_get_bins.argtypes = [vp, up, up, up, up]
        L.orc_sim_set_bins.argtypes = [vp, up, up, up]
        L.orc_sim_get_grid.argtypes = [vp, fp, fp]
        L.orc_sim_get_grid_f64.argtypes = [vp, dp, dp]
        L.orc_sim_set_grid.argtypes = [vp, fp, fp]
        L.orc_sim_set_grid_old.argtypes = [vp, fp]
        L.orc_collider_sdf.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.orc_collider_sdf.restype = C.c_double
        L.orc_collider_bc.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_int]
        L.orc_hash_build.argtypes = [fp, up, C.c_uint32, fp, C.c_float, up, C.c_uint32, up, up, up]
        L.orc_hash_build.restype = C.c_uint32
        L.orc_morton_encode.restype = C.c_uint32
        L.orc_morton_encode.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_morton_decode.argtypes = [C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_cell_id.restype = C.c_uint32
        L.orc_cell_id.argtypes = [C.POINTER(OrcScene), C.c_float, C.c_float, C.c_float]
        for nm, t in (("orc_bspline2_f32", C.c_float), ("orc_bspline2g_f32", C.c_float),
                      ("orc_bspline2_f64", C.c_double), ("orc_bspline2g_f64", C.c_double)):
            getattr(L, nm).restype = t
            getattr(L, nm).argtypes = [t]
        L.orc_mesh_sdf.argtypes = [fp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_int, fp]
        L.orc_emit.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, C.c_float, fp, fp]
        L.orc_emit.restype = C.c_uint32
        L.orc_noise_value.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_noise_value.restype = C.c_double
        L.orc_noise_fbm.argtypes = [C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_int]
        L.orc_noise_fbm.restype = C.c_double
        L.orc_noise_curl.argtypes = [C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_int, dp]
        L.orc_noise_hash.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.orc_svd3_f32.argtypes = [fp, fp, fp, fp]
        L.orc_svd3_f64.argtypes = [dp, dp, dp, dp]
        L.orc_hencky_f32.argtypes = [fp, C.c_float, C.c_float, fp]
        L.orc_hencky_f64.argtypes = [dp, C.c_double, C.c_double, dp]
        L.orc_project_vm_f32.argtypes = [fp, C.c_float, C.c_float, C.c_float, fp]
        L.orc_project_vm_f64.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_project_dp_f32.argtypes = [fp, C.c_float, C.c_float, C.c_float, C.c_float, fp]
        L.orc_project_dp_f64.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.orc_dual_tangents.restype = C.c_int
        L.orc_dual_tangents.argtypes = []
        L.orc_dual_substeps.argtypes = [C.POINTER(OrcScene), vp, vp, vp, fp, dp, up, C.c_uint32, dp, dp, C.c_uint32,
                                        dp, dp]
        L.orc_dual32_substeps.argtypes = [C.POINTER(OrcScene), vp, vp, vp, fp, dp, up, C.c_uint32, dp, dp, C.c_uint32,
                                        dp, dp]
        _lib = L
    return _lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32)) if a is not None else None


def cube_res_of(grid_res):
    m = max(max(grid_res), 1)
    p = 1
    while p < m:
        p <<= 1
    return p


def make_scene(dt, cell_size, grid_res, world_min=(0.0, 0.0, 0.0), world_max=None, flip_ratio=0.0):
    sc = OrcScene()
    sc.dt = dt
    sc.cell_size = cell_size
    for a in range(3):
        sc.grid_res[a] = int(grid_res[a])
        sc.world_min[a] = float(world_min[a])
        sc.world_max[a] = float(world_max[a]) if world_max is not None else float(world_min[a] + grid_res[a] * cell_size)
    sc.flip_ratio = flip_ratio
    sc.cube_res = cube_res_of(grid_res)
    return sc


class Sim:
    """One oracle simulation: the reference's substep on host arrays.

    ``materials``/``colliders``/``forces`` are raw bytes of the 64/64/32-byte
    records (the reference's MaterialParams / ColliderPrimitive / ForceGPU).
    ``fields`` is an (8, n, 4) float32 array: P, v, F0, F1, F2, B0, B1, B2.
    """

    def __init__(self, scene, materials: bytes, colliders: bytes = b"", forces: bytes = b"", sdf=None,
                 precision=32):
        L = lib()
        sc = OrcScene.from_buffer_copy(scene)
        sc.material_count = len(materials) // 64
        sc.collider_count = len(colliders) // 64
        sc.force_count = len(forces) // 32
        sc.has_sdf = 1 if sdf is not None else 0
        self.scene = sc
        self.precision = precision
        self.NC = sc.cube_res ** 3
        self._mats = C.create_string_buffer(materials, max(len(materials), 1))
        self._cols = C.create_string_buffer(colliders, max(len(colliders), 1))
        self._forces = C.create_string_buffer(forces, max(len(forces), 1))
        self._sdf = np.ascontiguousarray(sdf, dtype=np.float32) if sdf is not None else None
        self.h = L.orc_sim_create(C.byref(sc), self._mats, self._cols, self._forces,
                                  _fptr(self._sdf) if self._sdf is not None else None, precision)
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_sim_destroy(self.h)
            self.h = None

    def set_dt(self, dt):
        lib().orc_sim_set_dt(self.h, dt)

    def set_flip_ratio(self, r):
        lib().orc_sim_set_flip_ratio(self.h, r)

    def set_corrected_svd(self, on: bool):
        lib().orc_sim_set_corrected_svd(self.h, 1 if on else 0)

    def set_colliders(self, colliders: bytes):
        buf = C.create_string_buffer(colliders, max(len(colliders), 1))
        lib().orc_sim_set_colliders(self.h, buf, len(colliders) // 64)

    def set_state(self, fields):
        f = np.ascontiguousarray(fields, dtype=np.float32)
        assert f.ndim == 3 and f.shape[0] == 8 and f.shape[2] == 4
        self.n = f.shape[1]
        lib().orc_sim_set_state(self.h, _fptr(f), self.n)

    def set_state_f64(self, fields, matid):
        """FP64 state (precision 64): fields (8, n, 4) float64, matid (n,) uint32."""
        assert self.precision == 64
        f = np.ascontiguousarray(fields, dtype=np.float64)
        m = np.ascontiguousarray(matid, dtype=np.uint32)
        self.n = f.shape[1]
        lib().orc_sim_set_state_f64(self.h, _dptr(f), m.ctypes.data_as(C.POINTER(C.c_uint32)), self.n)

    def get_state(self):
        out = np.empty((8, self.n, 4), np.float32)
        lib().orc_sim_get_state(self.h, _fptr(out))
        return out

    def get_state_f64(self):
        assert self.precision == 64
        out = np.empty((8, self.n, 4), np.float64)
        lib().orc_sim_get_state_f64(self.h, _dptr(out))
        return out

    def bin(self):
        lib().orc_sim_bin(self.h)

    def get_bins(self):
        keys = np.empty(self.n, np.uint32)
        srt = np.empty(self.n, np.uint32)
        cnt = np.empty(self.NC, np.uint32)
        end = np.empty(self.NC, np.uint32)
        lib().orc_sim_get_bins(self.h, _uptr(keys), _uptr(srt), _uptr(cnt), _uptr(end))
        return keys, srt, cnt, end

    def set_bins(self, sorted_idx, cell_count, cell_end):
        s = np.ascontiguousarray(sorted_idx, np.uint32)
        c = np.ascontiguousarray(cell_count, np.uint32)
        e = np.ascontiguousarray(cell_end, np.uint32)
        lib().orc_sim_set_bins(self.h, _uptr(s), _uptr(c), _uptr(e))

    def p2g(self):
        lib().orc_sim_p2g(self.h)

    def grid_update(self):
        lib().orc_sim_grid_update(self.h)

    def sdf_bc(self):
        lib().orc_sim_sdf_bc(self.h)

    def g2p(self):
        lib().orc_sim_g2p(self.h)

    def substeps(self, n):
        """Run n full substeps; returns the wall time in seconds."""
        return lib().orc_sim_substeps(self.h, n)

    def get_grid(self):
        mom = np.empty((2 * self.NC, 4), np.float32)
        mass = np.empty(self.NC, np.float32)
        lib().orc_sim_get_grid(self.h, _fptr(mom), _fptr(mass))
        return mom, mass

    def get_grid_f64(self):
        mom = np.empty((2 * self.NC, 4), np.float64)
        mass = np.empty(self.NC, np.float64)
        lib().orc_sim_get_grid_f64(self.h, _dptr(mom), _dptr(mass))
        return mom, mass

    def set_grid(self, mom4, mass):
        m = np.ascontiguousarray(mom4, np.float32).reshape(-1, 4)[: self.NC]
        ms = np.ascontiguousarray(mass, np.float32)
        lib().orc_sim_set_grid(self.h, _fptr(np.ascontiguousarray(m)), _fptr(ms))

    def set_grid_old(self, old4):
        """The saved pre-force velocities (the [NC, 2NC) half FLIP reads)."""
        o = np.ascontiguousarray(np.asarray(old4, np.float32).reshape(-1, 4)[: self.NC])
        assert o.shape[0] == self.NC
        lib().orc_sim_set_grid_old(self.h, _fptr(o))


# ---- scalar helpers (property tests) ----

def morton_encode(x, y, z):
    return lib().orc_morton_encode(x, y, z)


def morton_decode(c):
    x, y, z = C.c_int(), C.c_int(), C.c_int()
    lib().orc_morton_decode(c, C.byref(x), C.byref(y), C.byref(z))
    return x.value, y.value, z.value


def cell_id(scene, x, y, z):
    return lib().orc_cell_id(C.byref(scene), x, y, z)


def bspline2(d, precision=32):
    return lib().orc_bspline2_f32(d) if precision == 32 else lib().orc_bspline2_f64(d)


def bspline2g(d, precision=32):
    return lib().orc_bspline2g_f32(d) if precision == 32 else lib().orc_bspline2g_f64(d)


def _m(a, dt):
    """numpy (3,3) row-major matrix -> column-major flat array."""
    return np.ascontiguousarray(np.asarray(a, dt).T.reshape(9))


def _unm(a):
    return a.reshape(3, 3).T.copy()


def svd3(F, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    f = _m(F, dt)
    U = np.empty(9, dt)
    V = np.empty(9, dt)
    s = np.empty(3, dt)
    if precision == 32:
        lib().orc_svd3_f32(_fptr(f), _fptr(U), _fptr(s), _fptr(V))
    else:
        lib().orc_svd3_f64(_dptr(f), _dptr(U), _dptr(s), _dptr(V))
    return _unm(U), s, _unm(V)


def hencky_stress(F, mu, lam, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    f = _m(F, dt)
    t = np.empty(9, dt)
    if precision == 32:
        lib().orc_hencky_f32(_fptr(f), mu, lam, _fptr(t))
    else:
        lib().orc_hencky_f64(_dptr(f), mu, lam, _dptr(t))
    return _unm(t)


def project_von_mises(F, mu, lam, q_max, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    f = _m(F, dt)
    o = np.empty(9, dt)
    if precision == 32:
        lib().orc_project_vm_f32(_fptr(f), mu, lam, q_max, _fptr(o))
    else:
        lib().orc_project_vm_f64(_dptr(f), mu, lam, q_max, _dptr(o))
    return _unm(o)


def project_drucker_prager(F, mu, lam, M, q_c, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    f = _m(F, dt)
    o = np.empty(9, dt)
    if precision == 32:
        lib().orc_project_dp_f32(_fptr(f), mu, lam, M, q_c, _fptr(o))
    else:
        lib().orc_project_dp_f64(_dptr(f), mu, lam, M, q_c, _dptr(o))
    return _unm(o)


def collider_sdf(record: bytes, x, precision=32):
    """(sdf, outward normal) of one 64-byte collider record at point x."""
    buf = C.create_string_buffer(bytes(record), 64)
    n = np.zeros(3, np.float64)
    d = lib().orc_collider_sdf(buf, float(x[0]), float(x[1]), float(x[2]), precision, _dptr(n))
    return d, n


def collider_bc(v, n, vc, restitution, friction, precision=32):
    """applyColliderBC on one velocity; returns the new velocity."""
    vv = np.array(v, np.float64)
    nn = np.array(n, np.float64)
    cc = np.array(vc, np.float64)
    lib().orc_collider_bc(_dptr(vv), _dptr(nn), _dptr(cc), float(restitution), float(friction), precision)
    return vv


def hash_build(pred4, type_flag, world_min, cell_size, grid_res):
    """The PBF / cloth spatial hash of (n, 4) float32 positions: (cell_count,
    cell_end, sorted[:m]) with cell_end the END offsets (hash_cells = the
    Morton cube of grid_res); particles with type_flag 6 are skipped."""
    p = np.ascontiguousarray(pred4, np.float32).reshape(-1, 4)
    n = p.shape[0]
    tf = None if type_flag is None else np.ascontiguousarray(type_flag, np.uint32)
    cube = cube_res_of(grid_res)
    nc = cube ** 3
    cnt = np.empty(nc, np.uint32)
    end = np.empty(nc, np.uint32)
    srt = np.zeros(max(n, 1), np.uint32)
    wm = np.ascontiguousarray(world_min, np.float32)
    gr = np.ascontiguousarray(grid_res, np.uint32)
    m = lib().orc_hash_build(_fptr(p), _uptr(tf), n, _fptr(wm), float(cell_size), _uptr(gr), nc, _uptr(cnt),
                             _uptr(end), _uptr(srt))
    return cnt, end, srt[:m]


# procedural noise of the turbulence / noise forces (shaders/force_noise.glsl)
def noise_value(p, precision=32):
    return lib().orc_noise_value(float(p[0]), float(p[1]), float(p[2]), precision)


def noise_fbm(p, octaves, precision=32):
    return lib().orc_noise_fbm(float(p[0]), float(p[1]), float(p[2]), octaves, precision)


def noise_curl(p, octaves, precision=32):
    out = np.zeros(3, np.float64)
    lib().orc_noise_curl(float(p[0]), float(p[1]), float(p[2]), octaves, precision, _dptr(out))
    return out


def noise_hash(p, precision=32):
    out = np.zeros(3, np.float64)
    lib().orc_noise_hash(float(p[0]), float(p[1]), float(p[2]), precision, _dptr(out))
    return out


def emit(records: bytes, n_emitters: int, n_calls: int, dt: float, capacity: int, live: int, Vp: float):
    """The particles MPMEngine::emitFromEmitters would append over n_calls
    calls: (pos4, vel4) float32 arrays."""
    room = max(capacity - live, 0)
    pos = np.zeros((room, 4), np.float32)
    vel = np.zeros((room, 4), np.float32)
    buf = C.create_string_buffer(bytes(records), max(len(records), 1))
    n = lib().orc_emit(buf, n_emitters, n_calls, dt, capacity, live, Vp, _fptr(pos), _fptr(vel))
    return pos[:n], vel[:n]


def mesh_sdf(tris, grid_res, cell_size, exact=False):
    """buildMeshSDF: the Morton SDF (cube^3 floats) of (n, 12) triangles;
    exact adds the closest-point region of edge v1-v2 the reference lacks."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 12)
    cube = cube_res_of(grid_res)
    out = np.empty(cube ** 3, np.float32)
    lib().orc_mesh_sdf(_fptr(t), t.shape[0], int(grid_res[0]), int(grid_res[1]), int(grid_res[2]), cube,
                       float(cell_size), int(bool(exact)), _fptr(out))
    return out


def dual_tangents():
    """Tangent directions one dual_substeps call carries."""
    return lib().orc_dual_tangents()


def dual_substeps(scene, materials: bytes, fields, matid, tan_fields, tan_mats=None, n_sub=1, colliders: bytes = b"",
                  forces: bytes = b"", sdf=None, corrected_svd=False, precision=64):
    """Forward-mode derivatives of n_sub substeps of the FP64 twin
    (oracle/dual.hpp): the values after n_sub substeps, (8, n, 4) float64, and
    their tangents, (K, 8, n, 4), from the start ``fields`` (8, n, 4) float64,
    material ids ``matid`` (n,) and input tangents ``tan_fields`` (K, 8, n, 4)
    and ``tan_mats`` (K, materials, 4: mu, lambda, M_friction, q_cohesion),
    K = dual_tangents().  Exact derivatives of the FP64 substep map (no step
    size); the material-id lanes carry no tangent.  precision=32: the FP32
    oracle's map, values and tangents in FP32 (Dual32), returned as float64."""
    L = lib()
    K = L.orc_dual_tangents()
    sc = OrcScene.from_buffer_copy(scene)
    nm = len(materials) // 64
    sc.material_count = nm
    sc.collider_count = len(colliders) // 64
    sc.force_count = len(forces) // 32
    sc.has_sdf = 1 if sdf is not None else 0
    sc.corrected_svd = 1 if corrected_svd else 0
    f = np.ascontiguousarray(fields, np.float64)
    n = f.shape[1]
    assert f.shape == (8, n, 4)
    m = np.ascontiguousarray(matid, np.uint32)
    tf = np.ascontiguousarray(tan_fields, np.float64)
    assert tf.shape == (K, 8, n, 4)
    tm = None
    if tan_mats is not None:
        # [material][parameter][tangent]
        tm = np.ascontiguousarray(np.transpose(np.asarray(tan_mats, np.float64), (1, 2, 0)))
        assert tm.shape == (nm, 4, K)
    mb = C.create_string_buffer(materials, max(len(materials), 1))
    cb = C.create_string_buffer(colliders, max(len(colliders), 1))
    fb = C.create_string_buffer(forces, max(len(forces), 1))
    sd = np.ascontiguousarray(sdf, np.float32) if sdf is not None else None
    out = np.empty((8, n, 4), np.float64)
    out_t = np.empty((K, 8, n, 4), np.float64)
    if precision not in (32, 64):
        raise ValueError(f"precision {precision}: 32 or 64")
    fn = L.orc_dual_substeps if precision == 64 else L.orc_dual32_substeps
    fn(C.byref(sc), mb, cb, fb, _fptr(sd) if sd is not None else None, _dptr(f),
                        m.ctypes.data_as(C.POINTER(C.c_uint32)), n, _dptr(tf), _dptr(tm) if tm is not None else None,
                        int(n_sub), _dptr(out), _dptr(out_t))
    return out, out_t
