"""Python mirror of the reference's MPMEngine surface over the C ABI.

``MPMEngine`` keeps the reference's method names and argument meaning
(src/engine/MPMEngine.h:59-186, src/engine/EngineBase.h:23-31) so that code
written against the reference reads the same; every call goes through
``libnebula_mpm.so`` (include/nebula_mpm.h).  There is no CPU fallback: if the
HIP library is missing or no GPU is present the calls raise.

Also here: the POD records (MaterialParams, ColliderPrimitive, ForceGPU) with
the reference's byte layouts, the material presets of
src/core/MaterialParams.h:42-103, ColliderSet (src/core/Collider.h:31-120),
the force classes of src/core/Force.h and the domain helpers of
src/core/Domain.h.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Iterable, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnebula_mpm.so")

NB_OK = 0
# status codes (include/nebula_mpm.h)
NB_ERR_ARG, NB_ERR_RANGE, NB_ERR_HIP, NB_ERR_STATE, NB_ERR_UNSUPPORTED = -1, -2, -3, -4, -5
FIELD_P, FIELD_V, FIELD_F0, FIELD_F1, FIELD_F2, FIELD_B0, FIELD_B1, FIELD_B2 = range(8)
FIELD_NAMES = ("P", "v", "F0", "F1", "F2", "B0", "B1", "B2")
MAX_MATERIALS, MAX_COLLIDERS, MAX_FORCES = 16, 64, 32


class MaterialModel:
    """src/core/MaterialParams.h:5-12"""

    ELASTIC = 0
    VON_MISES = 1
    DRUCKER_PRAGER = 2
    GRANULAR_POWDER = 3
    FLUID = 4
    VISCOPLASTIC_MUD = 5


class ColliderType:
    """src/core/Collider.h:8-13"""

    PLANE = 0
    SPHERE = 1
    BOX = 2
    CAPSULE = 3


class ForceType:
    """src/core/Force.h:13-19"""

    GRAVITY = 0
    CONSTANT_WIND = 1
    TURBULENCE = 2
    NOISE = 3
    Z_CLAMP = 4


class MaterialParams(C.Structure):
    """64-byte material record (src/core/MaterialParams.h:16-38)."""

    _fields_ = [
        ("mu", C.c_float), ("lambda_", C.c_float), ("rho0", C.c_float), ("model", C.c_uint32),
        ("M_friction", C.c_float), ("q_cohesion", C.c_float), ("q_max", C.c_float), ("bulkK", C.c_float),
        ("fluidGamma", C.c_float), ("viscosity", C.c_float), ("hardening", C.c_float), ("xi", C.c_float),
        ("pad", C.c_float * 4),
    ]


class ColliderPrimitive(C.Structure):
    """64-byte analytic collider record (src/core/Collider.h:19-29)."""

    _fields_ = [
        ("type", C.c_uint32), ("px", C.c_float), ("py", C.c_float), ("pz", C.c_float),
        ("nx", C.c_float), ("ny", C.c_float), ("nz", C.c_float), ("radius", C.c_float),
        ("vx", C.c_float), ("vy", C.c_float), ("vz", C.c_float), ("restitution", C.c_float),
        ("friction", C.c_float), ("pad", C.c_float * 3),
    ]


class ForceGPU(C.Structure):
    """32-byte force record (src/core/Force.h:27-32)."""

    _fields_ = [("type", C.c_uint16), ("affectMask", C.c_uint16), ("data", C.c_float * 7)]


class MPMConfig(C.Structure):
    """MPMConfig (src/engine/MPMEngine.h:22-55); defaults from its member initialisers."""

    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("domainSize", C.c_float * 3), ("cellSize", C.c_float), ("maxParticles", C.c_uint32),
        ("E", C.c_float), ("nu", C.c_float), ("rho0", C.c_float),
    ]

    def __init__(self, **kw):
        super().__init__()
        self.nx = self.ny = self.nz = 32
        self.domainSize[:] = (10.0, 10.0, 10.0)
        self.cellSize = 10.0 / 64.0
        self.maxParticles = 0
        self.E, self.nu, self.rho0 = 1e4, 0.3, 1000.0
        for k, v in kw.items():
            if k == "domainSize":
                self.domainSize[:] = tuple(float(x) for x in v)
            else:
                setattr(self, k, v)

    def particleCount(self):
        return self.nx * self.ny * self.nz

    def maxParticleCount(self):
        return self.maxParticles if self.maxParticles > 0 else self.particleCount()

    def gridRes(self):
        return gridRes(tuple(self.domainSize), self.cellSize)

    def totalCells(self):
        return hashCells(self.gridRes())

    def nGroups(self):
        return nGroups(self.totalCells())

    def mu(self):
        return calcMu(self.E, self.nu)

    def lame(self):
        return calcLambda(self.E, self.nu)


class _Knobs(C.Structure):
    _fields_ = [
        ("restitution", C.c_float), ("wall_friction", C.c_float), ("num_substeps", C.c_int32),
        ("plastic_model", C.c_uint32), ("q_max", C.c_float), ("M_friction", C.c_float),
        ("q_cohesion", C.c_float), ("flip_ratio", C.c_float), ("corrected_svd", C.c_uint32),
    ]


# ---------------------------------------------------------------------------
# domain helpers — src/core/Domain.h:14-39
# ---------------------------------------------------------------------------
def gridRes(domainSize: Sequence[float], cellSize: float):
    h = np.float32(max(np.float32(cellSize), np.float32(1e-6)))
    out = []
    for d in domainSize:
        r = math.ceil(float(np.float32(d) / h))
        out.append(max(int(r), 1))
    return tuple(out)


def mortonCubeRes(res: Sequence[int]) -> int:
    m = max(max(res), 1)
    p = 1
    while p < m:
        p <<= 1
    return p


def hashCells(res: Sequence[int]) -> int:
    c = mortonCubeRes(res)
    return c * c * c


def nGroups(cells: int) -> int:
    return (cells + 255) // 256


def mortonEncode(x: int, y: int, z: int) -> int:
    """Morton cell id (shaders/mpm_common.glsl:161-181)."""

    def expand(v):
        v &= 0x3FF
        v = (v | (v << 16)) & 0x030000FF
        v = (v | (v << 8)) & 0x0300F00F
        v = (v | (v << 4)) & 0x030C30C3
        v = (v | (v << 2)) & 0x09249249
        return v

    return expand(x) | (expand(y) << 1) | (expand(z) << 2)


def mortonEncodeArray(x, y, z):
    def expand(v):
        v = np.asarray(v, np.uint32) & np.uint32(0x3FF)
        v = (v | (v << np.uint32(16))) & np.uint32(0x030000FF)
        v = (v | (v << np.uint32(8))) & np.uint32(0x0300F00F)
        v = (v | (v << np.uint32(4))) & np.uint32(0x030C30C3)
        v = (v | (v << np.uint32(2))) & np.uint32(0x09249249)
        return v

    return expand(x) | (expand(y) << np.uint32(1)) | (expand(z) << np.uint32(2))


# ---------------------------------------------------------------------------
# materials — src/core/MaterialParams.h:40-103
# ---------------------------------------------------------------------------
def calcMu(E, nu):
    E, nu = np.float32(E), np.float32(nu)
    return float(E / (np.float32(2.0) * (np.float32(1.0) + nu)))


def calcLambda(E, nu):
    E, nu = np.float32(E), np.float32(nu)
    return float(E * nu / ((np.float32(1.0) + nu) * (np.float32(1.0) - np.float32(2.0) * nu)))


def _mat(**kw) -> MaterialParams:
    m = MaterialParams()
    for k, v in kw.items():
        setattr(m, "lambda_" if k == "lambda" else k, v)
    return m


def presetJelly(E=1e4, nu=0.4, rho0=1000.0):
    return _mat(mu=calcMu(E, nu), lambda_=calcLambda(E, nu), rho0=rho0, model=MaterialModel.ELASTIC)


def presetSand(E=1e5, nu=0.3, rho0=1600.0):
    return _mat(mu=calcMu(E, nu), lambda_=calcLambda(E, nu), rho0=rho0, model=MaterialModel.DRUCKER_PRAGER,
                M_friction=0.577, q_cohesion=0.0)


def presetWater(rho0=1000.0, bulkK=5e3):
    return _mat(mu=0.0, lambda_=0.0, rho0=rho0, model=MaterialModel.FLUID, bulkK=bulkK, fluidGamma=7.0,
                viscosity=0.01)


def presetMud(E=8e4, nu=0.35, rho0=1800.0):
    return _mat(mu=calcMu(E, nu), lambda_=calcLambda(E, nu), rho0=rho0, model=MaterialModel.VISCOPLASTIC_MUD,
                M_friction=0.4, q_cohesion=5e2, viscosity=0.1)


def presetPowder(E=5e4, nu=0.25, rho0=1200.0):
    return _mat(mu=calcMu(E, nu), lambda_=calcLambda(E, nu), rho0=rho0, model=MaterialModel.GRANULAR_POWDER,
                M_friction=0.8, q_cohesion=50.0)


def materials_bytes(mats: Iterable[MaterialParams]) -> bytes:
    return b"".join(bytes(m) for m in mats)


# ---------------------------------------------------------------------------
# colliders — src/core/Collider.h:31-120
# ---------------------------------------------------------------------------
class ColliderSet:
    def __init__(self):
        self._p: List[ColliderPrimitive] = []

    def _add(self, t, pos, n, radius, restitution, friction, vel):
        c = ColliderPrimitive()
        c.type = t
        c.px, c.py, c.pz = (float(x) for x in pos)
        c.nx, c.ny, c.nz = (float(x) for x in n)
        c.radius = radius
        c.vx, c.vy, c.vz = (float(x) for x in vel)
        c.restitution, c.friction = restitution, friction
        self._p.append(c)

    def addPlane(self, point, normal, restitution=0.3, friction=0.0, vel=(0, 0, 0)):
        n = np.asarray(normal, np.float32)
        n = n / np.float32(np.sqrt(np.float32(np.dot(n, n))))
        self._add(ColliderType.PLANE, point, n, 0.0, restitution, friction, vel)

    def addSphere(self, center, radius, restitution=0.3, friction=0.0, vel=(0, 0, 0)):
        self._add(ColliderType.SPHERE, center, (0, 0, 0), radius, restitution, friction, vel)

    def addBox(self, center, halfExt, restitution=0.3, friction=0.0, vel=(0, 0, 0)):
        self._add(ColliderType.BOX, center, halfExt, 0.0, restitution, friction, vel)

    def addCapsule(self, start, axisVec, radius, restitution=0.3, friction=0.0, vel=(0, 0, 0)):
        self._add(ColliderType.CAPSULE, start, axisVec, radius, restitution, friction, vel)

    def clear(self):
        self._p.clear()

    def count(self):
        return len(self._p)

    def empty(self):
        return not self._p

    def data(self):
        return list(self._p)

    def to_bytes(self) -> bytes:
        return b"".join(bytes(c) for c in self._p)


# ---------------------------------------------------------------------------
# forces — src/core/Force.h:48-256 (gravity, wind, turbulence and noise act on the grid)
# ---------------------------------------------------------------------------
class Force:
    type_id = -1
    affectMask = 0

    def pack_data(self, data):
        pass

    def advanceTime(self, dt):
        pass

    def pack(self) -> ForceGPU:
        g = ForceGPU()
        g.type = self.type_id
        g.affectMask = self.affectMask & 0xFFFF
        self.pack_data(g.data)
        return g


class GravityForce(Force):
    type_id = ForceType.GRAVITY

    def __init__(self, direction=(0.0, 0.0, -1.0), strength=9.8):
        self.direction = tuple(direction)
        self.strength = strength
        self.affectMask = 0

    def pack_data(self, data):
        data[0], data[1], data[2] = self.direction
        data[3] = self.strength

    @staticmethod
    def FromDirection(direction, magnitude):
        return GravityForce(direction, magnitude)


class ConstantWindForce(GravityForce):
    type_id = ForceType.CONSTANT_WIND

    def __init__(self, direction=(1.0, 0.0, 0.0), strength=1.0):
        super().__init__(direction, strength)

    @staticmethod
    def FromDirection(direction, magnitude):
        return ConstantWindForce(direction, magnitude)


class TurbulenceForce(Force):
    """Curl-noise wind (src/core/Force.h:180-215): data = strength, frequency,
    octaves (uint bits), seed; advanceTime moves the seed by speed * dt."""
    type_id = ForceType.TURBULENCE

    def __init__(self, strength=1.0, frequency=0.5, octaves=3, seed=0.0, speed=0.3):
        self.strength, self.frequency, self.octaves, self.seed, self.speed = strength, frequency, octaves, seed, speed
        self.affectMask = 0

    def pack_data(self, data):
        data[0] = self.strength
        data[1] = self.frequency
        data[2] = np.array([self.octaves], np.uint32).view(np.float32)[0]
        data[3] = self.seed

    def advanceTime(self, dt):
        self.seed = float(np.float32(self.seed) + np.float32(self.speed) * np.float32(dt))


class NoiseForce(Force):
    """fBm noise along a direction (src/core/Force.h:217-256): data = dir,
    strength, frequency, octaves (uint bits), seed."""
    type_id = ForceType.NOISE

    def __init__(self, direction=(0.0, 1.0, 0.0), strength=1.0, frequency=1.0, octaves=1, seed=0.0, speed=0.0):
        self.direction = tuple(direction)
        self.strength, self.frequency, self.octaves, self.seed, self.speed = strength, frequency, octaves, seed, speed
        self.affectMask = 0

    def pack_data(self, data):
        data[0], data[1], data[2] = self.direction
        data[3] = self.strength
        data[4] = self.frequency
        data[5] = np.array([self.octaves], np.uint32).view(np.float32)[0]
        data[6] = self.seed

    def advanceTime(self, dt):
        self.seed = float(np.float32(self.seed) + np.float32(self.speed) * np.float32(dt))


class ZClampForce(Force):
    """Position constraint; no effect on the MPM grid (the reference's grid
    shader has no post-force marker, src/core/Force.h:75-82)."""

    type_id = ForceType.Z_CLAMP

    def __init__(self, z=0.0):
        self.zValue = z
        self.affectMask = 0

    def pack_data(self, data):
        data[0] = self.zValue


def forces_bytes(forces: Iterable[Force]) -> bytes:
    return b"".join(bytes(f.pack()) for f in forces)


# ---------------------------------------------------------------------------
# the library
# ---------------------------------------------------------------------------
_lib = None

EXPORTED_SYMBOLS = (
    "nb_last_error", "nb_mpm_abi_version", "nb_mpm_config_default", "nb_mpm_knobs_default", "nb_mpm_create",
    "nb_mpm_destroy", "nb_mpm_step", "nb_mpm_substep", "nb_mpm_substeps", "nb_mpm_sync", "nb_mpm_set_stream",
    "nb_mpm_set_knobs",
    "nb_mpm_get_knobs", "nb_mpm_get_config", "nb_mpm_grid_res", "nb_mpm_total_cells",
    "nb_mpm_live_particle_count", "nb_mpm_set_materials", "nb_mpm_set_particle_material_ids",
    "nb_mpm_set_colliders", "nb_mpm_clear_analytic_colliders", "nb_mpm_set_collider_sdf",
    "nb_mpm_set_collider_sphere", "nb_mpm_clear_collider", "nb_mpm_set_forces", "nb_mpm_append_particles",
    "nb_mpm_reset_particles", "nb_mpm_get_field", "nb_mpm_set_field", "nb_mpm_field_device_ptr",
    "nb_mpm_set_live_particle_count", "nb_mpm_run_bin", "nb_mpm_get_bins", "nb_mpm_run_p2g",
    "nb_mpm_run_grid_update", "nb_mpm_run_g2p", "nb_mpm_get_grid", "nb_mpm_set_grid",
    "nb_mpm_profile_substeps", "nb_mpm_calibration_copy", "nb_mpm_active_node_count", "nb_mpm_active_tile_count",
    "nb_mpm_adjoint_begin", "nb_mpm_adjoint_begin_tapes", "nb_mpm_adjoint_tapes", "nb_mpm_forward_recorded", "nb_mpm_set_adjoint", "nb_mpm_get_adjoint",
    "nb_mpm_adjoint_device_ptr", "nb_mpm_backward", "nb_mpm_adjoint_end",
    "nb_mpm_create_batch", "nb_mpm_scene_count", "nb_mpm_scene_info",
    "nb_mpm_get_tuning", "nb_mpm_set_tuning", "nb_mpm_autotune", "nb_mpm_graph_substeps",
    "nb_mpm_add_emitter", "nb_mpm_clear_emitters", "nb_mpm_emit", "nb_mpm_get_emitter",
    "nb_mpm_set_collider_mesh", "nb_mpm_get_collider_sdf",
    "nb_mpm_set_svd_fallback", "nb_mpm_get_svd_fallback", "nb_mpm_svd_fallback_stats", "nb_mpm_autotune_cached",
    "nb_mpm_profile_calls", "nb_mpm_set_fused_g2p",
)


class EmitterGPU(C.Structure):
    """nb_emitter: one emitter record (shape, center, size, velocities, rate, ...)."""
    _fields_ = [("shape", C.c_uint32), ("center", C.c_float * 3), ("size", C.c_float * 3), ("vel", C.c_float * 3),
                ("center_vel", C.c_float * 3), ("particles_per_step", C.c_int32), ("particle_type", C.c_uint32),
                ("emit_along_normal", C.c_uint32), ("velocity_randomness", C.c_float), ("step_count", C.c_int32),
                ("reserved", C.c_uint32 * 2)]


class Emitter:
    """src/core/Emitter.h:39-96 (the MPM-side fields)."""
    shape = 0

    def __init__(self):
        self.center = (0.0, 0.0, 0.0)
        self.vel = (0.0, 0.0, 0.0)
        self.center_vel = (0.0, 0.0, 0.0)
        self.particles_per_step = 1024
        self.particleType = 1
        self.emitAlongNormal = False
        self.velocityRandomness = 0.0
        self.step_count = -1

    def _size(self):
        return (0.0, 0.0, 0.0)

    def pack(self) -> EmitterGPU:
        g = EmitterGPU()
        g.shape = self.shape
        g.center[:] = [float(x) for x in self.center]
        g.size[:] = [float(x) for x in self._size()]
        g.vel[:] = [float(x) for x in self.vel]
        g.center_vel[:] = [float(x) for x in self.center_vel]
        g.particles_per_step = int(self.particles_per_step)
        g.particle_type = int(self.particleType)
        g.emit_along_normal = 1 if self.emitAlongNormal else 0
        g.velocity_randomness = float(self.velocityRandomness)
        g.step_count = int(self.step_count)
        return g


class AABBEmitter(Emitter):
    """src/core/Emitter.h:97-138: uniform in a box of edge lengths ``size``."""
    shape = 0

    def __init__(self):
        super().__init__()
        self.size = (1.0, 1.0, 1.0)

    def _size(self):
        return self.size

    @staticmethod
    def FromAABB(mn, mx, vel, particles_per_step):
        e = AABBEmitter()
        e.center = tuple((a + b) * 0.5 for a, b in zip(mn, mx))
        e.size = tuple(b - a for a, b in zip(mn, mx))
        e.vel = tuple(vel)
        e.particles_per_step = particles_per_step
        return e


class SphereEmitter(Emitter):
    """src/core/Emitter.h:141-174: uniform in a ball."""
    shape = 1

    def __init__(self):
        super().__init__()
        self.radius = 1.0

    def _size(self):
        return (self.radius, self.radius, self.radius)

    @staticmethod
    def FromSphere(center, radius, vel, particles_per_step):
        e = SphereEmitter()
        e.center, e.radius, e.vel, e.particles_per_step = tuple(center), radius, tuple(vel), particles_per_step
        return e


class EllipseEmitter(Emitter):
    """src/core/Emitter.h:176-224 (sample(): semiA along x, semiB along y, z at the center)."""
    shape = 2

    def __init__(self):
        super().__init__()
        self.semiA, self.semiB = 1.0, 1.0

    def _size(self):
        return (self.semiA, self.semiB, 0.0)

TUNE_ROUNDING = 1


class Tuning(C.Structure):
    """Launch shape (nb_mpm_tuning): blocks of the per-tile kernels, threads
    per G2P block, particles per P2G work item, and the autotuner's timing."""
    _fields_ = [("tile_blocks", C.c_uint32), ("g2p_block", C.c_uint32), ("p2g_item", C.c_uint32),
                ("ms_per_substep", C.c_float)]

    def __repr__(self):
        return (f"Tuning(tile_blocks={self.tile_blocks}, g2p_block={self.g2p_block}, p2g_item={self.p2g_item}, "
                f"ms_per_substep={self.ms_per_substep:.4f})")

MAX_BATCH_MATERIALS = 1024


class SvdFallback(C.Structure):
    """nb_mpm_svd_fallback: the FP64 SVD of near-degenerate F in the
    Drucker-Prager family (an extension; the reference is FP32 throughout)."""
    _fields_ = [("enabled", C.c_uint32), ("gap_rel", C.c_float), ("zero_rel", C.c_float),
                ("count_stats", C.c_uint32)]


# the counters of nb_mpm_svd_fallback_stats, in order
SVD_FB_STATS = ("dp_lanes", "fp64_return_map_lanes", "fp64_stress_lanes", "fp64_waves", "dp_waves")


class BatchScene(C.Structure):
    """One scene of a batch (nb_mpm_scene): its domain [0, domain_size],
    cell size and number of particle rows."""
    _fields_ = [("domain_size", C.c_float * 3), ("cell_size", C.c_float), ("particle_count", C.c_uint32),
                ("reserved", C.c_uint32 * 3)]

    def __init__(self, domainSize=(1.0, 1.0, 1.0), cellSize=1.0 / 64, particleCount=0):
        super().__init__()
        self.domain_size[:] = [float(x) for x in domainSize]
        self.cell_size = float(cellSize)
        self.particle_count = int(particleCount)


class NebulaError(RuntimeError):
    pass


def hip_runtimes_mapped():
    """Distinct HIP runtime files mapped into this process (from
    /proc/self/maps; empty where that file does not exist)."""
    try:
        with open("/proc/self/maps") as fh:
            return sorted({os.path.realpath(ln.split()[-1]) for ln in fh if "libamdhip64" in ln and "/" in ln})
    except OSError:
        return []


def check_single_hip_runtime():
    """Raise when two HIP runtimes are mapped into this process.

    torch's ROCm wheel links its own copy of the runtime under the name
    libamdhip64.so, and this library asks the loader for libamdhip64.so.7.
    The dynamic loader takes the two names for different libraries: when
    this library is loaded before torch, importing torch maps a second HIP
    (and HSA) runtime, and device pointers, streams and events of one are
    meaningless to the other.  Importing torch first makes this library bind
    to torch's runtime, which is what nebula_amd.autograd does."""
    rts = hip_runtimes_mapped()
    if len(rts) > 1:
        raise NebulaError("two HIP runtimes are mapped into this process (" + ", ".join(rts) + "); torch was "
                          "imported after the nebula_amd library loaded: import torch first")


def lib():
    """Load libnebula_mpm.so (raises if it was not built, or if it would run
    beside a second HIP runtime)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NebulaError(f"HIP library missing: {LIB_PATH} — build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, fp, up = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        L.nb_last_error.restype = C.c_char_p
        L.nb_mpm_create.argtypes = [C.POINTER(MPMConfig), vp, C.POINTER(vp)]
        L.nb_mpm_create_batch.argtypes = [C.POINTER(BatchScene), C.c_uint32, vp, C.POINTER(vp)]
        L.nb_mpm_scene_count.argtypes = [vp, up]
        L.nb_mpm_scene_info.argtypes = [vp, C.c_uint32, up, up, up]
        L.nb_mpm_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.nb_mpm_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.nb_mpm_autotune.argtypes = [vp, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(Tuning)]
        L.nb_mpm_graph_substeps.argtypes = [vp, C.c_float, C.c_uint32]
        L.nb_mpm_autotune_cached.argtypes = [vp, C.c_float, C.c_uint32, C.c_uint32, C.c_char_p, C.POINTER(Tuning),
                                             C.POINTER(C.c_int)]
        L.nb_mpm_set_svd_fallback.argtypes = [vp, C.POINTER(SvdFallback)]
        L.nb_mpm_get_svd_fallback.argtypes = [vp, C.POINTER(SvdFallback)]
        L.nb_mpm_svd_fallback_stats.argtypes = [vp, C.POINTER(C.c_uint64), C.c_int]
        L.nb_mpm_set_fused_g2p.argtypes = [vp, C.c_int]
        L.nb_mpm_add_emitter.argtypes = [vp, C.POINTER(EmitterGPU)]
        L.nb_mpm_set_collider_mesh.argtypes = [vp, fp, C.c_uint32, C.c_int]
        L.nb_mpm_get_collider_sdf.argtypes = [vp, fp]
        L.nb_mpm_clear_emitters.argtypes = [vp]
        L.nb_mpm_emit.argtypes = [vp, C.c_float]
        L.nb_mpm_get_emitter.argtypes = [vp, C.c_uint32, C.POINTER(EmitterGPU), up]
        L.nb_mpm_config_default.argtypes = [C.POINTER(MPMConfig)]
        L.nb_mpm_config_default.restype = None
        L.nb_mpm_knobs_default.argtypes = [C.POINTER(_Knobs)]
        L.nb_mpm_knobs_default.restype = None
        for nm in ("nb_mpm_destroy", "nb_mpm_sync", "nb_mpm_clear_analytic_colliders", "nb_mpm_clear_collider",
                   "nb_mpm_reset_particles", "nb_mpm_run_bin"):
            getattr(L, nm).argtypes = [vp]
        for nm in ("nb_mpm_step", "nb_mpm_substep", "nb_mpm_run_p2g", "nb_mpm_run_grid_update", "nb_mpm_run_g2p"):
            getattr(L, nm).argtypes = [vp, C.c_float]
        L.nb_mpm_substeps.argtypes = [vp, C.c_float, C.c_uint32]
        L.nb_mpm_set_stream.argtypes = [vp, vp]
        L.nb_mpm_set_knobs.argtypes = [vp, C.POINTER(_Knobs)]
        L.nb_mpm_get_knobs.argtypes = [vp, C.POINTER(_Knobs)]
        L.nb_mpm_get_config.argtypes = [vp, C.POINTER(MPMConfig)]
        L.nb_mpm_grid_res.argtypes = [vp, up]
        L.nb_mpm_total_cells.argtypes = [vp, up]
        L.nb_mpm_live_particle_count.argtypes = [vp, up]
        L.nb_mpm_set_materials.argtypes = [vp, vp, C.c_uint32]
        L.nb_mpm_set_particle_material_ids.argtypes = [vp, up, C.c_uint32]
        L.nb_mpm_set_colliders.argtypes = [vp, vp, C.c_uint32]
        L.nb_mpm_set_collider_sdf.argtypes = [vp, fp, C.c_uint64]
        L.nb_mpm_set_collider_sphere.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
        L.nb_mpm_set_forces.argtypes = [vp, vp, C.c_uint32]
        L.nb_mpm_append_particles.argtypes = [vp, fp, fp, C.c_uint32, up]
        L.nb_mpm_get_field.argtypes = [vp, C.c_int, fp, C.c_uint32, C.c_uint32]
        L.nb_mpm_set_field.argtypes = [vp, C.c_int, fp, C.c_uint32, C.c_uint32]
        L.nb_mpm_field_device_ptr.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.nb_mpm_set_live_particle_count.argtypes = [vp, C.c_uint32]
        L.nb_mpm_get_bins.argtypes = [vp, up, up, up, up]
        L.nb_mpm_get_grid.argtypes = [vp, fp, fp]
        L.nb_mpm_set_grid.argtypes = [vp, fp, fp, fp]
        L.nb_mpm_profile_substeps.argtypes = [vp, C.c_float, C.c_uint32, fp, C.c_uint32, up, C.c_char_p, C.c_uint32]
        L.nb_mpm_profile_calls.argtypes = [vp, up, C.c_uint32]
        L.nb_mpm_active_node_count.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.nb_mpm_calibration_copy.argtypes = [vp, C.c_uint64, C.c_uint32]
        L.nb_mpm_active_tile_count.argtypes = [vp, up]
        L.nb_mpm_adjoint_begin.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.nb_mpm_adjoint_begin_tapes.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32]
        L.nb_mpm_adjoint_tapes.argtypes = [vp, C.POINTER(C.c_uint32)]
        L.nb_mpm_forward_recorded.argtypes = [vp, C.c_float]
        L.nb_mpm_set_adjoint.argtypes = [vp, C.c_int, vp, C.c_uint32, C.c_uint32]
        L.nb_mpm_get_adjoint.argtypes = [vp, C.c_int, vp, C.c_uint32, C.c_uint32]
        L.nb_mpm_adjoint_device_ptr.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.nb_mpm_backward.argtypes = [vp, fp, C.c_uint32]
        L.nb_mpm_adjoint_end.argtypes = [vp]
        _lib = L
        check_single_hip_runtime()
    return _lib


def _check(rc):
    if rc != NB_OK:
        raise NebulaError(f"nebula_mpm error {rc}: {lib().nb_last_error().decode()}")


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32)) if a is not None else None


class MPMEngine:
    """The reference's MPMEngine over the HIP library.

    ``init(cfg)`` allocates and seeds the default particle block exactly as
    src/engine/MPMEngine.cpp:34-156 does; ``step(dt)`` enqueues
    ``numSubsteps`` substeps on the engine's HIP stream (call ``sync()`` to
    wait).  Public knobs are properties with the reference's names.
    """

    def __init__(self, cfg: Optional[MPMConfig] = None, stream: Optional[int] = None):
        self._h = None
        self._forces: List[Force] = []
        if cfg is not None:
            self.init(cfg, stream)

    # lifecycle -----------------------------------------------------------
    def init(self, cfg: MPMConfig, stream: Optional[int] = None):
        h = C.c_void_p()
        L = lib()
        check_single_hip_runtime()
        _check(L.nb_mpm_create(C.byref(cfg), C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h
        self._cfg = cfg
        self._n_cells = self.totalCells()

    @classmethod
    def batch(cls, scenes: Sequence[BatchScene], stream: Optional[int] = None) -> "MPMEngine":
        """An engine over a batch of independent scenes (nb_mpm_create_batch):
        scene s owns particle rows [first, first + particle_count) in table
        order, positions in its own coordinates."""
        e = cls()
        arr = (BatchScene * len(scenes))(*scenes)
        h = C.c_void_p()
        L = lib()
        check_single_hip_runtime()
        _check(L.nb_mpm_create_batch(arr, len(scenes), C.c_void_p(stream) if stream else None, C.byref(h)))
        e._h = h
        e._cfg = e.config()
        e._n_cells = e.totalCells()
        return e

    # emitters (src/engine/MPMEngine.cpp:271-352) ------------------------------
    def addEmitter(self, e: "Emitter"):
        g = e.pack()
        _check(lib().nb_mpm_add_emitter(self._h, C.byref(g)))
        self.__dict__.setdefault("_emitters", []).append(e)

    def clearEmitters(self):
        _check(lib().nb_mpm_clear_emitters(self._h))
        self._emitters = []

    def emitFromEmitters(self, dt: float):
        _check(lib().nb_mpm_emit(self._h, dt))
        self._sync_emitters()

    def _sync_emitters(self):
        for i, e in enumerate(getattr(self, "_emitters", [])):
            g = EmitterGPU()
            _check(lib().nb_mpm_get_emitter(self._h, i, C.byref(g), None))
            e.center = tuple(g.center)

    # launch shapes and graph replay ---------------------------------------
    def tuning(self) -> "Tuning":
        t = Tuning()
        _check(lib().nb_mpm_get_tuning(self._h, C.byref(t)))
        return t

    def setTuning(self, t: "Tuning"):
        _check(lib().nb_mpm_set_tuning(self._h, C.byref(t)))

    def autotune(self, dt_sub: float, trial_substeps: int = 3, rounding: bool = False) -> "Tuning":
        """Time candidate launch shapes on the current state (left as it
        was) and keep the fastest (nb_mpm_autotune)."""
        t = Tuning()
        _check(lib().nb_mpm_autotune(self._h, dt_sub, trial_substeps, TUNE_ROUNDING if rounding else 0,
                                     C.byref(t)))
        return t

    def setFusedG2P(self, on: bool):
        """Run the grid update inside G2P for all but a call's last substep
        (nb_mpm_set_fused_g2p; off by default)."""
        _check(lib().nb_mpm_set_fused_g2p(self._h, int(bool(on))))

    # the FP64 SVD fallback ---------------------------------------------------
    def svdFallback(self) -> "SvdFallback":
        f = SvdFallback()
        _check(lib().nb_mpm_get_svd_fallback(self._h, C.byref(f)))
        return f

    def setSvdFallback(self, enabled: bool = True, gap_rel: float = 1e-4, zero_rel: float = 1e-6,
                       count_stats: bool = False):
        f = SvdFallback(int(bool(enabled)), gap_rel, zero_rel, int(bool(count_stats)))
        _check(lib().nb_mpm_set_svd_fallback(self._h, C.byref(f)))

    def svdFallbackStats(self, reset: bool = False) -> dict:
        """The fallback's counters since the last reset (counted while
        count_stats is on), by name (SVD_FB_STATS)."""
        out = (C.c_uint64 * len(SVD_FB_STATS))()
        _check(lib().nb_mpm_svd_fallback_stats(self._h, out, int(bool(reset))))
        return dict(zip(SVD_FB_STATS, (int(x) for x in out)))

    def autotuneCached(self, dt_sub: float, cache_path: str, trial_substeps: int = 3, rounding: bool = False):
        """autotune behind the persisted cache at cache_path
        (nb_mpm_autotune_cached): (tuning, hit), hit True when the file
        already held this workload's shape and nothing was timed."""
        t = Tuning()
        hit = C.c_int(0)
        _check(lib().nb_mpm_autotune_cached(self._h, dt_sub, trial_substeps, TUNE_ROUNDING if rounding else 0,
                                            os.fsencode(cache_path), C.byref(t), C.byref(hit)))
        return t, bool(hit.value)

    def graphSubstep(self, dt_sub: float, n: int = 1):
        """n substeps as replays of the captured substep graph (nb_mpm_graph_substeps)."""
        _check(lib().nb_mpm_graph_substeps(self._h, dt_sub, n))

    def sceneCount(self) -> int:
        n = C.c_uint32()
        _check(lib().nb_mpm_scene_count(self._h, C.byref(n)))
        return n.value

    def sceneInfo(self, scene: int):
        """(grid resolution, first particle row, particle count) of one scene."""
        res = (C.c_uint32 * 3)()
        f, c = C.c_uint32(), C.c_uint32()
        _check(lib().nb_mpm_scene_info(self._h, scene, res, C.byref(f), C.byref(c)))
        return tuple(res), f.value, c.value

    def cleanup(self):
        if self._h is not None:
            lib().nb_mpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # stepping ------------------------------------------------------------
    def step(self, dt: float):
        for f in self._forces:
            f.advanceTime(dt)
        if self._forces:
            self._upload_forces()
        _check(lib().nb_mpm_step(self._h, dt))
        self._sync_emitters()

    def substep(self, dt_sub: float, n: int = 1):
        _check(lib().nb_mpm_substeps(self._h, dt_sub, n))

    def sync(self):
        _check(lib().nb_mpm_sync(self._h))

    def setStream(self, stream: Optional[int] = None):
        """Enqueue later work on the hipStream_t `stream` (an integer handle,
        e.g. torch.cuda.Stream().cuda_stream), or on the engine's own stream
        when None (nb_mpm_set_stream)."""
        _check(lib().nb_mpm_set_stream(self._h, C.c_void_p(stream) if stream else None))

    # knobs ---------------------------------------------------------------
    def _knobs(self) -> _Knobs:
        k = _Knobs()
        _check(lib().nb_mpm_get_knobs(self._h, C.byref(k)))
        return k

    def _set_knob(self, name, value):
        k = self._knobs()
        setattr(k, name, value)
        _check(lib().nb_mpm_set_knobs(self._h, C.byref(k)))

    numSubsteps = property(lambda s: s._knobs().num_substeps, lambda s, v: s._set_knob("num_substeps", int(v)))
    flip_ratio = property(lambda s: s._knobs().flip_ratio, lambda s, v: s._set_knob("flip_ratio", float(v)))
    restitution = property(lambda s: s._knobs().restitution, lambda s, v: s._set_knob("restitution", float(v)))
    wall_friction = property(lambda s: s._knobs().wall_friction, lambda s, v: s._set_knob("wall_friction", float(v)))
    plasticModel = property(lambda s: s._knobs().plastic_model, lambda s, v: s._set_knob("plastic_model", int(v)))
    q_max = property(lambda s: s._knobs().q_max, lambda s, v: s._set_knob("q_max", float(v)))
    M_friction = property(lambda s: s._knobs().M_friction, lambda s, v: s._set_knob("M_friction", float(v)))
    q_cohesion = property(lambda s: s._knobs().q_cohesion, lambda s, v: s._set_knob("q_cohesion", float(v)))
    # extension: singular vectors used as columns (see include/nebula_mpm.h)
    corrected_svd = property(lambda s: bool(s._knobs().corrected_svd),
                             lambda s, v: s._set_knob("corrected_svd", 1 if v else 0))

    # queries -------------------------------------------------------------
    def config(self) -> MPMConfig:
        c = MPMConfig()
        _check(lib().nb_mpm_get_config(self._h, C.byref(c)))
        return c

    def liveParticleCount(self) -> int:
        n = C.c_uint32()
        _check(lib().nb_mpm_live_particle_count(self._h, C.byref(n)))
        return n.value

    def gridRes(self):
        r = (C.c_uint32 * 3)()
        _check(lib().nb_mpm_grid_res(self._h, r))
        return tuple(r)

    def totalCells(self) -> int:
        n = C.c_uint32()
        _check(lib().nb_mpm_total_cells(self._h, C.byref(n)))
        return n.value

    # scene data ----------------------------------------------------------
    def setMaterials(self, mats: Sequence[MaterialParams]):
        arr = (MaterialParams * max(len(mats), 1))(*mats)
        _check(lib().nb_mpm_set_materials(self._h, C.cast(arr, C.c_void_p), len(mats)))
        if mats:
            self._materials = [MaterialParams.from_buffer_copy(bytes(m)) for m in mats]

    def materials(self) -> List[MaterialParams]:
        """Copies of the material table last given to setMaterials (empty if none)."""
        return [MaterialParams.from_buffer_copy(bytes(m)) for m in getattr(self, "_materials", [])]

    def setParticleMaterialIds(self, ids):
        a = np.ascontiguousarray(ids, np.uint32)
        _check(lib().nb_mpm_set_particle_material_ids(self._h, _u(a), a.size))

    def setColliders(self, cols: ColliderSet):
        data = cols.data()
        arr = (ColliderPrimitive * max(len(data), 1))(*data)
        _check(lib().nb_mpm_set_colliders(self._h, C.cast(arr, C.c_void_p), len(data)))

    def clearAnalyticColliders(self):
        _check(lib().nb_mpm_clear_analytic_colliders(self._h))

    def setColliderSDF(self, mortonSDF):
        a = np.ascontiguousarray(mortonSDF, np.float32)
        _check(lib().nb_mpm_set_collider_sdf(self._h, _f(a), a.size))

    def setColliderSphere(self, radius, cx, cy, cz):
        _check(lib().nb_mpm_set_collider_sphere(self._h, radius, cx, cy, cz))

    def setColliderMesh(self, tris, exact=False):
        """buildMeshSDF on the GPU, installed as the collider SDF
        (nb_mpm_set_collider_mesh); tris (n, 12): normal, v0, v1, v2.
        exact=False keeps closestPtOnTriangle's missing edge-v1-v2 region
        (MeshSDF.h:48-74); exact=True measures the true distance."""
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 12)
        _check(lib().nb_mpm_set_collider_mesh(self._h, _f(t), t.shape[0], int(bool(exact))))

    def getColliderSDF(self):
        out = np.empty(self._n_cells, np.float32)
        _check(lib().nb_mpm_get_collider_sdf(self._h, _f(out)))
        return out

    def clearCollider(self):
        _check(lib().nb_mpm_clear_collider(self._h))

    def _upload_forces(self):
        packed = [f.pack() for f in self._forces]
        arr = (ForceGPU * max(len(packed), 1))(*packed)
        _check(lib().nb_mpm_set_forces(self._h, C.cast(arr, C.c_void_p), len(packed)))

    def addForce(self, f: Force):
        self._forces.append(f)
        self._upload_forces()

    def removeForce(self, f: Force):
        self._forces = [x for x in self._forces if x is not f]
        self._upload_forces()

    def setForces(self, forces: Sequence[Force]):
        self._forces = list(forces)
        self._upload_forces()

    def clearForces(self):
        self._forces = []
        self._upload_forces()

    def forces(self):
        return list(self._forces)

    def appendParticles(self, pos, vel) -> int:
        p = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        v = np.ascontiguousarray(vel, np.float32).reshape(-1, 4)
        if p.shape != v.shape:
            raise ValueError("pos and vel must have the same length")
        n = C.c_uint32()
        _check(lib().nb_mpm_append_particles(self._h, _f(p), _f(v), p.shape[0], C.byref(n)))
        return n.value

    def resetParticles(self):
        _check(lib().nb_mpm_reset_particles(self._h))

    # state access ----------------------------------------------------------
    def getField(self, field: int, first: int = 0, count: Optional[int] = None):
        if count is None:
            count = self.liveParticleCount() - first
        out = np.empty((count, 4), np.float32)
        if count:
            _check(lib().nb_mpm_get_field(self._h, field, _f(out), first, count))
        return out

    def setField(self, field: int, data, first: int = 0):
        a = np.ascontiguousarray(data, np.float32).reshape(-1, 4)
        _check(lib().nb_mpm_set_field(self._h, field, _f(a), first, a.shape[0]))

    def getState(self):
        """(8, n, 4) float32: P, v, F0, F1, F2, B0, B1, B2."""
        return np.stack([self.getField(f) for f in range(8)])

    def setStateFrom(self, ptr: int, n: int):
        """Set the live count to n and copy the state from a host or device
        address holding (8, n, 4) float32 (the eight arrays back to back)."""
        _check(lib().nb_mpm_set_live_particle_count(self._h, n))
        for i in range(8):
            _check(lib().nb_mpm_set_field(self._h, i, C.cast(C.c_void_p(ptr + i * n * 16), C.POINTER(C.c_float)), 0, n))

    def getStateTo(self, ptr: int):
        """Copy the (8, n, 4) state to a host or device address."""
        n = self.liveParticleCount()
        for i in range(8):
            _check(lib().nb_mpm_get_field(self._h, i, C.cast(C.c_void_p(ptr + i * n * 16), C.POINTER(C.c_float)), 0, n))

    def setState(self, fields):
        f = np.ascontiguousarray(fields, np.float32)
        n = f.shape[1]
        _check(lib().nb_mpm_set_live_particle_count(self._h, n))
        for i in range(8):
            self.setField(i, f[i])

    def getPositionBuffer(self) -> int:
        p = C.c_void_p()
        _check(lib().nb_mpm_field_device_ptr(self._h, FIELD_P, C.byref(p)))
        return p.value

    def getVelocityBuffer(self) -> int:
        p = C.c_void_p()
        _check(lib().nb_mpm_field_device_ptr(self._h, FIELD_V, C.byref(p)))
        return p.value

    # pass-level access (tests/helpers/MPMHarness.h:56-59) ----------------------
    def runBin(self):
        _check(lib().nb_mpm_run_bin(self._h))

    def getBins(self):
        n = self.liveParticleCount()
        NC = self._n_cells
        keys = np.empty(n, np.uint32)
        srt = np.empty(n, np.uint32)
        cnt = np.empty(NC, np.uint32)
        end = np.empty(NC, np.uint32)
        _check(lib().nb_mpm_get_bins(self._h, _u(keys), _u(srt), _u(cnt), _u(end)))
        return keys, srt, cnt, end

    def runP2G(self, dt):
        _check(lib().nb_mpm_run_p2g(self._h, dt))

    def runGridUpdate(self, dt):
        _check(lib().nb_mpm_run_grid_update(self._h, dt))

    def runG2P(self, dt):
        _check(lib().nb_mpm_run_g2p(self._h, dt))

    def getGrid(self):
        NC = self._n_cells
        mom = np.empty((2 * NC, 4), np.float32)
        mass = np.empty(NC, np.float32)
        _check(lib().nb_mpm_get_grid(self._h, _f(mom), _f(mass)))
        return mom, mass

    def setGrid(self, mom4, mass, old4=None):
        """mom4: NC float4 (or the full 2*NC export, whose second half is then
        taken as the saved velocity unless old4 is given), mass: NC floats."""
        NC = self._n_cells
        full = np.asarray(mom4, np.float32).reshape(-1, 4)
        m = np.ascontiguousarray(full[:NC])
        if old4 is None and full.shape[0] >= 2 * NC:
            old4 = full[NC:2 * NC]
        ms = np.ascontiguousarray(mass, np.float32)
        o = np.ascontiguousarray(np.asarray(old4, np.float32).reshape(-1, 4)[:NC]) if old4 is not None else None
        _check(lib().nb_mpm_set_grid(self._h, _f(m), _f(ms), _f(o) if o is not None else None))

    # measurement -----------------------------------------------------------
    def profileSubsteps(self, dt_sub: float, n: int):
        """Run n substeps with HIP events around every kernel; returns
        {kernel name: total ms}."""
        ms = np.zeros(32, np.float32)
        nk = C.c_uint32()
        names = C.create_string_buffer(1024)
        _check(lib().nb_mpm_profile_substeps(self._h, dt_sub, n, _f(ms), 32, C.byref(nk), names, 1024))
        keys = names.value.decode().split(",")[: nk.value]
        self._prof_names = keys
        return {k: float(ms[i]) for i, k in enumerate(keys)}

    def profileCalls(self):
        """{kernel name: launches} of the last profileSubsteps (its substeps
        but the last run the grid update inside G2P, as g2p_grid)."""
        calls = np.zeros(32, np.uint32)
        _check(lib().nb_mpm_profile_calls(self._h, _u(calls), 32))
        return {k: int(calls[i]) for i, k in enumerate(getattr(self, "_prof_names", []))}

    def calibrationCopy(self, nbytes: int, repeats: int = 1):
        """Plain copy kernel of nbytes (HBM byte-counter calibration)."""
        _check(lib().nb_mpm_calibration_copy(self._h, nbytes, repeats))

    def activeNodeCount(self) -> int:
        g = C.c_uint64()
        _check(lib().nb_mpm_active_node_count(self._h, C.byref(g)))
        return g.value

    def activeTileCount(self) -> int:
        t = C.c_uint32()
        _check(lib().nb_mpm_active_tile_count(self._h, C.byref(t)))
        return t.value

    # ---- reverse mode (no counterpart in the reference) ----
    def adjointBegin(self, n_substeps: int, segment: int, tapes: int = 1):
        """Reserve checkpoints for a recorded run of n_substeps, one every
        `segment` substeps (needs corrected_svd = 1; PIC/APIC; elastic, Von
        Mises, Drucker-Prager), and `tapes` segments' inner states at once
        (0: as many as fit in free device memory); the backward recomputes
        only the segments the tapes do not hold."""
        _check(lib().nb_mpm_adjoint_begin_tapes(self._h, n_substeps, segment, tapes))

    def adjointTapes(self) -> int:
        t = C.c_uint32()
        _check(lib().nb_mpm_adjoint_tapes(self._h, C.byref(t)))
        return t.value

    def forwardRecorded(self, dt_sub: float):
        _check(lib().nb_mpm_forward_recorded(self._h, dt_sub))

    def setAdjoint(self, grads):
        """dL/d(final state): (8, n, 4) float32 in the state's layout, a host
        array, or an int device pointer base per field via setAdjointField."""
        g = np.ascontiguousarray(grads, np.float32)
        for i in range(8):
            _check(lib().nb_mpm_set_adjoint(self._h, i, g[i].ctypes.data, 0, g.shape[1]))

    def setAdjointField(self, field: int, ptr: int, count: int, first: int = 0):
        """Copy `count` float4 rows from a host or device address."""
        _check(lib().nb_mpm_set_adjoint(self._h, field, ptr, first, count))

    def getAdjointField(self, field: int, ptr: int, count: int, first: int = 0):
        _check(lib().nb_mpm_get_adjoint(self._h, field, ptr, first, count))

    def getAdjoint(self):
        n = self.liveParticleCount()
        out = np.empty((8, n, 4), np.float32)
        for i in range(8):
            _check(lib().nb_mpm_get_adjoint(self._h, i, out[i].ctypes.data, 0, n))
        return out

    def backward(self):
        """Adjoint of the recorded run: returns dL/d(mu, lambda, M_friction,
        q_cohesion) per material, (material_count, 4)."""
        n = max(len(getattr(self, "_materials", [])), 1)
        buf = np.zeros(4 * n, np.float32)
        _check(lib().nb_mpm_backward(self._h, _f(buf), buf.size))
        return buf.reshape(n, 4)

    def adjointEnd(self):
        _check(lib().nb_mpm_adjoint_end(self._h))
