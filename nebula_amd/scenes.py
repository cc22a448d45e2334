"""Seeded synthetic scenes for the MPM substep workloads.

Each builder returns a ``Scene``: the engine configuration, material table,
colliders, forces, optional dense Morton SDF and the initial particle state in
the reference's eight-float4 layout (P, v, F0, F1, F2, B0, B1, B2).  Particles
are seeded cell by cell in Morton order of the cells (8 per cell, uniform
jitter inside the cell), so storage order follows space.

  elastic_block_64   unit box, 64^3 grid, a 0.25 cube of Hencky-elastic jelly
                     (E=1e4, nu=0.2, rho0=1) centred at (0.5, 0.6, 0.5):
                     32,768 particles, v0 = (0,-1,0) + N(0, 0.01^2), APIC,
                     gravity -9.8, dt = 1e-4
  multi_material_128 unit box, 128^3 grid, 262,144 particles: an elastic jelly
                     cube (E=5e3, nu=0.3) and a Drucker-Prager sand column
                     (E=3.5e4, nu=0.3, M=tan 30deg, cohesion 0), a sphere
                     collider r=0.1 moving at 0.5 m/s, APIC, dt = 5e-5
  snow_terrain_256   unit box, 256^3 grid, 4,194,304 Drucker-Prager snow
                     particles (E=1.4e5, nu=0.2, rho0=400, M=tan 30deg,
                     cohesion 500) in an 8-cell layer above a seeded six-mode
                     sine terrain given as a dense SDF, APIC, dt = 5e-5

The reference's own scene builders these follow: examples/mpm_elastic.cpp,
examples/mpm_multimaterial.cpp, examples/mpm_avalanche.cpp:176-205 (terrain
SDF as cell-centre height difference, snow placed 2 cells above the terrain).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .mpm import (ColliderSet, GravityForce, MaterialParams, MaterialModel, MPMConfig, calcLambda, calcMu,
                  gridRes as domain_grid_res, hashCells, mortonEncodeArray)


@dataclass
class Scene:
    name: str
    cfg: MPMConfig
    materials: List[MaterialParams]
    fields: np.ndarray  # (8, n, 4) float32
    dt: float
    flip_ratio: float
    substeps: int
    colliders: ColliderSet = field(default_factory=ColliderSet)
    forces: list = field(default_factory=list)
    sdf: Optional[np.ndarray] = None
    seed: int = 0
    collider_velocity: Optional[tuple] = None  # moving analytic collider (index 0)
    corrected_svd: bool = False  # singular vectors as columns (an extension; see include/nebula_mpm.h)

    @property
    def n(self):
        return self.fields.shape[1]

    def grid_res(self):
        return domain_grid_res(tuple(self.cfg.domainSize), self.cfg.cellSize)

    def total_cells(self):
        return hashCells(self.grid_res())

    def colliders_at(self, t: float) -> ColliderSet:
        """Collider set at time t (the moving collider advanced on the host,
        as examples/mpm_snow_impact.cpp:225-229 does per frame)."""
        if self.collider_velocity is None or self.colliders.empty():
            return self.colliders
        cs = ColliderSet()
        for i, c in enumerate(self.colliders.data()):
            c2 = type(c).from_buffer_copy(bytes(c))
            if i == 0:
                vx, vy, vz = self.collider_velocity
                c2.px = np.float32(c.px + vx * t)
                c2.py = np.float32(c.py + vy * t)
                c2.pz = np.float32(c.pz + vz * t)
            cs._p.append(c2)
        return cs


def _material(E, nu, rho0, model, M=0.0, qc=0.0, qmax=0.0):
    m = MaterialParams()
    m.mu, m.lambda_, m.rho0, m.model = calcMu(E, nu), calcLambda(E, nu), rho0, model
    m.M_friction, m.q_cohesion, m.q_max = M, qc, qmax
    return m


def _seed_cells(cells_xyz: np.ndarray, h: float, ppc: int, rng: np.random.Generator):
    """Positions for ``ppc`` uniformly jittered particles in each cell, cells
    visited in Morton order.  cells_xyz: (m, 3) int."""
    keys = mortonEncodeArray(cells_xyz[:, 0], cells_xyz[:, 1], cells_xyz[:, 2])
    order = np.argsort(keys, kind="stable")
    cells = cells_xyz[order].astype(np.float64)
    jit = rng.random((cells.shape[0], ppc, 3))
    pos = (cells[:, None, :] + jit) * h
    return pos.reshape(-1, 3).astype(np.float32)


def _fields(pos: np.ndarray, vel: np.ndarray, Vp: float, mat_ids: Optional[np.ndarray] = None):
    n = pos.shape[0]
    f = np.zeros((8, n, 4), np.float32)
    f[0, :, :3] = pos
    f[0, :, 3] = np.float32(Vp)
    f[1, :, :3] = vel
    ids = np.zeros(n, np.uint32) if mat_ids is None else mat_ids.astype(np.uint32)
    f[1, :, 3] = ids.view(np.float32)
    f[2, :, 0] = 1.0
    f[3, :, 1] = 1.0
    f[4, :, 2] = 1.0
    return f


def _box_cells(lo, hi):
    xs, ys, zs = np.meshgrid(np.arange(lo[0], hi[0]), np.arange(lo[1], hi[1]), np.arange(lo[2], hi[2]),
                             indexing="ij")
    return np.stack([xs.ravel(), ys.ravel(), zs.ravel()], axis=1)


def elastic_block_64(seed: int = 1, res: int = 64, ppc: int = 8) -> Scene:
    h = 1.0 / res
    side = res // 4  # 0.25 m cube
    c0 = np.array([res // 2 - side // 2, int(round(0.6 * res)) - side // 2, res // 2 - side // 2])
    cells = _box_cells(c0, c0 + side)
    rng = np.random.default_rng(seed)
    pos = _seed_cells(cells, h, ppc, rng)
    vel = np.zeros_like(pos)
    vel[:, 1] = -1.0
    vel += rng.normal(0.0, 0.01, size=vel.shape).astype(np.float32)
    cfg = MPMConfig(nx=0, ny=0, nz=0, domainSize=(1.0, 1.0, 1.0), cellSize=h, maxParticles=pos.shape[0],
                    E=1e4, nu=0.2, rho0=1.0)
    mats = [_material(1e4, 0.2, 1.0, MaterialModel.ELASTIC, qmax=1e5, M=0.577)]
    return Scene("elastic_block_64", cfg, mats, _fields(pos, vel, h ** 3 / ppc), dt=1e-4, flip_ratio=-1.0,
                 substeps=1000, forces=[GravityForce((0.0, 1.0, 0.0), -9.8)], seed=seed)


def multi_material_128(seed: int = 2, res: int = 128) -> Scene:
    h = 1.0 / res
    ppc = 8
    s = res / 128.0
    jelly_side = max(2, int(round(24 * s)))
    jc = np.array([int(0.3 * res) - jelly_side // 2, int(0.45 * res), res // 2 - jelly_side // 2])
    jelly = _box_cells(jc, jc + jelly_side)
    col_w, col_h = max(2, int(round(16 * s))), max(2, int(round(74 * s)))
    sc = np.array([int(0.7 * res) - col_w // 2, 3, res // 2 - col_w // 2])
    sand = _box_cells(sc, sc + np.array([col_w, col_h, col_w]))
    rng = np.random.default_rng(seed)
    pj = _seed_cells(jelly, h, ppc, rng)
    ps = _seed_cells(sand, h, ppc, rng)
    pos = np.concatenate([pj, ps])
    ids = np.concatenate([np.zeros(len(pj), np.uint32), np.ones(len(ps), np.uint32)])
    vel = np.zeros_like(pos)
    cfg = MPMConfig(nx=0, ny=0, nz=0, domainSize=(1.0, 1.0, 1.0), cellSize=h, maxParticles=pos.shape[0],
                    E=5e3, nu=0.3, rho0=1000.0)
    M = math.tan(math.radians(30.0))
    mats = [_material(5e3, 0.3, 1000.0, MaterialModel.ELASTIC),
            _material(3.5e4, 0.3, 1600.0, MaterialModel.DRUCKER_PRAGER, M=M, qc=0.0)]
    cols = ColliderSet()
    cols.addSphere((0.15, 0.15, 0.5), 0.1, restitution=0.0, friction=0.3, vel=(0.5, 0.0, 0.0))
    return Scene("multi_material_128", cfg, mats, _fields(pos, vel, h ** 3 / ppc, ids), dt=5e-5, flip_ratio=-1.0,
                 substeps=2000, colliders=cols, forces=[GravityForce((0.0, 1.0, 0.0), -9.8)], seed=seed,
                 collider_velocity=(0.5, 0.0, 0.0))


def terrain_height(x, z, params):
    """h(x,z) = 0.15 + sum_k a_k sin(f_k x + phi_k) cos(g_k z + psi_k)."""
    a, f, g, phi, psi = params
    hgt = np.full(np.broadcast(x, z).shape, 0.15, np.float64)
    for k in range(len(a)):
        hgt = hgt + a[k] * np.sin(f[k] * x + phi[k]) * np.cos(g[k] * z + psi[k])
    return hgt


def terrain_params(seed: int):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.01, 0.04, 6)
    f = rng.uniform(2.0, 12.0, 6)
    g = rng.uniform(2.0, 12.0, 6)
    phi = rng.uniform(0.0, 2 * math.pi, 6)
    psi = rng.uniform(0.0, 2 * math.pi, 6)
    return a, f, g, phi, psi


def terrain_sdf(res: int, params) -> np.ndarray:
    """Dense Morton SDF of the terrain: cell-centre y minus the terrain height
    (examples/mpm_avalanche.cpp:176-191, without its safety margin)."""
    h = 1.0 / res
    cube = 1
    while cube < res:
        cube <<= 1
    sdf = np.full(cube ** 3, 1e9, np.float32)
    c = (np.arange(res) + 0.5) * h
    X, Z = np.meshgrid(c, c, indexing="ij")  # (x, z)
    hgt = terrain_height(X, Z, params)  # (res, res)
    ix, iy, iz = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
    keys = mortonEncodeArray(ix.ravel(), iy.ravel(), iz.ravel())
    yc = (iy.ravel() + 0.5) * h
    val = (yc - hgt[ix.ravel(), iz.ravel()]).astype(np.float32)
    sdf[keys] = val
    return sdf


def snow_terrain_256(seed: int = 3, res: int = 256, layer: int = 8, ppc: int = 8) -> Scene:
    h = 1.0 / res
    params = terrain_params(seed)
    c = (np.arange(res) + 0.5) * h
    X, Z = np.meshgrid(c, c, indexing="ij")
    hgt = terrain_height(X, Z, params)
    start = np.clip(np.floor(hgt / h).astype(np.int64) + 2, 2, res - layer - 2)  # 2 cells above the terrain cell
    ix, iz = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    cells = []
    for k in range(layer):
        cells.append(np.stack([ix.ravel(), start.ravel() + k, iz.ravel()], axis=1))
    cells = np.concatenate(cells)
    rng = np.random.default_rng(seed)
    pos = _seed_cells(cells, h, ppc, rng)
    vel = np.zeros_like(pos)
    cfg = MPMConfig(nx=0, ny=0, nz=0, domainSize=(1.0, 1.0, 1.0), cellSize=h, maxParticles=pos.shape[0],
                    E=1.4e5, nu=0.2, rho0=400.0)
    M = math.tan(math.radians(30.0))
    mats = [_material(1.4e5, 0.2, 400.0, MaterialModel.DRUCKER_PRAGER, M=M, qc=500.0)]
    return Scene("snow_terrain_256", cfg, mats, _fields(pos, vel, h ** 3 / ppc), dt=5e-5, flip_ratio=-1.0,
                 substeps=500, forces=[GravityForce((0.0, 1.0, 0.0), -9.8)], sdf=terrain_sdf(res, params),
                 seed=seed)


def make_engine(scene: Scene, stream=None):
    """An MPMEngine holding the scene (materials, colliders, forces, SDF,
    particles) with the scene's transfer mode."""
    from .mpm import MPMEngine

    eng = MPMEngine(scene.cfg, stream)
    eng.setMaterials(scene.materials)
    eng.setForces(scene.forces)
    eng.setColliders(scene.colliders)
    if scene.sdf is not None:
        eng.setColliderSDF(scene.sdf)
    eng.setState(scene.fields)
    eng.flip_ratio = scene.flip_ratio
    eng.corrected_svd = scene.corrected_svd
    eng.numSubsteps = 1
    return eng


def make_batch_engine(scenes: List[Scene], stream=None):
    """One engine over a batch of the scenes (nb_mpm_create_batch): each
    scene's particles, in table order, with its materials appended to the
    shared table (material ids offset accordingly).  The scenes must agree on
    dt, transfer mode, forces and SVD convention; each keeps its own domain
    and cell size.  No SDF or colliders."""
    from .mpm import BatchScene, MPMEngine

    s0 = scenes[0]
    for sc in scenes:
        if sc.sdf is not None or sc.colliders.count():
            raise ValueError(f"scene {sc.name}: a batch takes no SDF or colliders")
        if sc.flip_ratio != s0.flip_ratio or sc.corrected_svd != s0.corrected_svd:
            raise ValueError("the scenes of a batch share the transfer mode and SVD convention")
    specs = [BatchScene(tuple(sc.cfg.domainSize), sc.cfg.cellSize, sc.n) for sc in scenes]
    eng = MPMEngine.batch(specs, stream)
    mats, fields = [], []
    for sc in scenes:
        f = sc.fields.copy()
        ids = f[1, :, 3].view(np.uint32) + np.uint32(len(mats))
        f[1, :, 3] = ids.view(np.float32)
        fields.append(f)
        mats.extend(sc.materials)
    eng.setMaterials(mats)
    eng.setForces(s0.forces)
    eng.setState(np.concatenate(fields, axis=1))
    eng.flip_ratio = s0.flip_ratio
    eng.corrected_svd = s0.corrected_svd
    eng.numSubsteps = 1
    return eng


def falling_box(name: str, grid: int, n: int, E: float, nu: float, model: int, rng: np.random.Generator,
                h: float = 1.0 / 64, dt: float = 1e-4, flip_ratio: float = -1.0, rho0: float = 1000.0,
                corrected_svd: bool = False) -> Scene:
    """A box of n particles (8 per cell, jittered) at a seeded place and
    velocity in the upper part of a grid^3-cell domain of cell size h,
    falling under gravity; one material of the given model (Von Mises
    q_max = 0.02 E, Drucker-Prager friction 30 deg, no cohesion)."""
    cells_n = -(-n // 8)
    e = max(1, round(cells_n ** (1.0 / 3.0)))
    dims = [e, e, -(-cells_n // (e * e))]
    lo = [int(rng.integers(3, max(4, grid - d - 3))) for d in dims]
    lo[1] = max(lo[1], grid // 3) if grid - dims[1] - 3 > grid // 3 else lo[1]
    g = [np.arange(lo[a], lo[a] + dims[a]) for a in range(3)]
    cx, cy, cz = np.meshgrid(*g, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)[:cells_n]
    pos = _seed_cells(cells, h, 8, rng)[:n]
    v0 = rng.uniform(-1.0, 1.0, 3).astype(np.float32)
    vel = np.broadcast_to(v0, (n, 3)).astype(np.float32)
    f = _fields(pos, vel, h ** 3 / 8)
    cfg = MPMConfig(nx=0, ny=0, nz=0, domainSize=(grid * h,) * 3, cellSize=h, maxParticles=n, E=E, nu=nu,
                    rho0=rho0)
    M = math.tan(math.radians(30)) if model == MaterialModel.DRUCKER_PRAGER else 0.0
    mats = [_material(E, nu, rho0, model, M=M, qc=0.0, qmax=0.02 * E)]
    return Scene(name, cfg, mats, f, dt=dt, flip_ratio=flip_ratio, substeps=200,
                 forces=[GravityForce((0.0, 1.0, 0.0), -9.8)], corrected_svd=corrected_svd)


def batch_mix_256(seed: int = 4, n_scenes: int = 256) -> List[Scene]:
    """The 256-scene mix: particle counts lognormal(ln 16000, 0.8) clipped to
    [2000, 64000], grid edges 32..96 cells (cell size 1/64), materials drawn
    uniformly from elastic / Von Mises / Drucker-Prager with E in [1e3, 1e5]
    and nu in [0.1, 0.4], each scene a falling box; dt = 1e-4, APIC."""
    rng = np.random.default_rng(seed)
    models = (MaterialModel.ELASTIC, MaterialModel.VON_MISES, MaterialModel.DRUCKER_PRAGER)
    out = []
    for s in range(n_scenes):
        n = int(np.clip(rng.lognormal(math.log(16000.0), 0.8), 2000, 64000))
        cells = -(-n // 8)
        edge = int(math.ceil(cells ** (1.0 / 3.0)))
        grid = int(rng.integers(max(32, edge + 10), 97))
        model = models[int(rng.integers(0, 3))]
        E = float(rng.uniform(1e3, 1e5))
        nu = float(rng.uniform(0.1, 0.4))
        out.append(falling_box(f"mix{s}", grid, n, E, nu, model, rng))
    return out


def layered_pair(name: str, E: float, nu: float, phi_deg: float, rng: np.random.Generator, h: float = 1.0 / 64,
                 dt: float = 1e-4) -> Scene:
    """One scene of the differentiable batch: on a 64^3 grid, an elastic
    layer resting on a Drucker-Prager layer (each 16 x 8 x 16 cells at 8
    jittered particles per cell: 32,768 particles), both falling at
    (0, -1, 0) plus seeded noise; E, nu shared by the layers, friction angle
    phi for the sand; proper polar SVD."""
    layers = []
    for y0 in (4, 12):
        g = [np.arange(24, 40), np.arange(y0, y0 + 8), np.arange(24, 40)]
        cx, cy, cz = np.meshgrid(*g, indexing="ij")
        layers.append(np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1))
    pos = np.concatenate([_seed_cells(c, h, 8, rng) for c in layers])
    n = pos.shape[0]
    vel = (np.array([0.0, -1.0, 0.0]) + 0.05 * rng.standard_normal((n, 3))).astype(np.float32)
    mid = np.concatenate([np.ones(n // 2, np.uint32), np.zeros(n - n // 2, np.uint32)])  # sand below, elastic above
    f = _fields(pos, vel, h ** 3 / 8, mid)
    cfg = MPMConfig(nx=0, ny=0, nz=0, domainSize=(64 * h,) * 3, cellSize=h, maxParticles=n, E=E, nu=nu, rho0=1000.0)
    mats = [_material(E, nu, 1000.0, MaterialModel.ELASTIC),
            _material(E, nu, 1000.0, MaterialModel.DRUCKER_PRAGER, M=math.tan(math.radians(phi_deg)), qc=0.0)]
    return Scene(name, cfg, mats, f, dt=dt, flip_ratio=-1.0, substeps=400,
                 forces=[GravityForce((0.0, 1.0, 0.0), -9.8)], corrected_svd=True)


def batch_layers_64(seed: int = 5, n_scenes: int = 64) -> List[Scene]:
    """The differentiable batch: 64 layered scenes (layered_pair) with hidden
    E in [1e4, 5e4], nu in [0.2, 0.35] and friction angle in [25, 40] deg
    drawn per scene; 2,097,152 particles in all."""
    rng = np.random.default_rng(seed)
    return [layered_pair(f"layers{s}", float(rng.uniform(1e4, 5e4)), float(rng.uniform(0.2, 0.35)),
                         float(rng.uniform(25.0, 40.0)), rng) for s in range(n_scenes)]
