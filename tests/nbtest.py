"""Shared helpers of the test suite: scene <-> oracle/engine plumbing and the
harness-style set-ups of the reference's per-pass tests
(tests/helpers/MPMHarness.cpp:41-90, 192-251)."""
from __future__ import annotations

import numpy as np

import oracle
from nebula_amd import mpm
from nebula_amd.mpm import (ColliderSet, GravityForce, MaterialModel, MaterialParams, MPMConfig, forces_bytes,
                            materials_bytes)
from nebula_amd.scenes import Scene


def oracle_sim(scene: Scene, precision=32, t=0.0):
    gr = scene.grid_res()
    sc = oracle.make_scene(scene.dt, scene.cfg.cellSize, gr, (0.0, 0.0, 0.0), tuple(scene.cfg.domainSize),
                           scene.flip_ratio)
    sim = oracle.Sim(sc, materials_bytes(scene.materials), scene.colliders_at(t).to_bytes(),
                     forces_bytes(scene.forces), scene.sdf, precision)
    sim.set_corrected_svd(scene.corrected_svd)
    sim.set_state(scene.fields)
    return sim


def engine(scene: Scene):
    from nebula_amd.scenes import make_engine

    return make_engine(scene)


def harness_scene(grid_res: int, world: float, particles, dt: float, rho0: float, mu: float, lam: float,
                  gravity: float, flip_ratio: float = 0.0) -> Scene:
    """The MPMHarness set-up: cube domain of grid_res cells of world/grid_res,
    one elastic material (mu, lam, rho0, q_max 1e5), one gravity force along +y
    with the given strength, PIC.  particles: list of dicts with pos, Vp, vel,
    F (3x3 row-major math matrix) and tau (3x3 symmetric)."""
    h = world / grid_res
    n = len(particles)
    f = np.zeros((8, max(n, 0), 4), np.float32)
    for i, p in enumerate(particles):
        F = np.asarray(p.get("F", np.eye(3)), np.float32)
        tau = np.asarray(p.get("tau", np.zeros((3, 3))), np.float32)
        f[0, i, :3] = p["pos"]
        f[0, i, 3] = p["Vp"]
        f[1, i, :3] = p.get("vel", (0, 0, 0))
        for c in range(3):
            f[2 + c, i, :3] = F[:, c]  # column c
        f[2, i, 3], f[3, i, 3], f[4, i, 3] = tau[0, 0], tau[1, 1], tau[2, 2]
        f[5, i, 3], f[6, i, 3], f[7, i, 3] = tau[0, 1], tau[0, 2], tau[1, 2]
    cfg = MPMConfig(nx=0, ny=0, nz=0, domainSize=(world, world, world), cellSize=h, maxParticles=max(n, 1),
                    E=1e4, nu=0.3, rho0=rho0)
    m = MaterialParams()
    m.mu, m.lambda_, m.rho0, m.model, m.q_max = mu, lam, rho0, MaterialModel.ELASTIC, 1e5
    return Scene("harness", cfg, [m], f, dt=dt, flip_ratio=flip_ratio, substeps=1,
                 forces=[GravityForce((0.0, 1.0, 0.0), gravity)])


def stress_of(fields, i):
    """Symmetric Kirchhoff stress of particle i from the w lanes
    (tests/helpers/MPMHarness.cpp:375-391)."""
    txx, tyy, tzz = fields[2, i, 3], fields[3, i, 3], fields[4, i, 3]
    txy, txz, tyz = fields[5, i, 3], fields[6, i, 3], fields[7, i, 3]
    return np.array([[txx, txy, txz], [txy, tyy, tyz], [txz, tyz, tzz]])


def F_of(fields, i):
    """Deformation gradient (math matrix, columns from F0..F2)."""
    return np.stack([fields[2, i, :3], fields[3, i, :3], fields[4, i, :3]], axis=1)


def rel_err(a, b, floor=1e-30):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), floor)
    return float(np.abs(a - b).max() / scale)
