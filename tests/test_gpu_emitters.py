"""Emitters through the engine (nb_mpm_add_emitter / nb_mpm_emit / step):
the appended particles equal the oracle's restatement of
MPMEngine::emitFromEmitters bit for bit (F = I, B = 0, tau = 0), and a step
with emitters equals appending the same particles and then substepping, as
MPMEngine::step emits before its substeps (src/engine/MPMEngine.cpp:404-406)."""
import numpy as np
import pytest

import oracle
from nebula_amd import mpm
from nebula_amd.mpm import MPMConfig
from nebula_amd.scenes import _material

pytestmark = pytest.mark.gpu


def _emitters():
    a = mpm.AABBEmitter.FromAABB((0.2, 0.5, 0.2), (0.4, 0.7, 0.4), (0.0, -1.0, 0.0), 300)
    a.step_count, a.center_vel = 0, (0.5, 0.0, 0.0)
    s = mpm.SphereEmitter.FromSphere((0.7, 0.6, 0.6), 0.08, (0.0, 0.0, 2.0), 200)
    s.emitAlongNormal, s.step_count, s.particleType = True, 2, 1
    el = mpm.EllipseEmitter()
    el.center, el.semiA, el.semiB, el.particles_per_step = (0.5, 0.3, 0.5), 0.1, 0.05, 150
    el.velocityRandomness, el.step_count = 0.3, -1
    return [a, s, el]


def _engine(cap=2000):
    cfg = MPMConfig(nx=8, ny=8, nz=8, domainSize=(1.0, 1.0, 1.0), cellSize=1.0 / 32, maxParticles=cap, E=2e3,
                    nu=0.3, rho0=1e3)
    e = mpm.MPMEngine(cfg)
    e.setMaterials([_material(2e3, 0.3, 1e3, mpm.MaterialModel.ELASTIC),
                    _material(4e3, 0.3, 1e3, mpm.MaterialModel.ELASTIC)])
    e.setForces([mpm.GravityForce((0.0, 1.0, 0.0), -9.8)])
    e.resetParticles()
    e.numSubsteps = 2
    return e


def _records(ems):
    return b"".join(bytes(x.pack()) for x in ems), len(ems)


def test_emitted_particles_equal_the_oracle(gpu_lib):
    e = _engine(cap=1300)  # calls emit 650, 500, then 300 cut to 150 by the capacity
    ems = _emitters()
    rec, n = _records(ems)
    for x in ems:
        e.addEmitter(x)
    Vp = np.float32(1.0 / 8) ** 3
    for _ in range(3):
        e.emitFromEmitters(0.01)
    pos, vel = oracle.emit(rec, n, 3, 0.01, 1300, 0, float(Vp))
    assert e.liveParticleCount() == pos.shape[0] == 1300
    st = e.getState()
    assert np.array_equal(st[0].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(st[1].view(np.uint32), vel.view(np.uint32))
    eye = np.zeros((3, 1300, 4), np.float32)
    for c in range(3):
        eye[c, :, c] = 1.0
    assert np.array_equal(st[2:5], eye) and not st[5:8].any()
    assert ems[0].center[0] == pytest.approx(0.3 + 3 * 0.5 * 0.01)  # the box moved with each call


def test_step_emits_then_substeps(gpu_lib):
    dt = 2e-3
    a = _engine()
    ems = _emitters()
    rec, n = _records(ems)
    for x in ems:
        a.addEmitter(x)
    b = _engine()
    Vp = float(np.float32(1.0 / 8) ** 3)
    prev = 0
    for k in range(1, 4):
        a.step(dt)
        pos, vel = oracle.emit(rec, n, k, dt, 2000, 0, Vp)
        b.appendParticles(pos[prev:], vel[prev:])
        prev = pos.shape[0]
        b.substep(dt / 2, 2)
    a.sync()
    b.sync()
    assert a.liveParticleCount() == b.liveParticleCount() == prev
    assert np.array_equal(a.getState().view(np.uint32), b.getState().view(np.uint32))
    a.resetParticles()
    assert a.liveParticleCount() == 0
