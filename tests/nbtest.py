"""Shared helpers of the test suite: scene <-> oracle/engine plumbing and the
harness-style set-ups of the reference's per-pass tests
(tests/helpers/MPMHarness.cpp:41-90, 192-251)."""
from __future__ import annotations

import math

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


def degenerate_svd_cases():
    """Two inputs that reach the Jacobi SVD's degenerate branches, with the
    result the reference's loop gives, worked by hand from its text.

    Stall (shaders/mpm_common.glsl:226, t = sign(tau) / ...): F = [[1, 1/4, 0],
    [1/4, 1, 0], [0, 0, 1]] has F^T F = [[17/16, 1/2, 0], [1/2, 17/16, 0],
    [0, 0, 1]], exact in binary.  The pivot (0, 1) has tau = 0, GLSL's
    sign(0) = 0 gives t = 0 (no rotation), and the pivot entry is set to zero
    anyway: the loop ends with D = (17/16, 17/16, 1) and V = I instead of the
    true singular values (5/4, 3/4, 1).  The singular vectors are u_i = F e_i /
    sigma_i.

    Zero singular value (shaders/mpm_common.glsl:274, U[i][i] = 1 when
    sigma_i <= 1e-8): F = diag(5e-9, 8000, 8000) is already diagonal,
    sigma = (5e-9, 8000, 8000) sorts to (8000, 8000, 5e-9) with V's vectors
    reordered to (e1, e2, e0), and the third singular vector is e2 itself
    rather than a completion of the basis: U's vectors are (e1, e2, e2),
    det(U) = 0, so no sign flip.  det F = 0.32 keeps G2P's Jacobian reset
    (|det F| outside [0.15, 10]) from discarding the result, and the Hencky
    log strain of 5e-9 is clamped at log(1e-6).

    Returned as (F, sigma, U, V) with the singular vectors in the rows of U and
    V, the storage svd3 returns."""
    r = math.sqrt(17.0 / 16.0)
    stall = (np.array([[1.0, 0.25, 0.0], [0.25, 1.0, 0.0], [0.0, 0.0, 1.0]]), np.array([r, r, 1.0]),
             np.array([[1.0 / r, 0.25 / r, 0.0], [0.25 / r, 1.0 / r, 0.0], [0.0, 0.0, 1.0]]), np.eye(3))
    zero = (np.diag([5e-9, 8000.0, 8000.0]), np.array([8000.0, 8000.0, 5e-9]),
            np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]),
            np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]))
    return {"stall": stall, "zero_sigma": zero}


def degenerate_hencky(name, mu, lam):
    """Hencky stress of a degenerate case from its hand-worked SVD:
    tau = U diag(lam tr(eps) + 2 mu eps) U^T with U as svd3 stores it,
    eps_i = log(max(|sigma_i|, 1e-6)) (shaders/mpm_common.glsl:284-293)."""
    _, s, U, _ = degenerate_svd_cases()[name]
    eps = np.log(np.maximum(np.abs(s), 1e-6))
    k = lam * eps.sum() + 2 * mu * eps
    return U @ np.diag(k) @ U.T


def jacobian_reset_state(fields):
    """A copy of fields at rest (v = 0, B = 0) whose particles carry F = s I,
    s cycling through 0.5, 0.6, 1.0, 2.1, 2.2 (det F = 0.125, 0.216, 1, 9.26,
    10.65: both sides of both thresholds of G2P's Jacobian reset), and the
    per-particle s."""
    st = fields.copy()
    n = st.shape[1]
    s = np.array([0.5, 0.6, 1.0, 2.1, 2.2], np.float32)[np.arange(n) % 5]
    st[1, :, :3] = 0.0
    st[5:8, :, :3] = 0.0
    st[2:5, :, :3] = 0.0
    for c in range(3):
        st[2 + c, :, c] = s
    return st, s


def check_jacobian_reset(out, s, mu, lam, rtol=2e-6):
    """G2P's result on a zero grid from jacobian_reset_state
    (shaders/mpm_g2p.comp:186-193): where |det F| < 0.15 or > 10, F = I and
    tau = 0 exactly; elsewhere F = s I exactly and tau = diag(3 lam log s +
    2 mu log s) within rtol of the largest such stress."""
    J = s.astype(np.float64) ** 3
    reset = (J < 0.15) | (J > 10.0)
    assert reset.any() and (~reset).any()
    F = out[2:5, :, :3]
    tau = out[2:8, :, 3].astype(np.float64)  # txx, tyy, tzz, txy, txz, tyz
    eye = np.eye(3, dtype=np.float32)
    for c in range(3):
        assert np.array_equal(F[c][reset], np.broadcast_to(eye[c], (int(reset.sum()), 3)))
        kept = np.zeros((int((~reset).sum()), 3), np.float32)
        kept[:, c] = s[~reset]
        assert np.array_equal(F[c][~reset], kept)
    assert np.array_equal(tau[:, reset], np.zeros((6, int(reset.sum()))))
    e = np.log(s[~reset].astype(np.float64))
    k = 3 * lam * e + 2 * mu * e
    want = np.concatenate([np.broadcast_to(k, (3, k.size)), np.zeros((3, k.size))])
    assert np.abs(tau[:, ~reset] - want).max() <= rtol * np.abs(k).max()

