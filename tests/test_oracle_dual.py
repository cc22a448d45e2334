"""The oracle's dual-number (forward-mode) twin, pinned on the CPU.

oracle.dual_substeps runs the FP64 twin with R = Dual (oracle/dual.hpp): the
value part must be the FP64 twin bit for bit (the same double operations),
and the tangents must be the derivatives of the FP64 substep map:
  - a lone particle in free fall, where the map is known in closed form
    (partition of unity and exact reproduction of x_p by the quadratic
    B-spline weights): after n substeps x_n = x_0 + n dt v_0 + dt^2 g n(n+1)/2,
    so dx_n/dx_0 = I, dx_n/dv_0 = n dt I, dv_n/dv_0 = I, dB_n/d(x_0, v_0) = 0;
  - central differences of the FP64 twin on a small elastic block (a smooth
    map: no yield surface, no walls), for every field and for mu and lambda.
"""
import numpy as np
import pytest

import oracle
from nbtest import oracle_sim
from nebula_amd import mpm
from nebula_amd.scenes import _fields, _material

K = None


def _k():
    global K
    if K is None:
        K = oracle.dual_tangents()
    return K


def _scene(res=16, dt=1e-3, flip=-1.0):
    return oracle.make_scene(dt, 1.0 / res, (res, res, res), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), flip)


def _block(seed, n_side=4, res=16):
    rng = np.random.default_rng(seed)
    h = 1.0 / res
    c0 = res // 2 - n_side // 2
    cells = np.stack(np.meshgrid(*[np.arange(c0, c0 + n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = ((cells[:, None, :] + rng.random((len(cells), 8, 3))) * h).reshape(-1, 3).astype(np.float32)
    vel = (0.3 * rng.standard_normal(pos.shape)).astype(np.float32)
    f = _fields(pos, vel, h ** 3 / 8)
    F = np.eye(3) + 0.05 * rng.standard_normal((len(pos), 3, 3))
    for c in range(3):
        f[2 + c, :, :3] = F[:, :, c]
    f[5:8, :, :3] = 0.02 * rng.standard_normal((3, len(pos), 3))
    f[2:8, :, 3] = 50.0 * rng.standard_normal((6, len(pos)))
    return f


def _mats():
    return mpm.materials_bytes([_material(2e3, 0.25, 10.0, mpm.MaterialModel.ELASTIC)])


GRAV = mpm.forces_bytes([mpm.GravityForce((0.0, 1.0, 0.0), -9.8)])


def test_dual_values_are_the_fp64_twin_bit_for_bit():
    f = _block(1)
    n = f.shape[1]
    mid = np.zeros(n, np.uint32)
    for corrected in (False, True):
        sc = _scene()
        tan = np.random.default_rng(2).standard_normal((_k(), 8, n, 4))
        val, _ = oracle.dual_substeps(sc, _mats(), f.astype(np.float64), mid, tan, None, 4, forces=GRAV,
                                      corrected_svd=corrected)
        sim = oracle.Sim(sc, _mats(), b"", GRAV, None, 64)
        sim.set_corrected_svd(corrected)
        sim.set_state_f64(f.astype(np.float64), mid)
        sim.substeps(4)
        want = sim.get_state_f64()
        assert np.array_equal(val[:, :, :3], want[:, :, :3]) and np.array_equal(val[2:, :, 3], want[2:, :, 3])
        assert np.array_equal(val[0, :, 3], want[0, :, 3])


def test_dual_free_fall_of_a_lone_particle_closed_form():
    res, dt, n_sub, g = 16, 1e-3, 6, -9.8
    f = _fields(np.array([[0.47, 0.53, 0.51]], np.float32), np.array([[0.4, -0.2, 0.3]], np.float32),
                (1.0 / res) ** 3 / 8)
    k = _k()
    tan = np.zeros((k, 8, 1, 4))
    tan[0, 0, 0, 0] = 1.0  # d/dx_0
    tan[1, 1, 0, 1] = 1.0  # d/dv_0,y
    tan[2, 0, 0, 2] = 1.0  # d/dz_0
    tan[3, 1, 0, 0] = 1.0  # d/dv_0,x
    mats = mpm.materials_bytes([_material(1e3, 0.2, 1.0, mpm.MaterialModel.ELASTIC)])
    val, dt_ = oracle.dual_substeps(_scene(res, dt), mats, f.astype(np.float64), np.zeros(1, np.uint32), tan, None,
                                    n_sub, forces=GRAV)
    x0 = f[0, 0, :3].astype(np.float64)
    v0 = f[1, 0, :3].astype(np.float64)
    want_x = x0 + n_sub * dt * v0 + np.array([0, g, 0]) * dt * dt * n_sub * (n_sub + 1) / 2
    assert np.allclose(val[0, 0, :3], want_x, atol=1e-12)
    assert np.allclose(val[1, 0, :3], v0 + np.array([0, g * n_sub * dt, 0]), atol=1e-12)
    e = np.eye(3)
    assert np.allclose(dt_[0, 0, 0, :3], e[0], atol=1e-9) and np.allclose(dt_[2, 0, 0, :3], e[2], atol=1e-9)
    assert np.allclose(dt_[1, 0, 0, :3], n_sub * dt * e[1], atol=1e-9)
    assert np.allclose(dt_[3, 0, 0, :3], n_sub * dt * e[0], atol=1e-9)
    assert np.allclose(dt_[0, 1, 0, :3], 0.0, atol=1e-9) and np.allclose(dt_[1, 1, 0, :3], e[1], atol=1e-9)
    assert np.abs(dt_[:, 5:8, 0, :3]).max() < 1e-9  # B stays 0: sum_i w_i (x_i - x_p) = 0 for every x_p


@pytest.mark.parametrize("corrected", [False, True], ids=["refsvd", "corrsvd"])
def test_dual_tangents_match_fp64_central_differences(corrected):
    """Elastic block, 3 substeps: each tangent (one per input kind, random
    directions) against (f(s + h d) - f(s - h d)) / 2h of the FP64 twin with
    h = 1e-6 (relative to the field's scale), every output field."""
    f = _block(3).astype(np.float64)
    n = f.shape[1]
    mid = np.zeros(n, np.uint32)
    rng = np.random.default_rng(4)
    k = _k()
    tan = np.zeros((k, 8, n, 4))
    tan[0, 0, :, :3] = rng.standard_normal((n, 3)) * (1.0 / 16) * 0.1  # positions
    tan[1, 1, :, :3] = rng.standard_normal((n, 3))                      # velocities
    tan[2, 2:5, :, :3] = 0.1 * rng.standard_normal((3, n, 3))           # F
    tan[2, 5:8, :, :3] = 0.05 * rng.standard_normal((3, n, 3))          # B
    tan[2, 2:8, :, 3] = 10.0 * rng.standard_normal((6, n))              # tau
    tm = np.zeros((k, 1, 4))
    tm[3, 0, :2] = [500.0, 300.0]                                       # mu, lambda
    sc = _scene()
    n_sub = 3
    _, dual = oracle.dual_substeps(sc, _mats(), f, mid, tan, tm, n_sub, forces=GRAV, corrected_svd=corrected)

    def run(fields, mats):
        sim = oracle.Sim(sc, mats, b"", GRAV, None, 64)
        sim.set_corrected_svd(corrected)
        sim.set_state_f64(fields, mid)
        sim.substeps(n_sub)
        return sim.get_state_f64()

    h = 1e-6
    base = mpm.materials_bytes([_material(2e3, 0.25, 10.0, mpm.MaterialModel.ELASTIC)])
    for j in range(k):
        if j < 3:
            fd = (run(f + h * tan[j], base) - run(f - h * tan[j], base)) / (2 * h)
        else:
            m0 = _material(2e3, 0.25, 10.0, mpm.MaterialModel.ELASTIC)
            mp, mm = (mpm.MaterialParams.from_buffer_copy(bytes(m0)) for _ in range(2))
            # mu and lambda are FP32 in the record: central differences with steps
            # of 0.5 and 0.3 (a 1e-3 multiple of the tangent), exact in FP32
            hs = 1e-3
            mp.mu, mp.lambda_ = m0.mu + hs * 500.0, m0.lambda_ + hs * 300.0
            mm.mu, mm.lambda_ = m0.mu - hs * 500.0, m0.lambda_ - hs * 300.0
            fd = (run(f, mpm.materials_bytes([mp])) - run(f, mpm.materials_bytes([mm]))) / (2 * hs)
        for lo, hi, name in ((0, 1, "x"), (1, 2, "v"), (2, 5, "F"), (5, 8, "B")):
            a = dual[j, lo:hi, :, :3]
            b = fd[lo:hi, :, :3]
            scale = max(np.abs(b).max(), 1e-12)
            err = np.abs(a - b).max() / scale
            assert err < (1e-5 if j < 3 else 2e-4), f"tangent {j} field {name}: {err:.2e}"
        a, b = dual[j, 2:8, :, 3], fd[2:8, :, 3]
        assert np.abs(a - b).max() / max(np.abs(b).max(), 1e-12) < (1e-5 if j < 3 else 2e-4), f"tangent {j} tau"


def test_dual_tangents_through_the_drucker_prager_return_map():
    """A Drucker-Prager block with finite strains (most particles yield or sit
    at the apex), 2 substeps, proper polar SVD: the tangents through the
    return map (log, exp, the radial return and the apex map) against FP64
    central differences (h = 1e-7; 1e-3 along the stress).  A particle within ~h of the yield
    surface makes the difference jump across the kink, so the bar is per
    particle: 99 % of the particles within 1e-5 of the field's scale."""
    f = _block(5).astype(np.float64)
    n = f.shape[1]
    rng = np.random.default_rng(6)
    F = np.eye(3) + 0.15 * rng.standard_normal((n, 3, 3))
    for c in range(3):
        f[2 + c, :, :3] = F[:, :, c]
    mid = np.zeros(n, np.uint32)
    mats = mpm.materials_bytes([_material(2e3, 0.25, 10.0, mpm.MaterialModel.DRUCKER_PRAGER, M=0.6, qc=5.0)])
    k = _k()
    tan = np.zeros((k, 8, n, 4))
    tan[0, 2:5, :, :3] = 0.1 * rng.standard_normal((3, n, 3))
    tan[1, 1, :, :3] = rng.standard_normal((n, 3))
    tan[2, 0, :, :3] = 0.006 * rng.standard_normal((n, 3))
    tan[3, 2:8, :, 3] = 10.0 * rng.standard_normal((6, n))
    sc = _scene()
    _, dual = oracle.dual_substeps(sc, mats, f, mid, tan, None, 2, forces=GRAV, corrected_svd=True)

    def run(fields):
        sim = oracle.Sim(sc, mats, b"", GRAV, None, 64)
        sim.set_corrected_svd(True)
        sim.set_state_f64(fields, mid)
        sim.substeps(2)
        return sim.get_state_f64()

    # the stress direction moves the state by dt^2 Vp / m times it: a larger
    # step keeps the differences clear of FP64 rounding
    for j, h in zip(range(k), (1e-7, 1e-7, 1e-7, 1e-3)):
        fd = (run(f + h * tan[j]) - run(f - h * tan[j])) / (2 * h)
        for lo, hi, name in ((0, 1, "x"), (1, 2, "v"), (2, 5, "F")):
            a, b = dual[j, lo:hi, :, :3], fd[lo:hi, :, :3]
            scale = max(np.abs(b).max(), 1e-12)
            per = np.abs(a - b).max(axis=(0, 2)) / scale
            assert np.mean(per <= 1e-5) >= 0.99, f"tangent {j} {name}: {np.mean(per <= 1e-5):.3f} within 1e-5"


@pytest.mark.parametrize("corrected", [False, True], ids=["refsvd", "corrsvd"])
def test_dual32_values_are_the_fp32_oracle_and_its_tangents_track_the_fp64_twin(corrected):
    """precision=32 (Dual32): the values are the FP32 oracle's bit for bit
    (the same float operations, each rounded as written), and on a
    well-conditioned elastic block (F = I + 0.05 noise, 4 substeps) the FP32
    tangents stay within 1e-3 of the FP64 twin's on every field."""
    f = _block(1)
    n = f.shape[1]
    mid = np.zeros(n, np.uint32)
    sc = _scene()
    tan = np.random.default_rng(12).standard_normal((_k(), 8, n, 4))
    tan[:, 1, :, 3] = 0.0
    val, t32 = oracle.dual_substeps(sc, _mats(), f.astype(np.float64), mid, tan, None, 4, forces=GRAV,
                                    corrected_svd=corrected, precision=32)
    sim = oracle.Sim(sc, _mats(), b"", GRAV, None, 32)
    sim.set_corrected_svd(corrected)
    fs = f.copy()
    fs[1, :, 3] = mid.view(np.float32)
    sim.set_state(fs)
    sim.substeps(4)
    want = sim.get_state()
    got = val.astype(np.float32)
    assert np.array_equal(got[:, :, :3], want[:, :, :3])
    assert np.array_equal(got[2:, :, 3], want[2:, :, 3]) and np.array_equal(got[0, :, 3], want[0, :, 3])
    _, t64 = oracle.dual_substeps(sc, _mats(), f.astype(np.float64), mid, tan, None, 4, forces=GRAV,
                                  corrected_svd=corrected)
    for lo, hi, name in ((0, 1, "x"), (1, 2, "v"), (2, 5, "F"), (5, 8, "B")):
        a, b = t32[:, lo:hi, :, :3], t64[:, lo:hi, :, :3]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        assert err < 1e-3, f"{name}: FP32 tangents {err:.2e} from FP64"
    with pytest.raises(ValueError):
        oracle.dual_substeps(sc, _mats(), f.astype(np.float64), mid, tan, None, 1, precision=16)
