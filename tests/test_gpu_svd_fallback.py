"""The FP64 SVD fallback of the Drucker-Prager return map (nb_mpm_svd_fallback)
against the FP64 oracle, through the C ABI.

G2P on a zero grid passes F through unchanged (the velocity gradient is
exactly zero), so every particle's return map and Hencky stress act on a
chosen F.  Particles are drawn from six classes, with the reference SVD
convention (F <- U^T exp(eps) V, tau = U^T k U: functions of the individual
singular vectors, shaders/mpm_common.glsl:250-293, shaders/mpm_g2p.comp:41-68):

  separated   singular values (0.99, 0.97, 0.95) +- 0.002, inside the cone
  pair 1e-6   two singular values 1e-6 apart (relative)
  pair 1e-5   two singular values 1e-5 apart
  tiny        singular values (8, 6, 5e-3): s2^2 / s0^2 = 3.9e-7
  stall       the sign(0) stall of degenerate_svd_cases (F^T F with equal
              diagonal entries: no rotation)
  zero sigma  diag(5e-9, 8000, 8000): the unit vector for sigma <= 1e-8

Checked: the separated class is bit-identical with the fallback on and off;
every other class takes the fallback (the counters say so, lane for lane)
and then agrees with the FP64 oracle to a few FP32 ulps in F and tau, where
the FP32 path is off by orders of magnitude more on the close pairs."""
import numpy as np
import pytest

from nbtest import degenerate_svd_cases, oracle_sim
from nebula_amd.mpm import MaterialModel
from nebula_amd.scenes import elastic_block_64, make_engine

pytestmark = pytest.mark.gpu

CLASSES = ("separated", "pair_1e-6", "pair_1e-5", "tiny", "stall", "zero_sigma")


def _rot(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 2] = -q[:, 2]
    return q


def _case():
    sc = elastic_block_64()
    sc.corrected_svd = False
    m = sc.materials[0]
    m.model, m.M_friction, m.q_cohesion = MaterialModel.DRUCKER_PRAGER, 0.577, 5.0
    n = sc.n
    rng = np.random.default_rng(41)
    cls = np.arange(n) % len(CLASSES)
    F = np.zeros((n, 3, 3))
    deg = degenerate_svd_cases()
    for i in range(n):
        c = CLASSES[cls[i]]
        if c == "stall" or c == "zero_sigma":
            F[i] = deg["stall" if c == "stall" else "zero_sigma"][0]
            continue
        if c == "separated":
            s = np.array([0.99, 0.97, 0.95]) + rng.uniform(-0.002, 0.002, 3)
        elif c == "tiny":
            s = np.array([8.0, 6.0, 5e-3])
        else:
            d = 1e-6 if c == "pair_1e-6" else 1e-5
            s0 = 0.97 + rng.uniform(-0.01, 0.01)
            s = np.array([s0 * (1 + d), s0, 0.93])
        F[i] = _rot(rng) @ np.diag(s) @ _rot(rng).T
    F = F.astype(np.float32)
    st = sc.fields.copy()
    st[1, :, :3] = 0.0
    st[5:8, :, :3] = 0.0
    st[2:8, :, 3] = 0.0
    for c in range(3):
        st[2 + c, :, :3] = F[:, :, c]  # column c of F
    return sc, st, cls


def _gpu(sc, st, enabled, stats=False):
    e = make_engine(sc)
    e.setSvdFallback(enabled=enabled, count_stats=stats)
    e.setState(st)
    nc = e.totalCells()
    e.setGrid(np.zeros((nc, 4), np.float32), np.zeros(nc, np.float32))
    e.runBin()
    e.runG2P(sc.dt)
    out = e.getState()
    return out, (e.svdFallbackStats() if stats else None)


def _oracle64(sc, st):
    s = oracle_sim(sc, 64)
    s.set_state_f64(st.astype(np.float64), st[1, :, 3].view(np.uint32))
    nc = s.NC
    s.set_grid(np.zeros((2 * nc, 4), np.float32), np.zeros(nc, np.float32))
    s.bin()
    s.g2p()
    return s.get_state_f64()


def _F_tau(a):
    F = np.asarray(a[2:5, :, :3], np.float64)                  # (3 cols, n, 3 rows)
    tau = np.asarray(a[2:8, :, 3], np.float64)                 # xx yy zz xy xz yz
    return F, tau


def test_svd_fallback_matches_fp64_oracle_and_leaves_other_particles_alone(gpu_lib):
    sc, st, cls = _case()
    on, stats = _gpu(sc, st, True, stats=True)
    off, _ = _gpu(sc, st, False)
    o64 = _oracle64(sc, st)
    n = sc.n
    sep = cls == 0
    # the separated class never takes the fallback: bit-identical on and off
    assert np.array_equal(on[:, sep].view(np.uint32), off[:, sep].view(np.uint32))
    want_fb = int(np.sum(~sep))
    print(f"  svd fallback counters: {stats} (expected {want_fb} FP64 return maps of {n})", flush=True)
    assert stats["dp_lanes"] == n and stats["fp64_return_map_lanes"] == want_fb
    assert stats["fp64_stress_lanes"] == 0 and 0 < stats["fp64_waves"] <= stats["dp_waves"]
    Fo, To = _F_tau(on)
    Ff, Tf = _F_tau(off)
    F6, T6 = _F_tau(o64)
    for k, name in enumerate(CLASSES):
        sel = cls == k
        fs = max(np.abs(F6[:, sel]).max(), 1e-30)
        ts = max(np.abs(T6[:, sel]).max(), 1e-30)
        e_on_F = np.abs(Fo[:, sel] - F6[:, sel]).max() / fs
        e_on_T = np.abs(To[:, sel] - T6[:, sel]).max() / ts
        e_off_F = np.abs(Ff[:, sel] - F6[:, sel]).max() / fs
        e_off_T = np.abs(Tf[:, sel] - T6[:, sel]).max() / ts
        print(f"  {name:10s}: vs FP64 oracle, F {e_on_F:.1e} (FP32 path {e_off_F:.1e}), "
              f"tau {e_on_T:.1e} (FP32 path {e_off_T:.1e})", flush=True)
        assert np.isfinite(Fo[:, sel]).all() and np.isfinite(To[:, sel]).all()
        if name == "separated":
            continue
        # FP64 result rounded once to FP32: a few ulps of the class's scale
        assert e_on_F <= 4e-7 and e_on_T <= 2e-6, name
        if name.startswith("pair"):
            assert e_off_F >= 100 * e_on_F, f"{name}: the FP32 path should be far off on a close pair"
