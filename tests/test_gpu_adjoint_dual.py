"""Reverse mode of the substep (nb_mpm_backward) against the oracle's
dual-number twin (forward-mode derivatives of the FP64 substep map,
oracle/dual.hpp, pinned in tests/test_oracle_dual.py).

For L = <W, state after n substeps> (W random), the GPU adjoint's directional
derivatives <dL/d(input), D> along four random directions (positions;
velocities; F, tau and B together; every material's mu, lambda, M and q_c)
are compared with the dual twin's exact derivatives of L along the same
directions.  No step size enters, so the comparison holds over many substeps
of plastic flow, where central differences move particles across the yield
surface (tests/test_gpu_adjoint.py stops at 2-4 substeps for that reason).
What remains is that the GPU differentiates its FP32 trajectory and the twin
its FP64 one: a particle whose trial state lies within FP32 rounding of the
yield surface can take the other branch."""
import numpy as np
import pytest

import oracle
from nebula_amd import mpm
from nebula_amd.scenes import batch_layers_64, make_engine
from test_gpu_adjoint import GROUPS, _mask, _scene

pytestmark = pytest.mark.gpu

PARAMS = ("mu", "lambda_", "M_friction", "q_cohesion")


def _directions(sc, rng):
    n = sc.n
    K = oracle.dual_tangents()
    assert K == 4
    h = sc.cfg.cellSize
    D = np.zeros((K, 8, n, 4))
    D[0, 0, :, :3] = 1e-2 * h * rng.standard_normal((n, 3))
    D[1, 1, :, :3] = rng.standard_normal((n, 3))
    for k, sl in GROUPS["F"] + GROUPS["B"]:
        D[2, k, :, sl] = 1e-2 * rng.standard_normal(D[2, k, :, sl].shape)
    for k, sl in GROUPS["tau"]:
        D[2, k, :, sl] = 10.0 * rng.standard_normal(D[2, k, :, sl].shape)
    nm = len(sc.materials)
    Dm = np.zeros((K, nm, 4))
    for mi, m in enumerate(sc.materials):
        for j, pn in enumerate(PARAMS):
            Dm[3, mi, j] = 0.01 * getattr(m, pn) * rng.standard_normal()
    return D, Dm


def _compare(sc, n_sub, seg, tol, seed=7):
    n = sc.n
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((8, n, 4)) * _mask(n)
    e = make_engine(sc)
    e.adjointBegin(n_sub, seg)
    e.forwardRecorded(sc.dt)
    e.setAdjoint(W.astype(np.float32))
    dpar = e.backward()[:len(sc.materials)].astype(np.float64)
    grad = e.getAdjoint().astype(np.float64)
    e.adjointEnd()
    del e

    D, Dm = _directions(sc, rng)
    f0 = sc.fields.astype(np.float64)
    mid = sc.fields[1, :, 3].view(np.uint32)
    f0[1, :, 3] = 0.0
    scn = oracle.make_scene(sc.dt, sc.cfg.cellSize, sc.grid_res(), (0.0, 0.0, 0.0), tuple(sc.cfg.domainSize),
                            sc.flip_ratio)
    _, tan = oracle.dual_substeps(scn, mpm.materials_bytes(sc.materials), f0, mid, D, Dm, n_sub,
                                  forces=mpm.forces_bytes(sc.forces), sdf=sc.sdf, corrected_svd=True)
    worst = 0.0
    for j, name in enumerate(("x", "v", "F/tau/B", "materials")):
        dual = float(np.sum(W * tan[j]))
        adj = float(np.sum(grad * D[j])) + float(np.sum(dpar * Dm[j]))
        err = abs(adj - dual) / max(abs(dual), 1e-30)
        worst = max(worst, err)
        print(f"  {sc.name} {n_sub} substeps (segments of {seg}): d/d{name}: adjoint {adj:+.7e}  dual {dual:+.7e}  "
              f"rel {err:.2e}", flush=True)
    assert worst <= tol, f"worst relative difference {worst:.2e} > {tol}"
    return worst


@pytest.mark.parametrize("flip_ratio", [-1.0, 0.0], ids=["apic", "pic"])
def test_adjoint_matches_dual_twin_over_20_substeps(gpu_lib, flip_ratio):
    """The 512-particle elastic + Drucker-Prager block (particles on both
    sides of the yield surface and at the apex), 20 substeps at dt 2e-3 in
    segments of 5 (recomputed from checkpoints)."""
    _compare(_scene(flip_ratio=flip_ratio), 20, 5, 1e-3)


def test_adjoint_matches_dual_twin_on_a_differentiable_batch_scene(gpu_lib):
    """One scene of the differentiable batch at full size (32,768 particles on
    64^3, an elastic layer on a Drucker-Prager layer, proper polar SVD), 20
    substeps in segments of 5."""
    _compare(batch_layers_64()[0], 20, 5, 1e-3)
