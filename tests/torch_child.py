"""Checks of the torch binding (nebula_amd.autograd), run in a child process
by tests/test_gpu_autograd.py.

torch is imported here before the engine library loads, so this process maps
one HIP runtime, torch's: torch asks the loader for libamdhip64.so and the
library for libamdhip64.so.7, and the loader maps a second runtime when the
library came first.  The pytest process itself never imports torch, so the
C-ABI tests run on the runtime the library was built against.

Usage: python tests/torch_child.py [case ...]   (no case: all of them).
Prints one line per case, "case: ok" or "case: FAIL <message>", and exits
with status 1 when any case failed."""
import os
import sys
import traceback

import torch  # noqa: E402  (first: see above)

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

from nebula_amd import mpm  # noqa: E402
from nebula_amd.autograd import material_params, material_params_from, rollout  # noqa: E402
from nebula_amd.scenes import batch_layers_64, make_engine  # noqa: E402
from test_gpu_adjoint import N_SUB, SEG, _mask, _scene  # noqa: E402


def _direct(sc, W):
    e = make_engine(sc)
    e.adjointBegin(N_SUB, SEG)
    e.forwardRecorded(sc.dt)
    final = e.getState()
    e.setAdjoint(W)
    dpar = e.backward()[: len(sc.materials)]
    grad = e.getAdjoint()
    e.adjointEnd()
    return final, grad, dpar


def rollout_matches_direct_reverse_mode():
    sc = _scene()
    n = sc.n
    W = (np.random.default_rng(3).standard_normal((8, n, 4)) * _mask(n)).astype(np.float32)
    final, grad, dpar = _direct(sc, W)

    for tapes in (1, 0):  # one segment's states kept at once, or as many as fit
        e = make_engine(sc)
        state = torch.tensor(sc.fields, device="cuda", requires_grad=True)
        params = material_params(e, "cuda").requires_grad_()
        out = rollout(e, state, params, N_SUB, sc.dt, SEG, tapes)
        loss = (torch.tensor(W, device="cuda") * out).sum()
        loss.backward()
        torch.cuda.synchronize()
        assert np.array_equal(out.detach().cpu().numpy().view(np.uint32), final.view(np.uint32))
        assert np.array_equal(state.grad.cpu().numpy(), grad)
        assert np.array_equal(params.grad.cpu().numpy(), dpar)
        assert float(state.grad[1, :, 3].abs().max()) == 0.0  # material id lane


def rollout_params_enter_the_forward():
    """A changed mu in params changes the run exactly as a changed material table does."""
    sc = _scene()
    e = make_engine(sc)
    params = material_params(e, "cuda")
    params[1, 0] *= 1.5
    out = rollout(e, torch.tensor(sc.fields, device="cuda"), params, N_SUB, sc.dt)
    mats = e.materials()
    sc2 = _scene()
    m1 = mats[1]
    m1.mu = float(params[1, 0])
    sc2.materials[1] = m1
    ref = make_engine(sc2)
    ref.substep(sc.dt, N_SUB)
    ref.sync()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.getState().view(np.uint32))


def rollout_backward_after_another_recording_raises():
    sc = _scene()
    e = make_engine(sc)
    state = torch.tensor(sc.fields, device="cuda", requires_grad=True)
    params = material_params(e, "cuda")
    out1 = rollout(e, state, params, 2, sc.dt)
    out2 = rollout(e, state, params, 2, sc.dt)
    try:
        out1.sum().backward()
    except RuntimeError:
        pass
    else:
        raise AssertionError("the backward of a superseded rollout must raise")
    out2.sum().backward()
    assert state.grad is not None


def rollout_rejects_params_of_the_wrong_shape():
    e = make_engine(_scene())
    try:
        rollout(e, torch.zeros(8, 10, 4, device="cuda"), torch.zeros(3, 4), 2, 1e-3)
    except ValueError:
        return
    raise AssertionError("params of the wrong shape must raise ValueError")


def chain_rule_to_young_poisson_friction():
    """Gradients with respect to (E, nu, friction angle, cohesion) through
    material_params_from equal the adjoint's dL/d(mu, lambda, M, q_c) times
    the closed-form Jacobian of mu = E/(2(1+nu)), lambda = E nu/((1+nu)(1-2nu)),
    M = tan(angle)."""
    sc = _scene()
    n = sc.n
    W = torch.tensor((np.random.default_rng(8).standard_normal((8, n, 4)) * _mask(n)).astype(np.float32),
                     device="cuda")
    e = make_engine(sc)
    E = torch.tensor([2e3, 2e3], dtype=torch.float64, device="cuda", requires_grad=True)
    nu = torch.tensor([0.3, 0.3], dtype=torch.float64, device="cuda", requires_grad=True)
    phi = torch.tensor([0.5, float(np.arctan(sc.materials[1].M_friction))], dtype=torch.float64, device="cuda",
                       requires_grad=True)
    qc = torch.tensor([0.0, 2.0], dtype=torch.float64, device="cuda", requires_grad=True)
    params = material_params_from(E, nu, phi, qc)
    params.retain_grad()
    out = rollout(e, torch.tensor(sc.fields, device="cuda"), params, N_SUB, sc.dt, SEG)
    (W * out).sum().backward()
    g = params.grad.cpu().numpy()
    Ev, nv, pv = 2e3, 0.3, phi.detach().cpu().numpy()
    dmu_dE, dmu_dnu = 1 / (2 * (1 + nv)), -Ev / (2 * (1 + nv) ** 2)
    den = (1 + nv) * (1 - 2 * nv)
    dlam_dE = nv / den
    dlam_dnu = Ev * (den - nv * (1 - 2 * nv - 2 * (1 + nv))) / den ** 2
    want_E = g[:, 0] * dmu_dE + g[:, 1] * dlam_dE
    want_nu = g[:, 0] * dmu_dnu + g[:, 1] * dlam_dnu
    want_phi = g[:, 2] / np.cos(pv) ** 2
    for got, want in ((E.grad, want_E), (nu.grad, want_nu), (phi.grad, want_phi), (qc.grad, g[:, 3])):
        assert np.allclose(got.cpu().numpy(), want, rtol=1e-9, atol=1e-12), (got, want)


def recovery_of_hidden_parameters():
    """One scene of the differentiable batch (32,768 particles, an elastic
    layer on a Drucker-Prager layer) with hidden E, nu and friction angle:
    the target is its final positions after 400 substeps; from a guess off by
    +25 % in E, +0.04 in nu and +4 degrees in the angle, Adam on (log E, nu,
    angle) with the loss = MSE of the final positions, gradients through the
    rollout (checkpoints every 20 substeps) and material_params_from."""
    sc = batch_layers_64()[0]
    e = make_engine(sc)
    m0, m1 = sc.materials
    # the scene's hidden values, recovered from its (mu, lambda) and M
    nu_h = m0.lambda_ / (2.0 * (m0.lambda_ + m0.mu))
    E_h = 2.0 * m0.mu * (1.0 + nu_h)
    phi_h = float(np.arctan(m1.M_friction))
    n_sub, seg = 400, 20
    state = torch.tensor(sc.fields, device="cuda")
    qc = torch.tensor([0.0, m1.q_cohesion], dtype=torch.float64, device="cuda")

    def params_of(logE, nu, phi):
        P = material_params_from(torch.exp(logE).expand(2), nu.expand(2), torch.stack([phi * 0.0, phi]), qc)
        return P.to(torch.float32)

    with torch.no_grad():
        target = rollout(e, state, params_of(torch.tensor(np.log(E_h), dtype=torch.float64, device="cuda"),
                                             torch.tensor(nu_h, dtype=torch.float64, device="cuda"),
                                             torch.tensor(phi_h, dtype=torch.float64, device="cuda")),
                         n_sub, sc.dt, seg)[0, :, :3].clone()
    e.adjointEnd()
    logE = torch.tensor(np.log(1.25 * E_h), dtype=torch.float64, device="cuda", requires_grad=True)
    nu = torch.tensor(nu_h + 0.04, dtype=torch.float64, device="cuda", requires_grad=True)
    phi = torch.tensor(phi_h + np.radians(4.0), dtype=torch.float64, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([logE, nu, phi], lr=0.02)
    first = None
    for it in range(40):
        opt.zero_grad()
        out = rollout(e, state, params_of(logE, nu, phi), n_sub, sc.dt, seg)
        loss = ((out[0, :, :3] - target) ** 2).mean()
        loss.backward()
        first = float(loss) if first is None else first
        if it % 5 == 0 or it == 39:
            print(f"  recovery iter {it:2d}: loss {float(loss):.3e}  E {float(torch.exp(logE)):.1f} (hidden {E_h:.1f})  "
                  f"nu {float(nu):.4f} ({nu_h:.4f})  angle {np.degrees(float(phi)):.2f} deg "
                  f"({np.degrees(phi_h):.2f})", flush=True)
        opt.step()
    last = float(loss)
    errs = (abs(float(torch.exp(logE)) / E_h - 1.0), abs(float(nu) - nu_h), abs(float(phi) - phi_h))
    print(f"  recovery: loss {first:.3e} -> {last:.3e}; |E/E_h - 1| {errs[0]:.3f} (start 0.25), "
          f"|nu - nu_h| {errs[1]:.4f} (start 0.04), |angle - angle_h| {np.degrees(errs[2]):.2f} deg (start 4)",
          flush=True)
    # the final positions pin the friction angle well and E, nu only weakly
    # (40 iterations from the guess above: loss 4.4e-7 -> 3.1e-9, angle 4 ->
    # 0.55 deg, E 25 % -> 22 %, nu 0.040 -> 0.045)
    assert last < 0.02 * first
    assert errs[2] < np.radians(1.0)


def one_hip_runtime_is_mapped():
    rts = mpm.hip_runtimes_mapped()
    assert len(rts) == 1, rts
    print(f"  HIP runtime: {rts[0]}", flush=True)


CASES = {f.__name__: f for f in (one_hip_runtime_is_mapped, rollout_matches_direct_reverse_mode,
                                 rollout_params_enter_the_forward, rollout_backward_after_another_recording_raises,
                                 rollout_rejects_params_of_the_wrong_shape, chain_rule_to_young_poisson_friction,
                                 recovery_of_hidden_parameters)}


def main(names):
    failed = 0
    for name in names or list(CASES):
        print(f"{name}: start", flush=True)
        try:
            CASES[name]()
            print(f"{name}: ok", flush=True)
        except Exception as exc:  # reported to the parent test, which fails
            failed += 1
            traceback.print_exc()
            print(f"{name}: FAIL {type(exc).__name__}: {exc}", flush=True)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
