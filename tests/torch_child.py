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
from nebula_amd.autograd import material_params, rollout  # noqa: E402
from nebula_amd.scenes import make_engine  # noqa: E402
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

    e = make_engine(sc)
    state = torch.tensor(sc.fields, device="cuda", requires_grad=True)
    params = material_params(e, "cuda").requires_grad_()
    out = rollout(e, state, params, N_SUB, sc.dt, SEG)
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


def one_hip_runtime_is_mapped():
    rts = mpm.hip_runtimes_mapped()
    assert len(rts) == 1, rts
    print(f"  HIP runtime: {rts[0]}", flush=True)


CASES = {f.__name__: f for f in (one_hip_runtime_is_mapped, rollout_matches_direct_reverse_mode,
                                 rollout_params_enter_the_forward, rollout_backward_after_another_recording_raises,
                                 rollout_rejects_params_of_the_wrong_shape)}


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
