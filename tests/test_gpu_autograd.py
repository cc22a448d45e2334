"""torch.autograd binding (nebula_amd.autograd.rollout) against the direct
reverse-mode entry points, which tests/test_gpu_adjoint.py checks against
FP64 finite differences of the oracle: the final state, dL/d(initial state)
and dL/d(material parameters) must agree bit for bit."""
import numpy as np
import pytest
import torch

from nebula_amd.autograd import material_params, rollout
from nebula_amd.scenes import make_engine
from test_gpu_adjoint import N_SUB, SEG, _mask, _scene

pytestmark = pytest.mark.gpu


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


def test_rollout_matches_direct_reverse_mode(gpu_lib):
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


def test_rollout_params_enter_the_forward(gpu_lib):
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


def test_rollout_backward_after_another_recording_raises(gpu_lib):
    sc = _scene()
    e = make_engine(sc)
    state = torch.tensor(sc.fields, device="cuda", requires_grad=True)
    params = material_params(e, "cuda")
    out1 = rollout(e, state, params, 2, sc.dt)
    out2 = rollout(e, state, params, 2, sc.dt)
    with pytest.raises(RuntimeError):
        out1.sum().backward()
    out2.sum().backward()
    assert state.grad is not None
