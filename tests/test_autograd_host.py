"""Host-side checks of the torch binding's argument validation (no GPU)."""
import pytest
import torch

from nebula_amd import mpm
from nebula_amd.autograd import rollout


class _FakeEngine:
    def materials(self):
        return [mpm.MaterialParams(), mpm.MaterialParams()]


def test_rollout_rejects_bad_state_shapes_and_devices():
    e = _FakeEngine()
    p = torch.zeros(2, 4)
    with pytest.raises(ValueError):
        rollout(e, torch.zeros(7, 10, 4), p, 2, 1e-3)
    with pytest.raises(ValueError):
        rollout(e, torch.zeros(8, 10, 4, dtype=torch.float64), p, 2, 1e-3)
    with pytest.raises(ValueError):
        rollout(e, torch.zeros(8, 10, 4), p, 2, 1e-3)  # host tensor


def test_rollout_rejects_params_of_the_wrong_shape():
    e = _FakeEngine()
    if not torch.cuda.is_available():
        # the device check comes first on a host tensor; the params check is reached on a GPU
        pytest.skip("needs a CUDA tensor to reach the params check")
    with pytest.raises(ValueError):
        rollout(e, torch.zeros(8, 10, 4, device="cuda"), torch.zeros(3, 4), 2, 1e-3)
