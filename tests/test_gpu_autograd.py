"""torch.autograd binding (nebula_amd.autograd.rollout) against the direct
reverse-mode entry points, which tests/test_gpu_adjoint.py checks against
FP64 finite differences of the oracle: the final state, dL/d(initial state)
and dL/d(material parameters) must agree bit for bit.

The checks run in one child process (tests/torch_child.py) that imports
torch before the engine library, so that it holds a single HIP runtime;
this pytest process never imports torch."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

CHILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torch_child.py")


@pytest.fixture(scope="module")
def child_report(gpu_lib):
    r = subprocess.run([sys.executable, "-u", CHILD], capture_output=True, text=True, timeout=900)
    print(r.stdout[-4000:], flush=True)
    if r.returncode not in (0, 1):
        print(r.stderr[-4000:], flush=True)
        pytest.fail(f"torch child process ended with status {r.returncode}")
    status = {}
    for ln in r.stdout.splitlines():
        name, _, rest = ln.partition(": ")
        if rest == "ok" or rest.startswith("FAIL"):
            status[name] = rest
    return status, r.stderr


@pytest.mark.parametrize("case", ["one_hip_runtime_is_mapped", "rollout_matches_direct_reverse_mode",
                                  "rollout_params_enter_the_forward",
                                  "rollout_backward_after_another_recording_raises",
                                  "rollout_rejects_params_of_the_wrong_shape",
                                  "chain_rule_to_young_poisson_friction", "recovery_of_hidden_parameters"])
def test_torch_binding(child_report, case):
    status, err = child_report
    assert case in status, f"{case} did not report (stderr tail: {err[-2000:]})"
    assert status[case] == "ok", status[case] + "\n" + err[-3000:]
