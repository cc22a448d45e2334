"""Host-side checks of the torch binding's argument validation (no GPU).

torch is imported inside the tests, not at collection: on a GPU machine the
GPU suite collects this file too, and a collection-time import would put
torch's HIP runtime into that process."""
import pytest

from nebula_amd import mpm

_HOST_CHECK = """
import sys
import torch
sys.path.insert(0, %r)
from nebula_amd import mpm
from nebula_amd.autograd import rollout

class FakeEngine:
    def materials(self):
        return [mpm.MaterialParams(), mpm.MaterialParams()]

e = FakeEngine()
p = torch.zeros(2, 4)
for bad in (torch.zeros(7, 10, 4), torch.zeros(8, 10, 4, dtype=torch.float64), torch.zeros(8, 10, 4)):
    try:
        rollout(e, bad, p, 2, 1e-3)
    except ValueError as exc:
        print("rejected:", exc)
    else:
        print("accepted a bad state")
"""


def test_rollout_rejects_bad_state_shapes_and_devices():
    """A wrong shape, a wrong dtype and a host tensor each raise ValueError
    (in a fresh process that imports torch first, as the binding requires)."""
    import subprocess
    import sys

    pytest.importorskip("torch")
    root = mpm.os.path.dirname(mpm.os.path.dirname(mpm.os.path.abspath(mpm.__file__)))
    r = subprocess.run([sys.executable, "-c", _HOST_CHECK % root], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("rejected:") == 3 and "accepted" not in r.stdout, r.stdout


def test_two_hip_runtimes_are_detected(monkeypatch):
    """check_single_hip_runtime raises when the maps list two runtime files."""
    monkeypatch.setattr(mpm, "hip_runtimes_mapped",
                        lambda: ["/opt/rocm/lib/libamdhip64.so.7.2", "/site-packages/torch/lib/libamdhip64.so"])
    with pytest.raises(mpm.NebulaError, match="two HIP runtimes"):
        mpm.check_single_hip_runtime()
    monkeypatch.setattr(mpm, "hip_runtimes_mapped", lambda: ["/opt/rocm/lib/libamdhip64.so.7.2"])
    mpm.check_single_hip_runtime()


def test_library_loaded_before_torch_maps_two_runtimes_and_is_refused():
    """In a fresh process: the library first, then torch, maps two HIP
    runtimes (the loader hazard check_single_hip_runtime guards against), and
    the torch binding then refuses to import."""
    import subprocess
    import sys

    pytest.importorskip("torch")
    root = mpm.os.path.dirname(mpm.os.path.dirname(mpm.os.path.abspath(mpm.__file__)))
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from nebula_amd import mpm; mpm.lib()\n"
            "import torch\n"
            "print('runtimes', len(mpm.hip_runtimes_mapped()))\n"
            "try:\n"
            "    import nebula_amd.autograd\n"
            "    print('autograd imported')\n"
            "except mpm.NebulaError as e:\n"
            "    print('refused', e)\n" % root)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "runtimes 2" in r.stdout, r.stdout
    assert "refused two HIP runtimes" in r.stdout, r.stdout
