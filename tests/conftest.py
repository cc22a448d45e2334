"""pytest setup: the `gpu` marker, repo root on sys.path, and the CPU oracle
built on demand (g++ only; the HIP library is built by __graft_entry__.build())."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")
    lib = os.path.join(ROOT, "oracle", "build", "liboracle_mpm.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; GPU tests fail loudly (no fallback) when it is absent."""
    from nebula_amd import mpm

    return mpm.lib()
