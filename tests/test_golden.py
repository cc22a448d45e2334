"""The oracle against the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py): regenerating every fixture from its
stored inputs must give the stored outputs bit for bit, so a change to the
oracle's arithmetic cannot go unnoticed."""
import os

import numpy as np
import pytest

from golden import make_golden

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["substep_elastic_apic", "substep_dp_pic", "bins_edge", "mesh_box_sdf", "noise", "emit", "forces_grid", "hash_pbf"]


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def regenerated():
    return make_golden.build_all()


@pytest.mark.parametrize("name", NAMES)
def test_fixture_reproduces_bit_for_bit(regenerated, name):
    stored = load(name)
    new = regenerated[name]
    assert sorted(stored.files) == sorted(new.keys())
    for k in stored.files:
        a, b = stored[k], np.asarray(new[k])
        assert a.dtype == b.dtype and a.shape == b.shape, f"{name}.{k}"
        assert a.tobytes() == b.tobytes(), f"{name}.{k} differs"


def test_substep_fixtures_fp32_close_to_fp64():
    for name in ("substep_elastic_apic", "substep_dp_pic"):
        d = load(name)
        h = 1.0 / int(d["res"])
        dx = np.abs(d["state32"][0, :, :3] - d["state64"][0, :, :3]).max() / h
        assert dx < 1e-4, f"{name}: FP32 and FP64 oracles {dx:.2e} dx apart"
