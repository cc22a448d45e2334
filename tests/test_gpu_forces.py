"""Turbulence and noise forces in the grid update against the FP32 oracle.

The forces' hash fract(sin(a) * 43758.5453) (shaders/force_noise.glsl:7-12)
magnifies a one-ulp difference in sin(a) 43758-fold, so the reference's own
results depend on its driver's sin.  This engine and the oracle both round
every step as written (no FMA contraction) and take the correctly rounded
FP32 sine (FP64 sin rounded once), so they agree to the last bit but for the
rare argument whose FP64 sines round differently.  The test states the
fraction of nodes that agree to 1e-5, requires all but 0.1 % of them to,
bounds the rest by the forces' own size, and checks that the forces did act.

The first test compares with tests/golden/forces_grid.npz (the oracle's grid
update, computed on the CPU by tests/golden/make_golden.py), so the GPU
process runs no oracle code.  The second computes the oracle's grid in the
same process and stays held off the GPU (reason in its skip marker)."""
import os

import numpy as np

from nebula_amd import mpm
from nebula_amd.scenes import elastic_block_64, make_engine
from nbtest import oracle_sim
from test_gpu_parity import _evolved
import pytest

pytestmark = pytest.mark.gpu


def _grid_after_update(sc, st, om, oms):
    sim = oracle_sim(sc)
    sim.set_state(st)
    NC = sim.NC
    sim.set_grid(om[:NC], oms)
    sim.grid_update()
    e = make_engine(sc)
    e.setGrid(om[:NC], oms)
    e.runGridUpdate(sc.dt)
    gm, _ = e.getGrid()
    o2, oms2 = sim.get_grid()
    return gm[:NC, :3], o2[:NC, :3], oms2 >= 1e-10


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forces_grid.npz")


def _fixture_forces():
    return [mpm.GravityForce((0.0, 1.0, 0.0), -9.8),
            mpm.TurbulenceForce(strength=3.0, frequency=1.5, octaves=3, seed=0.37),
            mpm.NoiseForce((0.3, 0.8, 0.1), strength=4.0, frequency=2.0, octaves=2, seed=1.1)]


def test_turbulence_and_noise_forces_match_golden_grid(gpu_lib):
    d = np.load(GOLDEN, allow_pickle=False)
    forces = _fixture_forces()
    assert mpm.forces_bytes(forces) == d["forces"].tobytes()
    res, dt = int(d["res"]), float(d["dt"])
    sc = elastic_block_64(res=res)
    sc.forces = forces
    e = make_engine(sc)
    NC = sc.total_cells()
    ids = d["node_ids"].astype(np.int64)
    mom = np.zeros((NC, 4), np.float32)
    mass = np.zeros(NC, np.float32)
    mom[ids], mass[ids] = d["mom_in"], d["mass_in"]
    e.setGrid(mom, mass)
    e.runGridUpdate(dt)
    gm, _ = e.getGrid()
    g = gm[ids, :3].astype(np.float64)
    o = d["vel_forces"][:, :3].astype(np.float64)
    kick = np.abs(o - d["vel_gravity"][:, :3]).max(axis=1)
    assert np.median(kick) > 1e-3, "the turbulence and noise forces must move the grid velocities"
    err = np.abs(g - o).max(axis=1)
    agree = float(np.mean(err <= 1e-5 * np.maximum(1.0, np.abs(o).max(axis=1))))
    exact = float(np.mean(np.all(gm[ids, :3] == d["vel_forces"][:, :3], axis=1)))
    print(f"  forces (golden): {ids.size} nodes, {agree:.4f} agree to 1e-5 ({exact:.4f} bit-exact), "
          f"max diff {err.max():.2e}, median kick {np.median(kick):.2e}", flush=True)
    assert agree >= 0.999
    assert err.max() <= 2.5 * kick.max() + 1e-5  # a differing node is off by at most its own kick


@pytest.mark.skip(reason="one full GPU suite run ended in a host segmentation fault while this test freed an "
                         "oracle Sim (orc_sim_destroy); no cause in this repository's code has been shown, so this "
                         "in-process form stays off the GPU; test_turbulence_and_noise_forces_match_golden_grid "
                         "checks the same forces against the oracle's grid computed on the CPU")
def test_turbulence_and_noise_forces_match_oracle(gpu_lib):
    base = elastic_block_64()
    st = _evolved(base, 10)
    sim = oracle_sim(base)
    sim.set_state(st)
    sim.bin()
    sim.p2g()
    om, oms = sim.get_grid()
    forces = [mpm.TurbulenceForce(strength=3.0, frequency=1.5, octaves=3, seed=0.37),
              mpm.NoiseForce((0.3, 0.8, 0.1), strength=4.0, frequency=2.0, octaves=2, seed=1.1)]
    sc = elastic_block_64()
    sc.forces = forces
    g, o, live = _grid_after_update(sc, st, om, oms)
    sc0 = elastic_block_64()
    sc0.forces = []
    _, o0, _ = _grid_after_update(sc0, st, om, oms)
    g, o, o0 = g[live].astype(np.float64), o[live].astype(np.float64), o0[live].astype(np.float64)
    kick = np.abs(o - o0).max(axis=1)
    assert np.median(kick) > 1e-4, "the forces must move the grid velocities"
    err = np.abs(g - o).max(axis=1)
    agree = float(np.mean(err <= 1e-5 * np.maximum(1.0, np.abs(o).max(axis=1))))
    print(f"  forces: {live.sum()} nodes, {agree:.4f} agree to 1e-5, max diff {err.max():.2e}, "
          f"median kick {np.median(kick):.2e}", flush=True)
    assert agree >= 0.9
    assert err.max() <= 2.5 * kick.max() + 1e-5  # a differing node is off by at most its own kick
