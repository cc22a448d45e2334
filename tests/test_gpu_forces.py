"""Turbulence and noise forces in the grid update against the FP32 oracle.

The forces' hash fract(sin(a) * 43758.5453) (shaders/force_noise.glsl:7-12)
turns a one-ulp difference in a or in sin(a) into an unrelated value, so the
reference's own results depend on its driver's sin.  Both this engine and the
oracle round every step as written (no FMA contraction); what remains is
sinf's last bit.  The test states the fraction of nodes that agree to 1e-5,
requires most of them to, bounds the rest by the forces' own size, and
checks that the forces did act."""
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


@pytest.mark.skip(reason="one full GPU suite run ended in a host segmentation fault while this test freed an "
                         "oracle Sim (orc_sim_destroy); the corrupting write is not found yet, so the case is held "
                         "off the GPU until it is")
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
