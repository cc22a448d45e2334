"""The grid update fused into G2P (k_g2p_tile, every substep of a call but
the last) against the separate k_grid + k_g2p launches (a call of one
substep), bit for bit: n substeps in one call must equal n calls of one
substep.  Scenes: APIC with analytic colliders and a moving sphere, the
terrain SDF (push-out and the dense-SDF BC), FLIP, PIC, a batch of scenes,
and particles placed outside the grid (whose stencils leave their tile's
node box and take the range-checked gather)."""
import numpy as np
import pytest

from nebula_amd.scenes import elastic_block_64, make_batch_engine, make_engine, multi_material_128, snow_terrain_256
from test_gpu_batch import _scenes

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _one_call_vs_single_calls(make, dt, n, prof_check=None):
    a, b = make(), make()
    a.setFusedG2P(True)
    a.substep(dt, n)
    a.sync()
    for _ in range(n):
        b.substep(dt, 1)
    b.sync()
    sa, sb = a.getState(), b.getState()
    diff = np.argwhere(sa.view(np.uint32) != sb.view(np.uint32))
    assert _same(sa, sb), f"{len(diff)} differing words, first {diff[:5].tolist()}"
    return a


@pytest.mark.parametrize("mode", [-1.0, 0.0, 0.95], ids=["apic", "pic", "flip"])
def test_fused_equals_separate_multi_material(gpu_lib, mode):
    sc = multi_material_128()
    sc.flip_ratio = mode
    _one_call_vs_single_calls(lambda: make_engine(sc), sc.dt, 6)


def test_fused_equals_separate_snow_terrain(gpu_lib):
    sc = snow_terrain_256(res=128, layer=6)
    e = _one_call_vs_single_calls(lambda: make_engine(sc), sc.dt, 5)
    prof = e.profileSubsteps(sc.dt, 4)
    calls = e.profileCalls()
    assert calls.get("g2p_grid") == 3 and calls.get("g2p") == 1 and calls.get("grid_update") == 1, calls
    assert prof["g2p_grid"] > 0.0


def test_fused_equals_separate_batch(gpu_lib):
    scenes = _scenes()
    _one_call_vs_single_calls(lambda: make_batch_engine(scenes), scenes[0].dt, 6)


def test_fused_equals_separate_with_particles_outside_the_grid(gpu_lib):
    sc = elastic_block_64()
    sc.corrected_svd = True
    f = sc.fields.copy()
    h = sc.cfg.cellSize
    f[0, :40, 0] = -0.5 * h - np.arange(40, dtype=np.float32) * 0.05 * h  # left of the grid
    f[0, 40:80, 1] = 1.0 + 0.3 * h                                        # above it
    sc.fields = f

    def make():
        e = make_engine(sc)
        e.setState(f)
        return e

    _one_call_vs_single_calls(make, sc.dt, 3)
