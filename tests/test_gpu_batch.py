"""Batches of scenes (nb_mpm_create_batch) against the same scenes run one by
one on plain engines.

A batch gives each scene its own range of cell keys and leaves every tile's
and every particle's arithmetic as it is, so each scene's state after a run
must equal the plain engine's bit for bit (material ids aside, which the
batch offsets into its shared table).  The same holds for the reverse mode's
state gradients; the material-parameter gradients are summed over blocks
grouped differently, so they agree to rounding."""
import numpy as np
import pytest

from nebula_amd import mpm
from nebula_amd.mpm import MaterialModel
from nebula_amd.scenes import falling_box, make_batch_engine, make_engine

pytestmark = pytest.mark.gpu

N_SUB = 30


def _scenes(flip_ratio=-1.0, corrected=False, seed=7):
    rng = np.random.default_rng(seed)
    specs = [  # grid edge, particles, model, E, cell size
        (16, 900, MaterialModel.ELASTIC, 2e3, 1 / 64),
        (24, 2000, MaterialModel.VON_MISES, 5e4, 1 / 64),
        (40, 6000, MaterialModel.DRUCKER_PRAGER, 3e4, 1 / 64),
        (20, 1500, MaterialModel.ELASTIC, 1e5, 1 / 32),
    ]
    return [falling_box(f"b{i}", g, n, E, 0.3, m, rng, h=h, flip_ratio=flip_ratio, corrected_svd=corrected)
            for i, (g, n, m, E, h) in enumerate(specs)]


def _compare_states(batch_state, singles, scenes):
    off = 0
    for s, (sc, st) in enumerate(zip(scenes, singles)):
        b = batch_state[:, off:off + sc.n]
        ids = b[1, :, 3].view(np.uint32)
        assert np.all(ids == s), f"scene {s}: material ids"
        b = b.copy()
        b[1, :, 3] = st[1, :, 3]
        diff = np.nonzero(b.view(np.uint32) != st.view(np.uint32))
        assert diff[0].size == 0, f"scene {s}: {diff[0].size} values differ, first at {[d[0] for d in diff]}"
        off += sc.n


@pytest.mark.parametrize("flip_ratio", [-1.0, 0.0, 0.95], ids=["apic", "pic", "flip"])
def test_batch_equals_scenes_run_alone(gpu_lib, flip_ratio):
    scenes = _scenes(flip_ratio)
    singles = []
    for sc in scenes:
        e = make_engine(sc)
        e.substep(sc.dt, N_SUB)
        e.sync()
        singles.append(e.getState())
    b = make_batch_engine(scenes)
    assert b.sceneCount() == len(scenes)
    b.substep(scenes[0].dt, N_SUB)
    b.sync()
    _compare_states(b.getState(), singles, scenes)


def test_batch_with_noise_forces_equals_scenes_run_alone(gpu_lib):
    """Turbulence and noise forces evaluate their noise at each node's world
    position in its scene's own coordinates, so the batch's noise form of the
    grid update (BATCH and NOISE) keeps each scene bit for bit."""
    forces = [mpm.GravityForce((0.0, 1.0, 0.0), -9.8),
              mpm.TurbulenceForce(strength=2.0, frequency=3.0, octaves=2, seed=0.3),
              mpm.NoiseForce((1.0, 0.0, 0.0), strength=1.5, frequency=2.0, octaves=2, seed=0.1)]
    scenes = _scenes()
    for sc in scenes:
        sc.forces = forces
    singles = []
    for sc in scenes:
        e = make_engine(sc)
        e.substep(sc.dt, N_SUB)
        e.sync()
        singles.append(e.getState())
    b = make_batch_engine(scenes)
    b.substep(scenes[0].dt, N_SUB)
    b.sync()
    _compare_states(b.getState(), singles, scenes)
    # the forces acted: the same run without them ends elsewhere
    plain = make_engine(_scenes()[0])
    plain.substep(scenes[0].dt, N_SUB)
    plain.sync()
    assert not np.array_equal(plain.getState()[0], singles[0][0])


def test_batch_reverse_mode_equals_scenes_run_alone(gpu_lib):
    scenes = _scenes(corrected=True)
    rng = np.random.default_rng(3)
    Ws = [rng.standard_normal((8, sc.n, 4)).astype(np.float32) for sc in scenes]
    for W in Ws:
        W[1, :, 3] = 0.0
    singles, grads, dpars = [], [], []
    for sc, W in zip(scenes, Ws):
        e = make_engine(sc)
        e.adjointBegin(8, 4)
        e.forwardRecorded(sc.dt)
        singles.append(e.getState())
        e.setAdjoint(W)
        dpars.append(e.backward()[0])
        grads.append(e.getAdjoint())
        e.adjointEnd()
    b = make_batch_engine(scenes)
    b.adjointBegin(8, 4)
    b.forwardRecorded(scenes[0].dt)
    _compare_states(b.getState(), singles, scenes)
    b.setAdjoint(np.concatenate(Ws, axis=1))
    dpar = b.backward()
    g = b.getAdjoint()
    b.adjointEnd()
    off = 0
    for s, sc in enumerate(scenes):
        assert np.array_equal(g[:, off:off + sc.n], grads[s]), f"scene {s}: state gradient"
        np.testing.assert_allclose(dpar[s], dpars[s], rtol=1e-5, atol=1e-7 * np.abs(dpars[s]).max(),
                                   err_msg=f"scene {s}: parameter gradient")
        off += sc.n


def test_batch_scene_table_and_unsupported_calls(gpu_lib):
    scenes = _scenes()
    b = make_batch_engine(scenes)
    first = 0
    for s, sc in enumerate(scenes):
        res, f, c = b.sceneInfo(s)
        edge = round(sc.cfg.domainSize[0] / sc.cfg.cellSize)
        assert res == (edge, edge, edge) and f == first and c == sc.n
        first += sc.n
    with pytest.raises(mpm.NebulaError):
        b.setColliderSphere(0.1, 0.5, 0.5, 0.5)
    with pytest.raises(mpm.NebulaError):
        b.appendParticles(np.zeros((4, 4), np.float32), np.zeros((4, 4), np.float32))
    with pytest.raises(mpm.NebulaError):
        b.getGrid()
    with pytest.raises(mpm.NebulaError):
        b.sceneInfo(len(scenes))
    e = make_engine(scenes[0])
    assert e.sceneCount() == 0


def test_full_256_scene_mix_equals_scenes_run_alone(gpu_lib):
    """The whole 256-scene mix (5.38 M particles, grids of 32 to 96 cells):
    after 5 substeps every scene of the batch equals the scene run alone."""
    from nebula_amd.scenes import batch_mix_256

    scenes = batch_mix_256()
    b = make_batch_engine(scenes)
    b.substep(scenes[0].dt, 5)
    b.sync()
    batch_state = b.getState()
    del b
    singles = []
    for sc in scenes:
        e = make_engine(sc)
        e.substep(sc.dt, 5)
        e.sync()
        singles.append(e.getState())
        del e
    _compare_states(batch_state, singles, scenes)


def test_full_64_scene_differentiable_batch_gradients_equal_scenes_alone(gpu_lib):
    """The 64-scene differentiable batch (2.1 M particles): reverse mode over
    4 substeps in 2-substep segments; every scene's state gradient equals the
    one its scene gets alone, bit for bit."""
    from nebula_amd.scenes import batch_layers_64

    scenes = batch_layers_64()
    rng = np.random.default_rng(9)
    W = rng.standard_normal((8, sum(sc.n for sc in scenes), 4)).astype(np.float32)
    W[1, :, 3] = 0.0
    b = make_batch_engine(scenes)
    b.adjointBegin(4, 2)
    b.forwardRecorded(scenes[0].dt)
    b.setAdjoint(W)
    b.backward()
    g = b.getAdjoint()
    b.adjointEnd()
    del b
    off = 0
    for s, sc in enumerate(scenes):
        e = make_engine(sc)
        e.adjointBegin(4, 2)
        e.forwardRecorded(sc.dt)
        e.setAdjoint(W[:, off:off + sc.n])
        e.backward()
        assert np.array_equal(g[:, off:off + sc.n], e.getAdjoint()), f"scene {s}"
        e.adjointEnd()
        del e
        off += sc.n
