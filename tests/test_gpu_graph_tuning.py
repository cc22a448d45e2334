"""Graph replay (nb_mpm_graph_substeps) and launch shapes / the autotuner
(nb_mpm_set_tuning, nb_mpm_autotune).

Replays of the captured substep must equal eager substeps bit for bit after
every round, with eager work (state downloads) between the rounds and with
changes that must and must not trigger a new capture.  The autotuner must
leave the state as it found it, and the shapes it explores without
NB_TUNE_ROUNDING must not change results."""
import numpy as np
import pytest

from nebula_amd import mpm
from nebula_amd.scenes import make_batch_engine, make_engine
from test_gpu_batch import _scenes

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("kind", ["plain", "batch"])
def test_graph_replays_equal_eager_substeps(gpu_lib, kind):
    scenes = _scenes()
    make = (lambda: make_engine(scenes[2])) if kind == "plain" else (lambda: make_batch_engine(scenes))
    eager, graph = make(), make()
    dt = scenes[0].dt
    for rnd in range(6):
        if rnd == 3:  # new dt: a new capture
            dt *= 0.5
        if rnd == 4:  # new material values, same count: the graph reads the table at run time
            mats = eager.materials()
            mats[0].mu *= 1.3
            eager.setMaterials(mats)
            graph.setMaterials(mats)
        eager.substep(dt, 3)
        eager.sync()
        graph.graphSubstep(dt, 3)
        graph.sync()
        assert _same(eager.getState(), graph.getState()), f"round {rnd}"


def test_graph_recaptures_after_state_and_tuning_changes(gpu_lib):
    sc = _scenes()[1]
    eager, graph = make_engine(sc), make_engine(sc)
    graph.graphSubstep(sc.dt, 2)
    eager.substep(sc.dt, 2)
    # the caller rewrites the state between calls
    st = eager.getState()
    st[1, :, :3] *= 0.5
    eager.setState(st)
    graph.setState(st)
    t = graph.tuning()
    t.tile_blocks, t.g2p_block = 64, 128
    graph.setTuning(t)
    eager.substep(sc.dt, 4)
    eager.sync()
    graph.graphSubstep(sc.dt, 4)
    graph.sync()
    assert _same(eager.getState(), graph.getState())


def test_autotune_keeps_state_and_results(gpu_lib):
    scenes = _scenes()
    ref, tuned = make_batch_engine(scenes), make_batch_engine(scenes)
    before = tuned.getState()
    t = tuned.autotune(scenes[0].dt, trial_substeps=2)
    assert t.ms_per_substep > 0.0
    assert t.g2p_block in (64, 128, 256) and t.tile_blocks >= 1 and t.p2g_item == 12
    assert _same(tuned.getState(), before)
    ref.substep(scenes[0].dt, 10)
    ref.sync()
    tuned.substep(scenes[0].dt, 10)
    tuned.sync()
    assert _same(ref.getState(), tuned.getState())


def test_p2g_item_changes_results_only_within_rounding(gpu_lib):
    sc = _scenes()[2]
    a, b = make_engine(sc), make_engine(sc)
    t = b.tuning()
    t.p2g_item = 4
    b.setTuning(t)
    a.substep(sc.dt, 5)
    b.substep(sc.dt, 5)
    a.sync()
    b.sync()
    sa, sb = a.getState(), b.getState()
    np.testing.assert_allclose(sb[0, :, :3], sa[0, :, :3], rtol=0, atol=1e-6)
    np.testing.assert_allclose(sb[1, :, :3], sa[1, :, :3], rtol=1e-4, atol=1e-6)


def test_tuning_rejects_invalid_shapes(gpu_lib):
    e = make_engine(_scenes()[0])
    for field, bad in (("g2p_block", 100), ("tile_blocks", 0), ("p2g_item", 2), ("p2g_item", 64)):
        t = e.tuning()
        setattr(t, field, bad)
        with pytest.raises(mpm.NebulaError):
            e.setTuning(t)


def test_autotune_cache_persists_choices(gpu_lib, tmp_path):
    """nb_mpm_autotune_cached: the first call times the shapes and appends the
    choice to the cache file; a second engine of the same workload finds it
    (nothing timed, same tuning); a line that does not parse is skipped; a
    workload of another shape (another transfer mode) is a miss."""
    scenes = _scenes()
    cache = str(tmp_path / "tuning.jsonl")
    a = make_batch_engine(scenes)
    before = a.getState()
    t1, hit1 = a.autotuneCached(scenes[0].dt, cache, trial_substeps=2)
    assert not hit1 and t1.ms_per_substep > 0.0
    assert _same(a.getState(), before)
    with open(cache) as f:
        lines = f.read().splitlines()
    assert len(lines) == 1 and '"key": "' in lines[0] and "|scenes" in lines[0]
    with open(cache, "a") as f:
        f.write("not json at all\n")
    b = make_batch_engine(scenes)
    t2, hit2 = b.autotuneCached(scenes[0].dt, cache, trial_substeps=2)
    assert hit2
    assert (t2.tile_blocks, t2.g2p_block, t2.p2g_item) == (t1.tile_blocks, t1.g2p_block, t1.p2g_item)
    bt = b.tuning()
    assert (bt.tile_blocks, bt.g2p_block, bt.p2g_item) == (t1.tile_blocks, t1.g2p_block, t1.p2g_item)
    c = make_engine(scenes[1])
    c.flip_ratio = 0.0  # PIC: another key
    _, hit3 = c.autotuneCached(scenes[1].dt, cache, trial_substeps=2)
    assert not hit3
    with open(cache) as f:
        assert len(f.read().splitlines()) == 3
