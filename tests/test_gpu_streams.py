"""The engine on the caller's stream (nb_mpm_set_stream): the counterpart of
MPMEngine::step recording into the caller's command buffer
(src/engine/MPMEngine.cpp:404).  Work moved to another stream gives the
same results as on the engine's own stream, before and after switching back."""
import numpy as np
import pytest
import torch  # before the engine library loads: torch brings its own HIP runtime

from nebula_amd.mpm import MaterialModel
from nebula_amd.scenes import falling_box, make_engine

pytestmark = pytest.mark.gpu


def test_engine_follows_the_callers_stream(gpu_lib):
    sc = falling_box("s", 24, 2000, 5e4, 0.3, MaterialModel.VON_MISES, np.random.default_rng(3))
    a, b = make_engine(sc), make_engine(sc)
    s = torch.cuda.Stream()  # default priority
    b.setStream(s.cuda_stream)
    a.substep(sc.dt, 10)
    b.substep(sc.dt, 10)
    s.synchronize()
    a.sync()
    assert np.array_equal(a.getState().view(np.uint32), b.getState().view(np.uint32))
    b.setStream(None)
    a.substep(sc.dt, 5)
    b.substep(sc.dt, 5)
    a.sync()
    b.sync()
    assert np.array_equal(a.getState().view(np.uint32), b.getState().view(np.uint32))
