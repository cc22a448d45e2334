"""The engine on the caller's stream (nb_mpm_set_stream): the counterpart of
MPMEngine::step recording into the caller's command buffer
(src/engine/MPMEngine.cpp:404).  Work moved to another stream gives the
same results as on the engine's own stream, before and after switching back.

The caller's stream is made with hipStreamCreate from the HIP runtime the
engine library already loaded (the loader returns the mapped copy for its
soname), so no second runtime enters the process."""
import ctypes as C

import numpy as np
import pytest

from nebula_amd import mpm
from nebula_amd.mpm import MaterialModel
from nebula_amd.scenes import falling_box, make_engine

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def test_engine_follows_the_callers_stream(gpu_lib):
    assert len(mpm.hip_runtimes_mapped()) == 1
    hip = _hip()
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0  # default priority
    try:
        sc = falling_box("s", 24, 2000, 5e4, 0.3, MaterialModel.VON_MISES, np.random.default_rng(3))
        a, b = make_engine(sc), make_engine(sc)
        b.setStream(s.value)
        a.substep(sc.dt, 10)
        b.substep(sc.dt, 10)
        assert hip.hipStreamSynchronize(s) == 0
        a.sync()
        assert np.array_equal(a.getState().view(np.uint32), b.getState().view(np.uint32))
        b.setStream(None)
        a.substep(sc.dt, 5)
        b.substep(sc.dt, 5)
        a.sync()
        b.sync()
        assert np.array_equal(a.getState().view(np.uint32), b.getState().view(np.uint32))
        del a, b
    finally:
        assert hip.hipStreamDestroy(s) == 0
