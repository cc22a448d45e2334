"""The HIP path against the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py from the oracle), through the C ABI:
bins, mesh SDFs and emitted particles bit for bit; substeps within the
tolerance the FP32 arithmetic allows, checked against both the FP32 oracle
and its FP64 twin."""
import ctypes as C
import os

import numpy as np
import pytest

from nebula_amd import mpm
from nebula_amd.scenes import Scene, make_engine

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def _engine(res, fields, materials=None, gravity=None, dt=1e-4, flip_ratio=0.0, corrected_svd=False):
    """A plain engine on a res^3 unit-cube grid holding the fixture's particles."""
    cfg = mpm.MPMConfig(nx=0, ny=0, nz=0, domainSize=(1.0, 1.0, 1.0), cellSize=1.0 / res,
                        maxParticles=fields.shape[1], E=1e4, nu=0.2, rho0=1.0)
    if materials is not None:
        raw = materials.tobytes()
        mats = [mpm.MaterialParams.from_buffer_copy(raw[i:i + 64]) for i in range(0, len(raw), 64)]
    else:
        mats = [mpm.presetJelly()]
    forces = [mpm.GravityForce(tuple(float(v) for v in gravity[0]), float(gravity[1]))] if gravity else []
    sc = Scene("golden", cfg, mats, fields, dt, flip_ratio, 1, forces=forces, corrected_svd=corrected_svd)
    return make_engine(sc)


@pytest.mark.parametrize("name", ["substep_elastic_apic", "substep_dp_pic"])
def test_substeps_match_golden(gpu_lib, name):
    d = load(name)
    e = _engine(int(d["res"]), d["fields"], d["materials"], (d["gravity_dir"], d["gravity"]), float(d["dt"]),
                float(d["flip_ratio"]), bool(d["corrected_svd"]))
    e.substep(float(d["dt"]), int(d["substeps"]))
    e.sync()
    g = e.getState().astype(np.float64)
    s32, s64 = d["state32"].astype(np.float64), d["state64"]
    h = 1.0 / int(d["res"])
    ex32 = np.abs(g[0, :, :3] - s32[0, :, :3]).max()
    ex64, o64 = np.abs(g[0, :, :3] - s64[0, :, :3]).max(), np.abs(s32[0, :, :3] - s64[0, :, :3]).max()
    vmax = np.abs(s64[1, :, :3]).max()
    ev32 = np.abs(g[1, :, :3] - s32[1, :, :3]).max()
    print(f"  {name}: x vs FP32 {ex32 / h:.2e} dx, vs FP64 {ex64 / h:.2e} dx (FP32 oracle {o64 / h:.2e} dx); "
          f"v vs FP32 {ev32:.2e} of {vmax:.2e}", flush=True)
    assert ex32 <= 1e-4 * h
    assert ex64 <= 3.0 * o64 + 1e-6 * h
    assert ev32 <= 1e-3 * vmax
    assert np.array_equal(g[1, :, 3].astype(np.float32).view(np.uint32), d["fields"][1, :, 3].view(np.uint32))


def test_bins_match_golden(gpu_lib):
    d = load("bins_edge")
    e = _engine(int(d["res"]), d["fields"])
    e.runBin()
    keys, srt, cnt, end = e.getBins()
    for k, v in (("keys", keys), ("sorted", srt), ("cell_count", cnt), ("cell_end", end)):
        assert np.array_equal(v, d[k]), k


@pytest.mark.parametrize("exact", [False, True], ids=["reference", "exact"])
def test_mesh_sdf_matches_golden(gpu_lib, exact):
    d = load("mesh_box_sdf")
    res = int(d["res"])
    f = np.zeros((8, 1, 4), np.float32)
    f[0, 0, :3] = 0.5
    e = _engine(res, f)
    e.setColliderMesh(d["tris"], exact=exact)
    got = e.getColliderSDF()
    want = d["sdf_exact"] if exact else d["sdf_reference"]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_emitted_particles_match_golden(gpu_lib):
    d = load("emit")
    cap = int(d["capacity"])
    cfg = mpm.MPMConfig(nx=8, ny=8, nz=8, domainSize=(1.0, 1.0, 1.0), cellSize=1.0 / 32, maxParticles=cap,
                        E=2e3, nu=0.3, rho0=1e3)
    e = mpm.MPMEngine(cfg)
    e.setMaterials([mpm.presetJelly(), mpm.presetJelly()])
    e.resetParticles()
    raw = d["records"].tobytes()
    size = C.sizeof(mpm.EmitterGPU)
    for i in range(int(d["n_emitters"])):
        rec = mpm.EmitterGPU.from_buffer_copy(raw[i * size:(i + 1) * size])
        assert mpm.lib().nb_mpm_add_emitter(e._h, C.byref(rec)) == 0
    for _ in range(int(d["calls"])):
        assert mpm.lib().nb_mpm_emit(e._h, C.c_float(float(d["dt"]))) == 0
    st = e.getState()
    n = d["pos"].shape[0]
    assert e.liveParticleCount() == n
    assert np.array_equal(st[0, :n].view(np.uint32), d["pos"].view(np.uint32))
    assert np.array_equal(st[1, :n].view(np.uint32), d["vel"].view(np.uint32))
