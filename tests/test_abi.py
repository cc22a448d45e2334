"""The C ABI boundary (CPU-only): the built library exports every entry point
include/nebula_mpm.h declares, the POD records keep the reference's byte
layouts, the host-only defaults equal the reference's, and without a device
creation fails with a status and a message instead of falling back, and so
does any engine call when the library itself is missing."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

from nebula_amd import mpm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nebula_mpm.h")
C_HEADERS = sorted(os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
                   if f.endswith(".h"))


def _declared():
    names = set()
    for hdr in C_HEADERS:
        src = re.sub(r"/\*.*?\*/", "", open(hdr).read(), flags=re.S)
        names |= set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(mpm.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "nebula_amd", "csrc")], check=True)
    return mpm.lib()


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", mpm.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for n in names:
        getattr(lib, n)
    assert set(mpm.EXPORTED_SYMBOLS) <= set(names)


def test_pod_layouts_match_reference():
    # src/core/MaterialParams.h:115, src/core/Collider.h:30, src/core/Force.h:32
    assert C.sizeof(mpm.MaterialParams) == 64
    assert C.sizeof(mpm.ColliderPrimitive) == 64
    assert C.sizeof(mpm.ForceGPU) == 32
    assert mpm.MaterialParams.model.offset == 12 and mpm.MaterialParams.q_max.offset == 24
    assert mpm.MaterialParams.bulkK.offset == 28 and mpm.MaterialParams.viscosity.offset == 36
    assert mpm.ColliderPrimitive.radius.offset == 28 and mpm.ColliderPrimitive.friction.offset == 48
    assert mpm.ForceGPU.data.offset == 4


def test_host_defaults_equal_reference(lib):
    c = mpm.MPMConfig()
    lib.nb_mpm_config_default(C.byref(c))
    # src/engine/MPMEngine.h:25-39
    assert (c.nx, c.ny, c.nz) == (32, 32, 32)
    assert tuple(c.domainSize) == (10.0, 10.0, 10.0)
    assert c.cellSize == pytest.approx(10.0 / 64.0) and c.maxParticles == 0
    assert (c.E, c.nu, c.rho0) == (pytest.approx(1e4), pytest.approx(0.3), pytest.approx(1000.0))
    k = mpm._Knobs()
    lib.nb_mpm_knobs_default(C.byref(k))
    # src/engine/MPMEngine.h:72-80
    assert k.num_substeps == 20 and k.flip_ratio == 0.0 and k.plastic_model == 0
    assert k.q_max == pytest.approx(1e5) and k.M_friction == pytest.approx(0.577) and k.q_cohesion == 0.0
    assert k.restitution == pytest.approx(0.3) and k.wall_friction == 0.0
    assert lib.nb_mpm_abi_version() == 1


def test_presets_equal_reference_closed_forms():
    # src/core/MaterialParams.h:119-180
    j = mpm.presetJelly()
    assert j.mu == pytest.approx(1e4 / (2 * 1.4)) and j.lambda_ == pytest.approx(1e4 * 0.4 / (1.4 * 0.2))
    s = mpm.presetSand()
    assert s.model == mpm.MaterialModel.DRUCKER_PRAGER and s.M_friction == pytest.approx(0.577)
    w = mpm.presetWater()
    assert w.model == mpm.MaterialModel.FLUID and w.bulkK == 5e3 and w.fluidGamma == 7.0
    m = mpm.presetMud()
    assert m.model == mpm.MaterialModel.VISCOPLASTIC_MUD and m.q_cohesion == 5e2 and m.viscosity == pytest.approx(0.1)
    p = mpm.presetPowder()
    assert p.model == mpm.MaterialModel.GRANULAR_POWDER and p.M_friction == pytest.approx(0.8)
    cs = mpm.ColliderSet()
    cs.addPlane((0, 1, 0), (0, 2, 0))
    assert cs.data()[0].ny == pytest.approx(1.0)  # normalised as ColliderSet::addPlane does


def test_create_without_device_fails_loudly(lib):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    c = mpm.MPMConfig()
    h = C.c_void_p()
    rc = lib.nb_mpm_create(C.byref(c), None, C.byref(h))
    assert rc == -3 and h.value is None
    assert b"device" in lib.nb_last_error()
    with pytest.raises(mpm.NebulaError):
        mpm.MPMEngine(c)


def test_missing_library_fails_loudly():
    """Without the HIP library every engine call raises; there is no CPU path."""
    code = ("import nebula_amd.mpm as m\n"
            "m.LIB_PATH = '/nonexistent/libnebula_mpm.so'\n"
            "try:\n"
            "    m.MPMEngine(m.MPMConfig())\n"
            "except m.NebulaError as e:\n"
            "    print('raised:', e)\n"
            "else:\n"
            "    print('no error')\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("raised: HIP library missing"), out.stdout


def _c_layout(struct, fields):
    """sizeof and offsetof of a header struct, from the C compiler itself."""
    import tempfile
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "nebula_mpm.h"', 'int main(void) {',
           f'  printf("%zu", sizeof({struct}));']
    src += [f'  printf(" %zu", offsetof({struct}, {f}));' for f in fields]
    src += ['  printf("\\n");', '  return 0;', '}']
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        with open(c, "w") as fh:
            fh.write("\n".join(src))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    return [int(v) for v in out]


@pytest.mark.parametrize("struct,cls", [("nb_emitter", "EmitterGPU"), ("nb_mpm_scene", "BatchScene"),
                                        ("nb_mpm_tuning", "Tuning")])
def test_extension_records_match_the_header(struct, cls):
    py = getattr(mpm, cls)
    names = [f[0] for f in py._fields_]
    got = _c_layout(struct, names)
    assert got[0] == C.sizeof(py), f"{struct}: sizeof"
    assert got[1:] == [getattr(py, n).offset for n in names], f"{struct}: offsets"
