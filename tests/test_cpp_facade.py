"""The C++ drop-in facade (include/nebula_mpm.hpp): compiles against the C ABI
with g++ (CPU) and runs the reference-style C++ tests on the GPU."""
import os
import subprocess

import pytest

from nebula_amd import mpm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "engine_cpp_test.cpp")
OUT = os.path.join(ROOT, "_build", "engine_cpp_test")


def _build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.dirname(mpm.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), SRC, "-o", OUT,
                    "-L", libdir, "-lnebula_mpm", f"-Wl,-rpath,{libdir}"], check=True)


def test_facade_compiles_and_links():
    if not os.path.exists(mpm.LIB_PATH):
        pytest.skip("HIP library not built")
    _build()
    assert os.path.exists(OUT)


@pytest.mark.gpu
def test_facade_runs_reference_style_tests(gpu_lib):
    _build()
    r = subprocess.run([OUT], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


def test_facade_emitters_run_reference_emitter_tests():
    """The reference's emitter tests (tests/test_mpm.cpp:175-226) through the
    facade's Emitter::sample, which draws exactly what the engine draws; a
    header-only program, no GPU."""
    src = os.path.join(ROOT, "tests", "cpp", "emitter_cpp_test.cpp")
    out = os.path.join(ROOT, "_build", "emitter_cpp_test")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", out],
                   check=True)
    r = subprocess.run([out], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


def test_facade_host_code_is_clean_under_sanitizers():
    """The facade's host-side code (emitter sampling, record packing) built
    with AddressSanitizer and UndefinedBehaviorSanitizer: the emitter test
    program runs without a report.  Host code only, no GPU."""
    src = os.path.join(ROOT, "tests", "cpp", "emitter_cpp_test.cpp")
    out = os.path.join(ROOT, "_build", "emitter_cpp_test_san")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", os.path.join(ROOT, "include"), src, "-o", out], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([out], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
