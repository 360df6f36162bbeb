"""The CPU restatement built with AddressSanitizer and UndefinedBehaviorSanitizer
(host code only) runs a few float and double steps with OpenMP threads, a
wall hit and the CFL step, without a report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_scene_generators_under_asan_ubsan(tmp_path):
    """The engine's host-side scene generators (reference lattices and the
    seeded ball sampler with its worker threads) built with g++ and ASan +
    UBSan run every case of tests/cpp/scene_sanitize.cpp without a report."""
    exe = str(tmp_path / "scene_sanitize")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "simplefluid_amd", "csrc", "sph_scene.cpp"),
           os.path.join(ROOT, "tests", "cpp", "scene_sanitize.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout + r.stderr)[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
