"""The C++ adapter include/sph_solver.hpp (QtSPHSolver method names over the
C ABI): it compiles and links against libsph_engine.so on the CPU, and on the
GPU a C++ program driving it like Source/Simulator.cpp:33-63 reproduces the
Python binding's results bit for bit."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "qt_adapter_demo.cpp")
THREADED_SRC = os.path.join(ROOT, "tests", "cpp", "qt_threaded_demo.cpp")


def build_demo(out_dir, src=SRC):
    from simplefluid_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        from simplefluid_amd import build
        build.build()
    libdir = os.path.dirname(engine.LIB_PATH)
    exe = os.path.join(str(out_dir), os.path.splitext(os.path.basename(src))[0])
    cmd = ["g++", "-std=c++17", "-O2", "-pthread", "-Wall", "-Wextra", "-Werror", "-I",
           os.path.join(ROOT, "include"), src, "-L", libdir, "-lsph_engine", f"-Wl,-rpath,{libdir}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_adapter_compiles_and_links(tmp_path):
    exe = build_demo(tmp_path)
    assert os.path.exists(exe)
    assert os.path.exists(build_demo(tmp_path, THREADED_SRC))


def test_adapter_header_is_self_contained(tmp_path):
    c = tmp_path / "t.cpp"
    c.write_text('#include "sph_solver.hpp"\nint main(){sph::SPHParameters p; p.setResolution(32);'
                 ' return p.particleRadius == p.kernelRadius / 4 ? 0 : 1;}\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(c)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_adapter_reports_engine_errors_as_exceptions(tmp_path):
    """Without a GPU, constructing the solver throws sph::Error carrying the
    engine's code and message (the reference has no error channel)."""
    from simplefluid_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        from simplefluid_amd import build
        build.build()
    libdir = os.path.dirname(engine.LIB_PATH)
    src = tmp_path / "err.cpp"
    src.write_text(
        '#include <cstdio>\n#include <memory>\n#include "sph_solver.hpp"\n'
        "int main() {\n"
        "  auto p = std::make_shared<sph::SPHParameters>();\n"
        "  try { sph::QtSPHSolver<> s(p); return 1; }\n"
        "  catch (const sph::Error& e) { std::printf(\"%d %s\\n\", e.code(), e.what());\n"
        "    return e.code() == SPH_E_HIP ? 0 : 2; }\n"
        "}\n")
    exe = str(tmp_path / "err")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", libdir, "-lsph_engine", f"-Wl,-rpath,{libdir}", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "HIP device" in r.stdout


@pytest.mark.gpu
def test_cpp_driver_matches_python_binding(tmp_path):
    from simplefluid_amd import engine as E
    exe = build_demo(tmp_path)
    out = tmp_path / "state.bin"
    scene, res, frames = E.SCENE_DOUBLE_DAMBREAK, 24, 3
    print(f"[cpp adapter] scene {scene} res {res} frames {frames}", flush=True)
    r = subprocess.run([exe, str(scene), str(res), str(frames), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    n, nf = np.frombuffer(raw[:8], np.uint32)
    ft = np.frombuffer(raw[8:8 + 4 * nf], np.float32)
    body = np.frombuffer(raw[8 + 4 * nf:], np.float32)
    x_cpp = body[:3 * n].reshape(n, 3)
    v_cpp = body[3 * n:].reshape(n, 3)
    assert n == 23958 and nf == frames

    p = E.default_params(res)
    s = E.SPHSolver(p)
    s.setupScene(scene)
    s.makeReady()
    for f in range(frames):
        t = np.float32(0.0)
        while t < np.float32(0.0333333333):
            t = np.float32(t + np.float32(s.advanceFrame()))
        assert t == ft[f]
    assert np.array_equal(s.getParticles(), x_cpp)
    assert np.array_equal(s.getVelocity(), v_cpp)
    lo, hi = np.float32(-1.0) + p.particle_radius, np.float32(1.0) - p.particle_radius
    assert (x_cpp >= lo).all() and (x_cpp <= hi).all()


@pytest.mark.gpu
def test_threaded_driver_snapshots_and_atomic_stop(tmp_path):
    """Worker thread steps frames and publishes snapshots; the GUI thread reads
    six of them and stops the worker through a std::atomic flag; the worker's
    final state equals a single-threaded replay of the same frame count."""
    from simplefluid_amd import engine as E
    exe = build_demo(tmp_path, THREADED_SRC)
    print("[threaded adapter] DoubleDambreak res 16, 6 frames", flush=True)
    r = subprocess.run([exe, str(E.SCENE_DOUBLE_DAMBREAK), "16", "6"], capture_output=True, text=True,
                       timeout=150)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "frames read by the GUI thread 6" in r.stdout
