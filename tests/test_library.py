"""CPU-side checks of the C ABI library: it builds, loads, exports every
symbol include/sph_engine.h declares, and its host-only entry points (scene
generation, parameter defaults) agree with the oracle.  No GPU call here."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "sph_engine.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from simplefluid_amd import engine
    so = engine.LIB_PATH
    if not os.path.exists(so):
        from simplefluid_amd import build
        build.build()
    return engine.load_library()


def test_exports_every_declared_symbol(lib):
    from simplefluid_amd import engine
    names = header_functions()
    assert len(names) >= 38
    assert sorted(engine.EXPORTS) == names
    for n in names:
        assert hasattr(lib, n), n


def test_header_is_c_compatible(tmp_path):
    """The ABI header compiles as C (no C++ types in the signatures)."""
    import subprocess
    c = tmp_path / "t.c"
    c.write_text('#include "sph_engine.h"\nint main(void){sph_params p; return sph_params_default(&p);}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(c)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_defaults_match_reference_panels(lib):
    from simplefluid_amd import engine
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "scene_counts.json")))["defaults"]
    p = engine.default_params()
    assert p.kernel_radius == np.float32(2.0) / np.float32(24)
    assert p.particle_radius == p.kernel_radius / np.float32(4)
    assert p.pressure_stiffness == pytest.approx(g["pressure_stiffness"])
    assert p.viscosity == pytest.approx(g["viscosity"])
    assert p.boundary_restitution == pytest.approx(g["boundary_restitution"])
    assert p.use_attractive_pressure == g["attractive_pressure"]
    assert p.stop_time == pytest.approx(g["stop_time"])


@pytest.mark.parametrize("res", [16, 24, 40])
def test_scene_generator_bit_exact_vs_oracle(lib, res):
    from oracle import scenes
    from simplefluid_amd import engine
    for sid, name in enumerate(engine.SCENE_NAMES):
        ours = engine.scene_generate(sid, res)
        ref = scenes.reference_scene(name, res)
        assert ours.shape == ref.shape, name
        assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32)), name


def test_scene_counts_fixture(lib):
    from simplefluid_amd import engine
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "scene_counts.json")))["res24"]
    for sid, name in enumerate(engine.SCENE_NAMES):
        assert engine.scene_generate(sid, 24).shape[0] == g[name]


def test_scene_generate_rejects_bad_input(lib):
    n = C.c_uint32(0)
    assert lib.sph_scene_generate(7, 24, None, 0, C.byref(n)) == -1
    assert lib.sph_scene_generate(1, 0, None, 0, C.byref(n)) == -1
    assert lib.sph_scene_generate(1, 8, None, 0, C.byref(n)) == 0 and n.value == 8 ** 3
    buf = (C.c_float * 3)()
    assert lib.sph_scene_generate(1, 8, buf, 1, C.byref(n)) == -6


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from simplefluid_amd import engine
    monkeypatch.setattr(engine, "_LIB", None)
    with pytest.raises(RuntimeError, match="missing"):
        engine.load_library(str(tmp_path / "nope.so"))


def test_synthetic_configs_shapes():
    from simplefluid_amd import configs
    c0 = configs.dambreak_65k()
    assert c0["pos"].shape == (65536, 3)
    c1 = configs.double_dambreak_1m()
    assert c1["pos"].shape == (1048576, 3)
    assert (c1["pos"] >= 0).all() and (c1["pos"][:, 1] <= 1).all() and (c1["pos"][:, [0, 2]] <= 2).all()
    e = configs.ensemble(nscenes=32, seed=3)
    ns = [s["pos"].shape[0] for s in e["scenes"]]
    assert min(ns) >= 3000 and max(ns) <= 70000
    for s in e["scenes"]:
        assert (s["pos"] < 1.0).all()


def test_slab_share_partitions_and_holds_every_second_neighbour():
    """The slabs of an 8-way x sharding own disjoint eighths of the scene, and
    each share holds every particle within 2h of one it owns (the neighbours of
    its neighbours, whose densities the share recomputes)."""
    from scipy.spatial import cKDTree
    from simplefluid_amd import configs
    c = configs.dambreak_65k(jitter=0.002, seed=5)
    pos = c["pos"]
    h = float(c["params"].kernel_radius)
    tree = cKDTree(pos.astype(np.float64))
    seen = np.zeros(len(pos), bool)
    for s in range(8):
        sh = configs.slab_share(c, 8, s)
        assert sh["owned"] == len(pos) // 8
        own, sel = sh["pos"][:sh["owned"]], sh["pos"]
        idx = {tuple(p) for p in sel.tolist()}
        assert len(idx) == len(sel)
        x0, x1 = sh["x_range"]
        assert ((own[:, 0] >= x0) & (own[:, 0] < x1)).all()
        oi = np.flatnonzero((pos[:, 0] >= x0) & (pos[:, 0] < (x1 if s < 7 else np.inf)))
        assert not seen[oi].any()
        seen[oi] = True
        need = set()
        for lst in tree.query_ball_point(own.astype(np.float64), 2 * h):
            need.update(lst)
        assert {tuple(p) for p in pos[sorted(need)].tolist()} <= idx
        # Nothing farther than 2h from the slab is carried.
        g = sh["pos"][sh["owned"]:, 0]
        assert ((g >= x0 - 2 * h) & (g < x1 + 2 * h)).all()
        lo, hi = sh["params"].box_min[0], sh["params"].box_max[0]
        assert (sel[:, 0] > lo).all() and (sel[:, 0] < hi).all()
    assert seen.all()
    worst = configs.slab_share(c, 8)
    assert worst["owned"] + worst["ghosts"] == max(
        configs.slab_share(c, 8, s)["pos"].shape[0] for s in range(8))


def test_null_handle_is_rejected(lib):
    """Entry points given no handle return SPH_E_INVALID without touching a GPU."""
    from simplefluid_amd import engine
    v = C.c_uint64(7)
    assert lib.sph_set_device_memory_limit(None, 1 << 20) == engine.SPH_E_INVALID
    assert lib.sph_device_memory_used(None, C.byref(v)) == engine.SPH_E_INVALID and v.value == 7
    assert lib.sph_advance_steps(None, 1) == engine.SPH_E_INVALID
    assert lib.sph_make_ready(None) == engine.SPH_E_INVALID
    assert lib.sph_num_particles(None) == 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_no_gpu_fails_loudly(lib):
    """Without a GPU the engine refuses to run (no CPU fallback) and says why."""
    from simplefluid_amd import engine
    with pytest.raises(engine.SPHError) as ei:
        engine.Engine(engine.make_params())
    assert ei.value.code == engine.SPH_E_HIP
    assert "HIP device" in str(ei.value)


def test_golden_frame_fixture_loads_and_matches_fp32_oracle(lib, oracle_lib):
    """The committed frame (export format) loads with the binding's reader and
    an FP32 oracle rollout of the same scene lands within FP32 round-off of it."""
    import oracle
    from simplefluid_amd import engine
    x, v = engine.load_frame(os.path.join(ROOT, "tests", "golden", "frame_cubedrop16_step40.npy"))
    assert x.shape == (4096, 3) and v.shape == (4096, 3)
    p = engine.default_params(16)
    o = oracle.Oracle(oracle.from_params(p, 4), fp64=False, threads=8)
    o.set_state(engine.scene_generate(engine.SCENE_CUBE_DROP, 16).astype(np.float64))
    o.step(40)
    xo, vo = o.state_by_id()
    assert np.abs(xo - x).max() < 2e-6
    assert np.abs(vo - v).max() < 1e-4


def test_ball_sample_spacing_uniformity_and_seed(lib):
    """sph_scene_sample_ball: every point inside the ball, no pair of the
    float32 output closer than the minimum spacing (all pairs checked with a
    k-d tree in float64), the same points
    for the same seed and others for another, and uniform: the fraction
    within half the radius is 1/8 and the octants hold equal shares, within
    a few standard deviations of a multinomial."""
    from scipy.spatial import cKDTree
    from simplefluid_amd import engine
    c, R, s, n = (0.3, -0.2, 0.1), 0.1, 0.0034, 60000
    x = engine.sample_ball(7, c, R, s, n)
    assert x.shape == (n, 3) and x.dtype == np.float32
    d = x.astype(np.float64) - np.asarray(c, np.float32)
    rad = np.sqrt((d * d).sum(1))
    assert rad.max() < R * (1 + 1e-6)
    # The sampler checks the spacing on the float32 coordinates it returns,
    # against the float32 spacing.
    s32 = float(np.float32(s))
    pairs = cKDTree(x.astype(np.float64)).query_pairs(s32 * (1 - 1e-12))
    assert len(pairs) == 0
    # The packing is dense: most darts that still fit were placed.
    phi = n * (np.pi / 6) * s ** 3 / (4 / 3 * np.pi * R ** 3)
    assert phi > 0.25
    assert np.array_equal(x, engine.sample_ball(7, c, R, s, n))
    assert not np.array_equal(x, engine.sample_ball(8, c, R, s, n))
    inner = (rad < R / 2).mean()
    # Points are excluded from a shell s/2 thick at the surface less than
    # inside; allow that bias plus 4 sigma.
    assert abs(inner - 0.125) < 4 * np.sqrt(0.125 * 0.875 / n) + 0.01
    octant = (d[:, 0] > 0).astype(int) * 4 + (d[:, 1] > 0) * 2 + (d[:, 2] > 0)
    share = np.bincount(octant, minlength=8) / n
    assert np.abs(share - 0.125).max() < 4 * np.sqrt(0.125 * 0.875 / n)


def test_ball_sample_too_full_is_refused(lib):
    """More points than the ball can hold at the spacing: SPH_E_CAPACITY."""
    from simplefluid_amd import engine
    with pytest.raises(engine.SPHError) as ei:
        engine.sample_ball(1, (0, 0, 0), 0.05, 0.01, 200000)
    assert ei.value.code == engine.SPH_E_CAPACITY
