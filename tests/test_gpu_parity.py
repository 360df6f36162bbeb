"""GPU parity of the HIP engine against the CPU oracle, through the C ABI.

Bar: cell keys, the sort permutation, sorted keys, cell start/end tables
and per-particle neighbour sets bit-exact; density within 2e-6 relative;
acceleration within 1e-4 of the per-particle force scale (the oracle uses
1/sqrt, the engine the hardware rsqrt); positions within 1e-6 absolute after
a step from the identical input state.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RHO_RTOL = 2e-6
ACC_TOL = 1e-4
POS_ATOL = 1e-6


def say(*a):
    print(*a, flush=True)


def to_oracle(p, xsub=1):
    import oracle
    return oracle.from_params(p, xsub)


@pytest.fixture(scope="module")
def E():
    from simplefluid_amd import engine
    engine.load_library()
    return engine


def make_engine(E, params, pos, vel=None):
    e = E.Engine(params)
    e.set_particles(pos, vel)
    e.make_ready()
    return e


def step_and_compare(E, e, params, label, threads=16, vs_fp64=False):
    """State of e -> oracle; one step on both; compare every intermediate.

    vs_fp64: for states whose net accelerations are small differences of
    large pair terms, judge the acceleration against an FP64 step from the
    same state, allowing the GPU 4x the FP32 oracle's own largest error, and
    bound velocity and position by dt and dt^2 times that."""
    import oracle
    xs, vs, ids = e.debug_get_state()
    o = oracle.Oracle(to_oracle(params, e.debug_constants()["xsub"]), threads=threads)
    o.record_neighbors(True)
    o.set_state(xs.astype(np.float64), vs.astype(np.float64), ids)
    e.debug_keep_acc(True)
    e.advance_steps(1)
    g = e.debug_step_outputs(acc=True)
    o.step()
    d = o.step_debug()
    say(f"[{label}] n={len(ids)} compare keys/perm")
    assert np.array_equal(g["key"], d["key"]), "cell keys differ"
    assert np.array_equal(g["perm"], d["perm"]), "sort permutation differs"
    assert np.array_equal(g["sorted_key"], d["sorted_key"])
    gs, ge = e.debug_cells()
    os_, oe = o.cells()
    assert np.array_equal(gs, os_) and np.array_equal(ge, oe), "cell table differs"
    goff, glst = e.debug_neighbors()
    ooff, olst = o.neighbors()
    assert np.array_equal(goff, ooff), "neighbour counts differ"
    assert np.array_equal(glst, olst), "neighbour sets differ"
    rho_o = d["rho"]
    assert np.all(np.abs(g["rho"] - rho_o) <= RHO_RTOL * np.abs(rho_o)), \
        f"density max rel err {np.max(np.abs(g['rho'] - rho_o) / np.abs(rho_o))}"
    acc_o = d["acc"]
    scale = np.maximum(np.abs(acc_o).max(1, keepdims=True), 10.0)
    err = np.abs(g["acc"] - acc_o) / scale
    if vs_fp64:
        o64 = oracle.Oracle(to_oracle(params, e.debug_constants()["xsub"]), fp64=True, threads=threads)
        o64.set_state(xs.astype(np.float64), vs.astype(np.float64), ids)
        o64.step()
        d64 = o64.step_debug()
        assert np.array_equal(d64["perm"], d["perm"])
        e_gpu = np.abs(g["acc"] - d64["acc"]).max()
        e_f32 = np.abs(acc_o - d64["acc"]).max()
        say(f"[{label}] |a| max {np.abs(d64['acc']).max():.3e}; max err vs FP64: GPU {e_gpu:.3e}, "
            f"FP32 oracle {e_f32:.3e}")
        assert e_gpu <= 4.0 * e_f32 + 1e-6 * np.abs(d64["acc"]).max(), "acc error above 4x FP32's"
    else:
        assert err.max() <= ACC_TOL, f"acc max scaled err {err.max()}"
    xg, vg, ig = e.debug_get_state()
    xo, vo, io = o.state()
    assert np.array_equal(ig, io), "particle ids differ"
    if vs_fp64:
        dt = float(params.dt)
        dv = dt * 2.0 * (4.0 * e_f32 + 1e-6 * np.abs(d64["acc"]).max()) + 1e-6
        assert np.abs(vg - vo).max() <= dv, f"vel err {np.abs(vg - vo).max():.3e} above {dv:.3e}"
        assert np.abs(xg - xo).max() <= dt * dv + POS_ATOL, "pos err above dt times the velocity bound"
    else:
        assert np.abs(xg - xo).max() <= POS_ATOL, f"pos err {np.abs(xg - xo).max()}"
        assert np.abs(vg - vo).max() <= 1e-2 * ACC_TOL * max(1.0, np.abs(vo).max()) + 1e-6
    e.debug_keep_acc(False)
    say(f"[{label}] ok: rho rel {np.max(np.abs(g['rho'] - rho_o) / rho_o):.2e}, acc {err.max():.2e}, "
        f"nbrs {len(glst)}")
    return d


def test_constants_match_oracle(E):
    import oracle
    from simplefluid_amd import configs
    for cfg in (configs.dambreak_65k(), configs.double_dambreak_1m()):
        e = E.Engine(cfg["params"])
        e.set_particles(cfg["pos"])
        c = e.debug_constants()
        assert c["xsub"] == 4
        oc = oracle.Oracle(to_oracle(cfg["params"], c["xsub"])).constants()
        for k in ("h", "h2", "inv_cs", "mass", "cD", "cP", "cV", "inv_rho0"):
            assert np.float32(c[k]) == np.float32(oc[k]), k
        assert np.array_equal(c["G"], oc["G"])
        e.close()


def test_dambreak_65k_first_step_parity(E):
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k()
    e = make_engine(E, cfg["params"], cfg["pos"])
    step_and_compare(E, e, cfg["params"], "dambreak_65k step 1")


def test_dambreak_65k_parity_along_rollout(E):
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=11)
    e = make_engine(E, cfg["params"], cfg["pos"])
    for k in (1, 60, 300):
        e.advance_steps(k)
        step_and_compare(E, e, cfg["params"], f"dambreak_65k after {k}")


def test_dambreak_65k_drift_vs_fp64(E):
    """The full 65k dam-break rollout (1000 steps of 1e-4 s): drift of the
    GPU FP32 trajectory from the FP64 oracle, stated beside the FP32 oracle's
    own drift from FP64 over the same rollout (distances in units of h)."""
    import oracle
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k()
    steps = cfg["steps"]
    e = make_engine(E, cfg["params"], cfg["pos"])
    xsub = e.debug_constants()["xsub"]
    e.advance_steps(steps)
    xg = e.get_particles().astype(np.float64)
    e.close()
    ref = {}
    for fp64 in (True, False):
        say(f"[drift] oracle fp64={fp64} {steps} steps")
        o = oracle.Oracle(to_oracle(cfg["params"], xsub), fp64=fp64, threads=16)
        o.set_state(cfg["pos"].astype(np.float64))
        o.step(steps)
        ref[fp64] = o.state_by_id()[0]
    h = float(cfg["params"].kernel_radius)

    def drift(a, b):
        d = np.linalg.norm(a - b, axis=1) / h
        return d.max(), float(np.sqrt((d * d).mean()))

    gmax, grms = drift(xg, ref[True])
    omax, orms = drift(ref[False], ref[True])
    say(f"[drift] after {steps} steps vs FP64 oracle (units of h): GPU max {gmax:.3e} rms {grms:.3e}; "
        f"FP32 oracle max {omax:.3e} rms {orms:.3e}")
    assert np.isfinite(xg).all()
    assert grms <= max(3.0 * orms, 0.05), "GPU drifts from FP64 well beyond FP32 round-off growth"


def test_double_dambreak_1m_long_run_stays_physical(E):
    """2000 steps (0.2 s) of the full 1M workload: every particle finite and
    inside the clamp bounds, particle ids a permutation, and the fluid's
    centre of mass falling with its kinetic energy growing (a dam break)."""
    from simplefluid_amd import configs
    cfg = configs.double_dambreak_1m()
    p = cfg["params"]
    e = make_engine(E, p, cfg["pos"])
    y0 = float(e.get_particles()[:, 1].mean())
    e.advance_steps(2000)
    x, v = e.get_particles(), e.get_velocity()
    _, _, ids = e.debug_get_state()
    assert np.isfinite(x).all() and np.isfinite(v).all()
    lo = np.array(p.box_min[:], np.float32) + np.float32(p.particle_radius)
    hi = np.array(p.box_max[:], np.float32) - np.float32(p.particle_radius)
    assert (x >= lo).all() and (x <= hi).all()
    assert np.array_equal(np.sort(ids), np.arange(len(ids), dtype=np.uint32))
    y1 = float(x[:, 1].mean())
    ke = 0.5 * float((v.astype(np.float64) ** 2).sum(1).mean())
    say(f"[long run] mean height {y0:.4f} -> {y1:.4f}, mean kinetic energy per mass {ke:.4e}")
    assert y1 < y0 and ke > 0.01
    e.close()


def test_sphere_drop_4m_full_run_stays_physical(E):
    """The 4.4M-particle sphere drop over its full 300 steps (the largest
    plain workload, x subdivision 4, autotuned): every particle finite and
    inside the clamp bounds, ids a permutation, the sphere's particles
    falling towards the resting layer."""
    from simplefluid_amd import configs
    cfg = configs.sphere_drop_4m()
    p = cfg["params"]
    e = make_engine(E, p, cfg["pos"])
    e.autotune()
    ns = cfg["n_sphere"]
    ys0 = float(e.get_particles()[:ns, 1].mean())
    e.advance_steps(cfg["steps"])
    x, v = e.get_particles(), e.get_velocity()
    _, _, ids = e.debug_get_state()
    assert np.isfinite(x).all() and np.isfinite(v).all()
    lo = np.array(p.box_min[:], np.float32) + np.float32(p.particle_radius)
    hi = np.array(p.box_max[:], np.float32) - np.float32(p.particle_radius)
    assert (x >= lo).all() and (x <= hi).all()
    assert np.array_equal(np.sort(ids), np.arange(len(ids), dtype=np.uint32))
    ys1 = float(x[:ns, 1].mean())
    say(f"[sphere 4m] {len(x)} particles, sphere mean height {ys0:.4f} -> {ys1:.4f} after {cfg['steps']} steps")
    assert ys1 < ys0
    e.close()


def test_ensemble_256_full_run_stays_physical(E):
    """The 256-scene ensemble over its full 1000 steps in one handle (5.7M
    particles): every scene's particles finite, inside that scene's clamp
    bounds and falling, ids a permutation of the scene's own range (no
    particle crosses into another scene)."""
    from simplefluid_amd import configs
    ens = configs.ensemble(nscenes=256, seed=3)
    params = [s["params"] for s in ens["scenes"]]
    counts = [s["pos"].shape[0] for s in ens["scenes"]]
    b = E.BatchEngine(params, counts)
    b.set_particles(np.concatenate([s["pos"] for s in ens["scenes"]]))
    b.make_ready()
    say(f"[ensemble run] {sum(counts)} particles, {ens['steps']} steps")
    b.advance_steps(ens["steps"])
    x, v = b.get_particles(), b.get_velocity()
    _, _, ids = b.debug_get_state()
    off = b.scene_offsets()
    assert np.isfinite(x).all() and np.isfinite(v).all()
    for i, s in enumerate(ens["scenes"]):
        sl = slice(off[i], off[i + 1])
        p = s["params"]
        lo = np.array(p.box_min[:], np.float32) + np.float32(p.particle_radius)
        hi = np.array(p.box_max[:], np.float32) - np.float32(p.particle_radius)
        assert (x[sl] >= lo).all() and (x[sl] <= hi).all(), i
        assert np.array_equal(np.sort(ids[sl]), np.arange(off[i], off[i + 1], dtype=np.uint32)), i
        assert x[sl, 1].mean() < s["pos"][:, 1].mean(), i
    b.close()


def compressed_block_scene():
    """A 2r lattice block beside a block compressed to spacing 1.25r (about
    4x rest density) in the same cell rows: up to 147 neighbours and rows of
    more than 32 candidates, which the neighbour masks do not cover, next to
    ordinary rows in the same waves."""
    from simplefluid_amd import configs
    R = configs.R
    a = configs.lattice_block((16, 24, 16), (0.1, 0.0, 0.1))
    s = 1.25 * R
    i, j, k = np.meshgrid(np.arange(20), np.arange(24), np.arange(20), indexing="ij")
    b = np.array([0.1 + 33 * R, R, 0.1 + R]) + s * np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    pos = np.concatenate([a, b.astype(np.float32)])
    return configs.make_params(pressure_stiffness=50.0, viscosity=0.01, dt=1e-4), pos


def test_clustered_cloud_parity(E):
    """Irregular input: a uniform background, dense Gaussian blobs and
    particles on two box faces, on two grids; exact keys, cells and
    neighbour sets, identical density, forces judged against FP64.  Its
    short, uneven y/z columns make many waves span several columns (their
    row unions too wide to stage, so the staged passes read those rows from
    global memory): staged and gathered passes give the same bits over a
    few steps."""
    rng = np.random.default_rng(23)
    parts = [rng.uniform(0.0, 1.0, size=(20000, 3))]
    for c in rng.uniform(0.2, 0.8, size=(6, 3)):
        parts.append(c + rng.normal(0.0, 0.03, size=(1500, 3)))
    face = rng.uniform(0.0, 1.0, size=(2000, 3))
    face[:1000, 0] = 0.0
    face[1000:, 1] = 1.0
    parts.append(face)
    pos = np.clip(np.concatenate(parts), 0.0, 1.0).astype(np.float32)
    params = E.make_params(pressure_stiffness=50.0, viscosity=0.01, dt=1e-4)
    for xsub in (1, 4):
        e = E.Engine(params)
        e.set_particles(pos)
        e.set_cell_subdivision(xsub)
        e.make_ready()
        step_and_compare(E, e, params, f"clustered cloud xsub {xsub}", vs_fp64=True)
        e.close()
    out = []
    for staged in (0, 1):
        e = E.Engine(params)
        e.set_particles(pos)
        e.make_ready()
        e.set_tuning(256, 256, 256, 8, 0, staged, staged)
        e.advance_steps(5)
        out.append((e.get_particles(), e.get_velocity(), e.get_density()))
        e.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_rows_longer_than_the_mask_parity(E):
    """Rows of more than 32 candidates (their head re-tested in the force
    pass), with the density and force passes reading candidates from global
    memory and with their rows staged in LDS (rows up to 128 candidates
    staged, longer ones read from global memory); the two forms give the
    same bits."""
    params, pos = compressed_block_scene()
    for xsub in (1, 4):
        out = []
        for staged in (0, 1):
            e = E.Engine(params)
            e.set_particles(pos)
            e.set_cell_subdivision(xsub)
            e.make_ready()
            e.set_tuning(256, 256, 256, 8, 0, staged, staged)
            d = step_and_compare(E, e, params, f"compressed block xsub {xsub} staged {staged}", vs_fp64=True)
            step_and_compare(E, e, params, f"compressed block xsub {xsub} staged {staged} step 2", vs_fp64=True)
            e.advance_steps(3)
            out.append((e.get_particles(), e.get_velocity()))
            e.close()
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.abs(d["acc"]).max() > 1e3


def test_double_dambreak_subset_parity(E):
    """The 1M configuration's physics on a 1/8 slab (131,072 particles) with
    its jitter and box, checked after 40 steps."""
    from simplefluid_amd import configs
    cfg = configs.double_dambreak_1m()
    pos = cfg["pos"][: 524288 // 4]
    e = make_engine(E, cfg["params"], pos)
    e.advance_steps(40)
    step_and_compare(E, e, cfg["params"], "double_dambreak slab")


@pytest.mark.parametrize("name,steps,staged", [("double_dambreak_1m", 30, 0), ("sphere_drop_4m", 20, 0),
                                               ("double_dambreak_1m", 30, 1), ("sphere_drop_4m", 20, 1)])
def test_full_size_step_parity(E, name, steps, staged):
    """The whole 1,048,576-particle double dam break and the whole 4.4M
    sphere drop (4,194,304 sampled points plus the resting layer): after
    some engine steps, one step from the identical state on the GPU and on
    the oracle, every particle: cell keys, permutation, cell table and
    neighbour sets bit-exact, density, acceleration and state within the
    module's tolerances.  staged: the density and force passes read each
    row's candidates from LDS."""
    from simplefluid_amd import configs
    cfg = getattr(configs, name)()
    p = cfg["params"]
    say(f"[{name}] {len(cfg['pos'])} particles, {steps} engine steps first, staged passes {staged}")
    e = make_engine(E, p, cfg["pos"])
    e.set_tuning(256, 256, 256, 8, 0, staged, staged)
    t = e.get_tuning()
    assert t["force_staged"] == staged and t["density_staged"] == staged
    e.advance_steps(steps)
    step_and_compare(E, e, p, f"{name} full")
    e.close()


def test_ensemble_256_step_parity(E):
    """All 256 scenes of the ensemble (5.7M particles, one launch per pass):
    after 20 batched steps, one more batched step against the oracle run
    scene by scene from the same state.  Each scene's slice of the batch's
    keys (less the scene's cell offset), permutation, cell table and
    neighbour lists (less the scene's first index) equals the oracle's bit
    for bit; density, acceleration and positions within the tolerances."""
    import oracle
    from simplefluid_amd import configs
    ens = configs.ensemble()
    params = [s["params"] for s in ens["scenes"]]
    counts = [s["pos"].shape[0] for s in ens["scenes"]]
    b = E.BatchEngine(params, counts)
    b.set_particles(np.concatenate([s["pos"] for s in ens["scenes"]]))
    b.make_ready()
    b.advance_steps(20)
    xs, vs, ids = b.debug_get_state()
    off = b.scene_offsets()
    b.debug_keep_acc(True)
    b.advance_steps(1)
    g = b.debug_step_outputs(acc=True)
    gs, ge = b.debug_cells()
    goff, glst = b.debug_neighbors()
    xg, vg, ig = b.debug_get_state()
    say(f"[ensemble] {len(ids)} particles in {len(counts)} scenes, {len(glst)} neighbour entries")
    coff = 0
    worst = 0.0
    for i, s in enumerate(ens["scenes"]):
        c = b.debug_constants(i)
        ncell = int(np.prod(c["G"]))
        a, z = int(off[i]), int(off[i + 1])
        sl = slice(a, z)
        o = oracle.Oracle(to_oracle(s["params"], c["xsub"]), threads=16)
        o.record_neighbors(True)
        o.set_state(xs[sl].astype(np.float64), vs[sl].astype(np.float64), ids[sl] - a)
        o.step()
        d = o.step_debug()
        assert np.array_equal(g["key"][sl] - coff, d["key"]), f"scene {i} keys"
        assert np.array_equal(g["perm"][sl] - a, d["perm"]), f"scene {i} perm"
        os_, oe = o.cells()
        cs_, ce_ = gs[coff:coff + ncell], ge[coff:coff + ncell]
        occ = oe > 0
        assert np.array_equal(occ, ce_ > 0), f"scene {i} occupied cells"
        assert np.array_equal(cs_[occ] - a, os_[occ]) and np.array_equal(ce_[occ] - a, oe[occ]), f"scene {i} cells"
        ooff, olst = o.neighbors()
        assert np.array_equal(goff[a:z + 1] - goff[a], ooff), f"scene {i} neighbour counts"
        assert np.array_equal(glst[goff[a]:goff[z]] - a, olst), f"scene {i} neighbour sets"
        assert np.all(np.abs(g["rho"][sl] - d["rho"]) <= RHO_RTOL * np.abs(d["rho"])), f"scene {i} density"
        scale = np.maximum(np.abs(d["acc"]).max(1, keepdims=True), 10.0)
        err = (np.abs(g["acc"][sl] - d["acc"]) / scale).max()
        assert err <= ACC_TOL, f"scene {i} acc {err}"
        xo, vo, io = o.state()
        assert np.array_equal(ig[sl] - a, io)
        assert np.abs(xg[sl] - xo).max() <= POS_ATOL, f"scene {i} positions"
        worst = max(worst, err)
        coff += ncell
    assert coff == len(gs)
    say(f"[ensemble] all scenes equal the oracle; worst scaled acc err {worst:.2e}")
    b.close()


PARAM_VARIANTS = {
    "attractive": dict(pressure_stiffness=5000.0, viscosity=0.01, use_attractive_pressure=1,
                       attractive_pressure_ratio=0.3),
    "inviscid_bouncy": dict(pressure_stiffness=500.0, viscosity=0.0, boundary_restitution=0.5,
                            gravity=(0.0, 0.0, -9.81)),
    "light_fluid": dict(pressure_stiffness=200.0, viscosity=0.2, rest_density=500.0, dt=2e-4,
                        gravity=(1.0, -5.0, 0.5)),
}


@pytest.mark.parametrize("variant", sorted(PARAM_VARIANTS))
def test_parameter_variants_parity(E, variant):
    """Attractive pressure (p < 0 scaled by the ratio instead of clamped),
    zero viscosity with a bouncy wall and gravity along z, and another rest
    density, step, viscosity and gravity: a step from the same state as the
    oracle after 30 steps of the engine."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=31)
    p = E.make_params(**PARAM_VARIANTS[variant])
    e = make_engine(E, p, cfg["pos"])
    e.advance_steps(30)
    step_and_compare(E, e, p, variant)
    e.close()


def test_sphere_drop_subset_parity(E):
    """The 4.4M sphere drop's physics (r 0.0014, h 0.0056, dt 5e-5) on the
    particles of its lower corner: the sphere's underside above the resting
    layer on the floor; checked after 20 steps."""
    from simplefluid_amd import configs
    cfg = configs.sphere_drop_4m()
    pos = cfg["pos"]
    sel = (pos[:, 0] < 0.45) & (pos[:, 1] < 0.35) & (pos[:, 2] < 0.45)
    sub = pos[sel]
    say(f"[sphere subset] {len(sub)} of {len(pos)} particles")
    assert 50_000 < len(sub) < 400_000
    e = make_engine(E, cfg["params"], sub)
    e.advance_steps(20)
    step_and_compare(E, e, cfg["params"], "sphere_drop corner")
    e.close()


@pytest.mark.parametrize("scene", ["SphereDrop", "CubeDrop", "Dambreak", "DoubleDambreak"])
def test_reference_scene_defaults_parity(E, scene):
    """Each reference scene (Source/SceneManager.cpp:21-173) at res 24 with
    the GUI defaults (stiffness 50000, viscosity 0.05, box [-1,1]^3): one step
    from the same state as the oracle after 25 steps of the engine."""
    p = E.default_params(24)
    pos = E.scene_generate(E.SCENE_NAMES.index(scene), 24)
    e = make_engine(E, p, pos)
    e.advance_steps(25)
    step_and_compare(E, e, p, f"{scene} res24")
    e.close()


def test_graph_replay_equals_eager(E):
    """Graph replays, with the program's own eager work between them
    (downloads, snapshots, synchronisation), equal the eagerly launched step
    after every replay: the one-step graph, the eight-step graph (calls of
    8, 9 and 17 steps) and mixes of both."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=3)
    g = make_engine(E, cfg["params"], cfg["pos"])
    r = make_engine(E, cfg["params"], cfg["pos"])
    g.set_use_graph(True)
    r.set_use_graph(False)
    for i in range(16):
        k = (1, 2, 3, 8, 9, 17)[i % 6]
        g.advance_steps(k)
        r.advance_steps(k)
        xg = g.get_particles()
        assert np.array_equal(xg, r.get_particles()), i
        assert np.array_equal(g.get_velocity(), r.get_velocity()), i
        if i % 5 == 2:
            g.snapshot_begin()
            g.advance_steps(2)
            r.advance_steps(2)
            g.synchronize()
            assert np.array_equal(g.snapshot_wait(), xg), i
    g.close()
    r.close()


def test_profile_counts_the_hash_only_when_it_runs(E):
    """Per-pass timings: the hash pass runs once after a state upload (later
    keys come from the force pass), so only that step reports time for it."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k()
    e = make_engine(E, cfg["params"], cfg["pos"])
    e.profile_enable(True)
    e.advance_steps(1)
    ms, n = e.profile_read()
    assert n == 1 and all(v > 0 for v in ms.values()), ms
    e.advance_steps(5)
    ms, n = e.profile_read()
    assert n == 5 and ms["hash"] == 0.0, ms
    assert all(ms[k] > 0 for k in ("sort", "reorder", "density", "force")), ms
    e.set_particles(cfg["pos"])
    e.make_ready()
    e.advance_steps(2)
    ms, n = e.profile_read()
    assert n == 2 and ms["hash"] > 0, ms
    e.profile_enable(False)
    e.close()


def test_runs_are_deterministic_and_tuning_invariant(E):
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=4)
    res = []
    for tune in ((256, 256, 256, 8), (256, 256, 256, 8), (64, 512, 128, 16), (512, 128, 64, 4, 1),
                 (256, 256, 256, 8, 0, 1), (64, 64, 128, 16, 1, 1), (512, 512, 256, 8, 0, 1),
                 (256, 256, 256, 8, 0, 0, 1), (64, 512, 128, 16, 1, 1, 1), (512, 64, 256, 8, 0, 1, 1)):
        e = make_engine(E, cfg["params"], cfg["pos"])
        e.set_tuning(*tune)
        e.advance_steps(30)
        res.append(e.get_particles())
        e.close()
    for r in res[1:]:
        assert np.array_equal(res[0], r)


@pytest.mark.parametrize("xsub", [1, 2, 4])
def test_staged_equals_gathered_on_random_boxes(E, xsub):
    """The staged passes (each row's candidates copied into LDS; a wave in
    one y/z column bounds the density pass's copy by two reads of the
    cell-start table) against the gathered ones, bit for bit, on boxes of
    several shapes whose particles crowd the grid's first and last cells
    (the table's first and last entries) at a few times rest density."""
    rng = np.random.default_rng(100 + xsub)
    for box in ((0.31, 0.47, 0.23), (0.5, 0.2, 0.35), (0.12, 0.6, 0.6)):
        bmax = np.array(box)
        pos = np.concatenate([
            rng.uniform(0.0, 1.0, size=(12000, 3)) * bmax,
            bmax - rng.uniform(0.0, 0.08, size=(3000, 3)),
            rng.uniform(0.0, 0.08, size=(1500, 3)),
        ]).astype(np.float32)
        say(f"random box {box} xsub {xsub}: {len(pos)} particles")
        params = E.make_params(box_max=box)
        out = []
        for staged in (0, 1):
            e = E.Engine(params)
            e.set_particles(pos)
            e.set_cell_subdivision(xsub)
            e.make_ready()
            e.set_tuning(256, 256, 256, 8, 0, staged, staged)
            e.advance_steps(4)
            out.append((e.get_particles(), e.get_velocity(), e.get_density()))
            e.close()
        assert all(np.isfinite(x).all() for x in out[0]), (box, xsub)
        for a, b in zip(*out):
            assert np.array_equal(a, b), (box, xsub)


def test_persistent_launch_matches_one_block_per_tile(E):
    """Density and force as persistent blocks over per-XCD tile queues give
    the same bits as one block per tile (1M particles: more tiles than
    resident blocks, so the queues are in use)."""
    from simplefluid_amd import configs
    cfg = configs.double_dambreak_1m()
    res = []
    for tune in ((64, 64, 256, 8, 0), (64, 64, 256, 8, 1), (512, 256, 256, 8, 1), (512, 64, 256, 8, 1, 1),
                 (512, 512, 256, 8, 0, 1), (64, 256, 256, 8, 1, 1, 1), (512, 256, 256, 8, 0, 0, 1)):
        e = make_engine(E, cfg["params"], cfg["pos"])
        e.set_tuning(*tune)
        e.advance_steps(8)
        res.append((e.get_particles(), e.get_density()))
        e.close()
    for r in res[1:]:
        assert np.array_equal(res[0][0], r[0]) and np.array_equal(res[0][1], r[1])


def test_sort_modes_agree(E):
    """Cell-count sort and 8-bit radix sort: same permutation, same results."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=8)
    out = []
    for mode in (E.SORT_CELL_COUNT, E.SORT_RADIX):
        e = make_engine(E, cfg["params"], cfg["pos"])
        e.set_sort_mode(mode)
        assert e.get_sort_mode() == mode
        e.advance_steps(20)
        if mode == E.SORT_RADIX:
            step_and_compare(E, e, cfg["params"], "radix sort step")
        else:
            e.advance_steps(1)
        out.append((e.get_particles(), e.debug_step_outputs()["perm"]))
        e.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("xsub", [1, 2, 4])
def test_cell_subdivision_parity(E, xsub):
    """Every x subdivision gives oracle-exact neighbour sets and keys."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=12)
    e = make_engine(E, cfg["params"], cfg["pos"])
    e.set_cell_subdivision(xsub)
    assert e.debug_constants()["xsub"] == xsub
    e.advance_steps(30)
    step_and_compare(E, e, cfg["params"], f"xsub {xsub}")


def test_cfl_frames_on_device_match_the_host_loop(E):
    """Adaptive step: sph_run_frame replays a step graph that stops itself
    once the frame reaches 1/30 s (reading the frame state once per batch);
    frame times, step counts and the state equal the step-by-step host loop
    bit for bit, including the first frame after an upload."""
    import time
    p = E.default_params(24)
    p.dt = 5e-3            # the CFL rule sets the step
    p.cfl_number = 0.4
    pos = E.scene_generate(E.SCENE_DAMBREAK, 24)
    res = {}
    for graph in (True, False):
        e = make_engine(E, p, pos)
        e.set_use_graph(graph)
        frames = [e.run_frame() for _ in range(3)]
        e.synchronize()
        t0 = time.perf_counter()
        frames += [e.run_frame() for _ in range(5)]
        ms = (time.perf_counter() - t0) / 5 * 1e3
        res[graph] = (frames, e.get_particles(), e.get_velocity())
        say(f"[cfl frames] device loop {graph}: {ms:.2f} ms per frame, steps {[f[1] for f in frames]}")
        e.close()
    assert res[True][0] == res[False][0]
    assert all(t >= np.float32(0.0333333333) for t, _ in res[True][0])
    assert len(set(k for _, k in res[True][0])) > 1   # the step length varies
    assert np.array_equal(res[True][1], res[False][1]) and np.array_equal(res[True][2], res[False][2])


def test_fixed_step_frame_runs_its_steps_in_one_call(E):
    """With a fixed step, a frame's step count follows from the float sum
    of dt alone: run_frame gives the bits of the same number of single
    steps, and a dt too small to move the sum is refused, not looped on."""
    p = E.default_params(16)
    pos = E.scene_generate(E.SCENE_DAMBREAK, 16)
    a = make_engine(E, p, pos)
    b = make_engine(E, p, pos)
    t, steps = a.run_frame()
    ft = np.float32(0.0)
    n = 0
    while ft < 0.0333333333:
        ft = np.float32(ft + np.float32(p.dt))
        n += 1
    assert steps == n and t == float(ft)
    for _ in range(n):
        b.advance_steps(1)
    assert np.array_equal(a.get_particles(), b.get_particles())
    assert np.array_equal(a.get_velocity(), b.get_velocity())
    a.close()
    b.close()
    q = E.default_params(16)
    q.dt = 1e-12
    c = make_engine(E, q, pos)
    with pytest.raises(E.SPHError) as ei:
        c.run_frame()
    assert ei.value.code == E.SPH_E_INVALID and "too small" in str(ei.value)
    c.close()


def test_tuning_changes_between_captured_frames(E):
    """Block sizes and launch forms changed between frames of the adaptive
    step's device loop (each change re-captures the step graph, with tile
    tables for block sizes not used before) give the bits of a run with
    fixed tuning."""
    p = E.default_params(24)
    p.dt = 5e-3
    p.cfl_number = 0.4
    pos = E.scene_generate(E.SCENE_DAMBREAK, 24)
    res = []
    for plan in (((256, 256, 256, 8),), ((64, 128, 256, 8), (512, 64, 256, 8, 1), (128, 512, 256, 8)),
                 ((64, 128, 256, 8, 0, 1), (512, 64, 256, 8, 1, 1, 1), (128, 512, 256, 8, 0, 0, 1))):
        e = make_engine(E, p, pos)
        frames = []
        for i in range(6):
            e.set_tuning(*plan[i % len(plan)])
            frames.append(e.run_frame())
        res.append((frames, e.get_particles(), e.get_velocity()))
        e.close()
    for r in res[1:]:
        assert res[0][0] == r[0]
        assert np.array_equal(res[0][1], r[1]) and np.array_equal(res[0][2], r[2])


def test_cfl_step_parity(E):
    """Adaptive step: the engine's step lengths follow the oracle's CFL rule
    (identical first step from rest, then within 1e-4 as the trajectories
    differ by round-off), a step from an identical state matches the oracle
    in every intermediate, and a frame sums steps to >= 1/30 s."""
    import oracle
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=21)
    p = E.make_params(pressure_stiffness=5000.0, viscosity=0.01, dt=5e-3, cfl_number=0.3)
    e = make_engine(E, p, cfg["pos"])
    o = oracle.Oracle(to_oracle(p, e.debug_constants()["xsub"]), threads=16)
    o.set_state(cfg["pos"].astype(np.float64))
    dts = []
    for s in range(20):
        dts.append(e.advance_frame())
        o.step()
        if s == 0:
            assert dts[0] == np.float32(o.last_dt())
        assert abs(dts[-1] - o.last_dt()) <= 1e-4 * o.last_dt(), (s, dts[-1], o.last_dt())
    say(f"[cfl] steps {dts[0]:.3e} .. {dts[-1]:.3e}")
    assert dts[-1] < dts[0] <= 5e-3          # the fluid speeds up, steps shrink
    step_and_compare(E, e, p, "cfl step")
    t, n = e.run_frame()
    assert t >= np.float32(0.0333333333) and n >= 1
    with pytest.raises(E.SPHError):
        e.rollout_forward(10, 5)            # the adjoint needs a fixed dt
    e.close()
    with pytest.raises(E.SPHError):
        E.BatchEngine([p, p], [10, 10])     # a batch shares one dt


def test_frame_export_matches_golden(E, tmp_path):
    """sph_export_frame writes the state as (n, 6) float32 .npy in original
    order; after 40 steps of the reference CubeDrop (resolution 16, GUI
    defaults) it matches the FP64 oracle fixture to FP32 round-off."""
    import os
    p = E.default_params(16)
    e = make_engine(E, p, E.scene_generate(E.SCENE_CUBE_DROP, 16))
    e.set_cell_subdivision(4)
    e.advance_steps(40)
    path = str(tmp_path / "frame.npy")
    e.export_frame(path)
    x, v = E.load_frame(path)
    assert np.array_equal(x, e.get_particles()) and np.array_equal(v, e.get_velocity())
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    xg, vg = E.load_frame(os.path.join(root, "tests", "golden", "frame_cubedrop16_step40.npy"))
    say(f"[frame] pos err {np.abs(x - xg).max():.2e} vel err {np.abs(v - vg).max():.2e}")
    assert np.abs(x - xg).max() < 2e-6
    assert np.abs(v - vg).max() < 1e-4
    e.close()


def test_snapshot_overlaps_later_steps(E):
    """A snapshot holds the positions of the moment it began, whatever steps
    run before it is collected; a second snapshot replaces the first."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=14)
    e = make_engine(E, cfg["params"], cfg["pos"])
    e.advance_steps(5)
    x0 = e.get_particles()
    e.snapshot_begin()
    e.advance_steps(20)
    s0 = e.snapshot_wait()
    assert np.array_equal(s0, x0)
    x1 = e.get_particles()
    assert not np.array_equal(x1, x0)
    e.snapshot_begin()
    e.advance_steps(3)
    e.snapshot_begin()                     # waits for the first copy on the GPU
    x2 = e.get_particles()
    assert np.array_equal(e.snapshot_wait(), x2)
    with pytest.raises(E.SPHError):
        e.snapshot_wait()                  # nothing in flight
    # A new particle set between begin and wait: the snapshot is refused and
    # dropped, and the next one has the new count.
    big = np.concatenate([cfg["pos"], cfg["pos"][:1000] + np.float32(0.4)])
    e.snapshot_begin()
    e.set_particles(big)
    e.make_ready()
    with pytest.raises(E.SPHError) as ei:
        e.snapshot_wait()
    assert ei.value.code == E.SPH_E_INVALID and "count changed" in str(ei.value)
    with pytest.raises(E.SPHError) as ei:
        e.snapshot_wait()
    assert ei.value.code == E.SPH_E_NOT_READY
    e.advance_steps(2)
    e.snapshot_begin()
    s2 = e.snapshot_wait()
    assert s2.shape == big.shape and np.array_equal(s2, e.get_particles())
    e.close()


def test_get_particles_original_order(E):
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k()
    e = make_engine(E, cfg["params"], cfg["pos"])
    assert np.array_equal(e.get_particles(), cfg["pos"])
    e.advance_steps(10)
    xs, vs, ids = e.debug_get_state()
    p = e.get_particles()
    assert np.array_equal(p[ids], xs)
    assert np.array_equal(e.get_velocity()[ids], vs)
    rho = e.get_density()
    assert np.isfinite(rho).all() and rho.min() > 0


def test_run_frame_matches_reference_loop(E):
    """Frame = steps until the float sum of dt reaches 0.0333333333
    (Source/Simulator.cpp:46-51)."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k()
    e = make_engine(E, cfg["params"], cfg["pos"][:4096])
    t, s = e.run_frame()
    ft, n = np.float32(0), 0
    while ft < 0.0333333333:
        ft = np.float32(ft + np.float32(cfg["params"].dt))
        n += 1
    assert s == n and t == ft


def _session_to_stop_time(E, p, pos, every=15):
    """Frames until the float sim time reaches stopTime (Source/Simulator.cpp:44-51,
    m_SimTime is a float); returns frame count, the distinct step counts per
    frame and (v^2/2 + g.y) per unit mass every `every` frames, after checking
    that the state is finite and inside the clamp bounds."""
    e = make_engine(E, p, pos)
    g = -float(p.gravity[1])
    r = np.float32(p.particle_radius)
    lo = np.array(p.box_min[:], np.float32) + r
    hi = np.array(p.box_max[:], np.float32) - r
    sim_time, frames, steps, energies = np.float32(0), 0, set(), []
    while sim_time < np.float32(p.stop_time):
        t, k = e.run_frame()
        sim_time = np.float32(sim_time + np.float32(t))
        frames += 1
        steps.add(k)
        if frames % every == 0:
            x, v = e.get_particles(), e.get_velocity()
            assert np.isfinite(x).all() and np.isfinite(v).all()
            assert (x >= lo).all() and (x <= hi).all()
            ke = 0.5 * float(np.mean((v.astype(np.float64) ** 2).sum(1)))
            energies.append(ke + g * float(np.mean(x[:, 1].astype(np.float64) - lo[1])))
    e.close()
    return frames, steps, energies


def test_reference_double_dambreak_runs_to_stop_time(E):
    """The reference's default session: DoubleDambreak at res 24 with the GUI
    defaults runs its 5 s in 150 frames of equal step counts, every state
    finite and inside the clamp bounds, and v^2/2 + g.y per unit mass stays
    bounded: below twice its value at rest, and lower at the end than half a
    second in.  (The collapse itself adds energy at this stiffness: poly6
    density with a spiky pressure gradient is not an energy-conserving
    pairing; the walls' half-space terms keep it from growing afterwards.)"""
    p = E.default_params(24)
    pos = E.scene_generate(E.SCENE_DOUBLE_DAMBREAK, 24)
    g = -float(p.gravity[1])
    e0 = g * float(np.mean(pos[:, 1].astype(np.float64) - (np.float32(p.box_min[1]) + np.float32(p.particle_radius))))
    frames, steps, energies = _session_to_stop_time(E, p, pos)
    say(f"[stop time, B 50000] {frames} frames of {sorted(steps)} steps, energy per unit mass {e0:.3f} -> "
        f"{[round(v, 2) for v in energies]}")
    assert frames == 150 and len(steps) == 1
    assert max(energies) < 2.0 * e0 and energies[-1] < energies[0]


def test_softer_dambreak_session_loses_energy(E):
    """The same session at stiffness 5000 only loses energy (viscosity and
    the inelastic walls): v^2/2 + g.y per unit mass stays below its value at
    rest and ends lower than it was after the first half second."""
    p = E.default_params(24, pressure_stiffness=5000.0)
    pos = E.scene_generate(E.SCENE_DOUBLE_DAMBREAK, 24)
    g = -float(p.gravity[1])
    e0 = g * float(np.mean(pos[:, 1].astype(np.float64) - (np.float32(p.box_min[1]) + np.float32(p.particle_radius))))
    frames, _, energies = _session_to_stop_time(E, p, pos)
    say(f"[stop time, B 5000] energy per unit mass {e0:.3f} -> {[round(v, 3) for v in energies]}")
    assert frames == 150
    assert max(energies) < e0 and energies[-1] < energies[0]


def test_qt_shaped_solver(E):
    p = E.default_params(16)
    s = E.SPHSolver(p)
    s.setupScene(E.SCENE_CUBE_DROP)
    assert s.getNumParticles() == 16 ** 3
    s.makeReady()
    dt = s.advanceFrame()
    assert dt == np.float32(1e-4)
    x = s.getParticles()
    assert x.shape == (4096, 3) and np.isfinite(x).all()
    assert s.getVelocity()[:, 1].mean() < 0


def test_particle_count_change_recaptures_graph(E):
    """A handle stepped through its captured graph, then given a different
    particle count, steps exactly like a fresh handle with that count."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=6)
    small, big = cfg["pos"][:4096], cfg["pos"][:12288]
    e = make_engine(E, cfg["params"], small)
    e.advance_steps(10)
    e.set_particles(big)
    e.make_ready()
    e.advance_steps(10)
    f = make_engine(E, cfg["params"], big)
    f.advance_steps(10)
    assert np.array_equal(e.get_particles(), f.get_particles())
    assert np.array_equal(e.get_velocity(), f.get_velocity())
    e.close()
    f.close()


def test_single_particle_and_empty_input(E):
    """One particle has only itself as neighbour: density m W(0) plus the
    floor's half-space share (it starts 0.6 h above the floor; the value is
    the float32 oracle's, same operations), p clamped to 0 (below rest
    density), so no pressure of its own or of the wall acts and it falls
    under gravity along the float32 symplectic-Euler closed form, onto the
    floor clamp.  Zero particles cannot be made ready."""
    import oracle
    p = E.make_params()
    x0 = np.array([[0.5, 0.012, 0.5]], np.float32)
    e = make_engine(E, p, x0)
    c = e.debug_constants()
    x = x0.copy()
    v = np.zeros((1, 3), np.float32)
    g = np.array(p.gravity[:], np.float32)
    dt = np.float32(p.dt)
    lo = np.float32(p.box_min[1]) + np.float32(p.particle_radius)
    t = np.float32(c["h2"])
    for k in range(200):
        e.advance_steps(1)
        if k == 0:
            o = oracle.Oracle(to_oracle(p, c["xsub"]))
            o.set_state(x0.astype(np.float64))
            o.step()
            rho_o = np.float32(o.step_debug()["rho"][0])
            assert rho_o > np.float32(c["cD"]) * np.float32(np.float32(t * t) * t)
            assert e.get_density()[0] == rho_o
        v = v + dt * g                             # a = (0 + 0) + g: no pressure, no viscosity
        x = x + dt * v
        if x[0, 1] < lo:
            x[0, 1] = lo
            if v[0, 1] < 0:
                v[0, 1] = -np.float32(p.boundary_restitution) * v[0, 1]
    assert np.array_equal(e.get_particles(), x) and np.array_equal(e.get_velocity(), v)
    e.close()
    z = E.Engine(p)
    z.set_particles(np.zeros((0, 3), np.float32))
    with pytest.raises(E.SPHError) as ei:
        z.make_ready()
    assert ei.value.code == E.SPH_E_NOT_READY
    z.close()


def test_largest_grid_steps_and_a_larger_one_is_refused(E):
    """Maximum grid: 1024 x 1024 x 1023 cells, just under the 2^30 - 2 limit
    that keeps every cell-start offset inside 32-bit buffer addressing.  A
    close pair and a lone particle at the far corner (the top of the
    cell-start table) step bit for bit as the same particles do in a small
    box with the same upper corner (two-term sums do not depend on the
    order); a 1024^3 grid is refused when the handle is created."""
    h = 0.001
    cs = h * (1.0 + 2.0 ** -10)
    hi = (1023.5 * cs, 1023.5 * cs, 1022.5 * cs)
    c = np.array(hi) - 10 * h
    pos = np.array([c, c + [0.6 * h, 0.0, 0.0], c - [0.0, 0.0, 20 * h]], np.float32)
    vel = np.array([[0.1, 0.0, 0.0], [-0.1, 0.05, 0.0], [0.0, 0.0, 0.2]], np.float32)
    kw = dict(kernel_radius=h, particle_radius=h / 4, pressure_stiffness=50.0, dt=1e-4)
    res = []
    for lo in ((0.0, 0.0, 0.0), tuple(np.array(hi) - 40 * h)):
        e = E.Engine(E.make_params(box_min=lo, box_max=hi, **kw))
        e.set_cell_subdivision(1)
        e.set_particles(pos, vel)
        e.make_ready()
        if lo == (0.0, 0.0, 0.0):
            assert int(np.prod(np.asarray(e.debug_constants()["G"], np.int64))) == 1024 * 1024 * 1023
        e.advance_steps(20)
        res.append((e.get_particles(), e.get_velocity(), e.get_density()))
        e.close()
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)
    assert res[0][2][0] > res[0][2][2]   # the pair sees each other, the lone particle only itself
    with pytest.raises(E.SPHError) as ei:
        E.Engine(E.make_params(box_min=(0.0, 0.0, 0.0), box_max=(hi[0], hi[1], hi[0]), **kw))
    assert ei.value.code == E.SPH_E_INVALID and "grid cells" in str(ei.value)


def test_device_memory_limit(E):
    """A handle's buffers are charged to it; an allocation past its cap fails
    with SPH_E_OOM and a message naming the limit, and the handle works once
    the cap is raised, bit for bit like an uncapped one."""
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k()
    n = cfg["pos"].shape[0]
    e = make_engine(E, cfg["params"], cfg["pos"])
    e.advance_steps(3)
    used = e.device_memory_used()
    say(f"device bytes per particle (65k dam break): {used / n:.1f}")
    assert 100 * n <= used <= 1000 * n
    e.set_particles(cfg["pos"])
    e.make_ready()
    e.advance_steps(3)
    assert e.device_memory_used() == used, "re-uploading the same count must not grow the buffers"
    x_ref, v_ref, _ = e.debug_get_state()

    c = E.Engine(cfg["params"])
    c.set_device_memory_limit(used // 2)
    with pytest.raises(E.SPHError) as ei:
        c.set_particles(cfg["pos"])
        c.make_ready()
        c.advance_steps(3)
    assert ei.value.code == E.SPH_E_OOM and "device memory limit" in str(ei.value), str(ei.value)
    assert c.device_memory_used() <= used // 2
    c.set_device_memory_limit(0)
    c.set_particles(cfg["pos"])
    c.make_ready()
    c.advance_steps(3)
    assert c.device_memory_used() == used
    x, v, _ = c.debug_get_state()
    assert np.array_equal(x, x_ref) and np.array_equal(v, v_ref)
    c.close()
    e.close()


def test_error_codes(E):
    from simplefluid_amd import engine
    p = engine.make_params(kernel_radius=-1.0)
    with pytest.raises(engine.SPHError):
        engine.Engine(p)
    e = engine.Engine(engine.make_params())
    with pytest.raises(engine.SPHError) as ei:
        e.advance_steps(1)
    assert ei.value.code == engine.SPH_E_NOT_READY
    with pytest.raises(engine.SPHError):
        e.set_params(engine.make_params(dt=0.0))


def test_nan_and_out_of_box_particles_are_contained(E):
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k()
    pos = cfg["pos"][:8192].copy()
    pos[5] = [np.nan, 0.5, 0.5]
    pos[6] = [5.0, -3.0, 0.5]
    e = make_engine(E, cfg["params"], pos)
    e.advance_steps(3)
    x = e.get_particles()
    ok = np.ones(len(x), bool)
    ok[5] = False
    assert np.isfinite(x[ok]).all()
    assert (x[6] <= 1.0).all() and (x[6] >= 0.0).all()


def _nan32(negative):
    return np.array([0xFFC00000 if negative else 0x7FC00000], np.uint32).view(np.float32)[0]


def test_nan_positions_of_both_signs_inside_the_block(E):
    """Particles inside the 65k dam-break block get NaN coordinates of both
    signs, on one axis or all three; a NaN y clamps into the floor's cell
    row, beside the floor layer's particles.  One step from the same state:
    keys, permutation, cell table and neighbour sets equal the oracle's (its
    predicate is d^2 < h^2, which no NaN distance passes), density, force
    and state agree where finite and are NaN at the same places; particles
    farther than h from where the NaN particles were get bit for bit the
    density a NaN-free step gives them (and the state, farther than 2h); and
    every finite particle stays finite over 20 more steps."""
    import oracle
    from simplefluid_amd import configs
    cfg = configs.dambreak_65k(jitter=1e-6, seed=12)
    p = cfg["params"]
    pos = cfg["pos"].copy()
    h = float(p.kernel_radius)
    plant = [((0.155, 0.305, 0.155), (1,), True), ((0.105, 0.205, 0.205), (1,), False),
             ((0.205, 0.405, 0.105), (0,), True), ((0.055, 0.105, 0.255), (0, 1, 2), True),
             ((0.255, 0.505, 0.055), (2,), False), ((0.155, 0.005, 0.155), (1,), True),
             ((0.305, 0.005, 0.045), (1,), False)]
    bad = []
    pos_nan = pos.copy()
    for c, axes, neg in plant:
        i = int(np.argmin(((pos - np.asarray(c, np.float32)) ** 2).sum(1)))
        bad.append(i)
        for ax in axes:
            pos_nan[i, ax] = _nan32(neg)
    bad = np.array(bad)
    assert np.signbit(pos_nan[bad[0], 1]) and not np.signbit(pos_nan[bad[1], 1])
    say(f"[nan] NaN particles {bad.tolist()}")

    e = make_engine(E, p, pos_nan)
    xs, vs, ids = e.debug_get_state()
    o = oracle.Oracle(to_oracle(p, e.debug_constants()["xsub"]), threads=16)
    o.record_neighbors(True)
    o.set_state(xs.astype(np.float64), vs.astype(np.float64), ids)
    e.debug_keep_acc(True)
    e.advance_steps(1)
    g = e.debug_step_outputs(acc=True)
    o.step()
    d = o.step_debug()
    assert np.array_equal(g["key"], d["key"]) and np.array_equal(g["perm"], d["perm"])
    gs, ge = e.debug_cells()
    os_, oe = o.cells()
    assert np.array_equal(gs, os_) and np.array_equal(ge, oe)
    goff, glst = e.debug_neighbors()
    ooff, olst = o.neighbors()
    assert np.array_equal(goff, ooff) and np.array_equal(glst, olst), "neighbour sets differ"
    sid = d["sorted_id"]
    nan_sorted = np.isin(sid, bad)
    # The NaN particles have no neighbours, not even themselves, and are no
    # finite particle's neighbour.
    assert (np.diff(goff)[nan_sorted] == 0).all()
    assert not np.isin(glst, np.flatnonzero(nan_sorted)).any()
    assert np.all(np.abs(g["rho"] - d["rho"]) <= RHO_RTOL * np.abs(d["rho"]))
    fin = np.isfinite(d["acc"]).all(1)
    assert np.array_equal(fin, np.isfinite(g["acc"]).all(1))
    scale = np.maximum(np.abs(d["acc"][fin]).max(1, keepdims=True), 10.0)
    assert (np.abs(g["acc"][fin] - d["acc"][fin]) / scale).max() <= ACC_TOL
    xg, vg, _ = e.debug_get_state()
    xo, vo, _ = o.state()
    np.testing.assert_allclose(xg, xo, rtol=0, atol=POS_ATOL, equal_nan=True)
    e.debug_keep_acc(False)

    # Against the same step without the NaNs, in original order.
    x_nan, v_nan, rho_nan = e.get_particles(), e.get_velocity(), e.get_density()
    ref = make_engine(E, p, pos)
    ref.advance_steps(1)
    x_ref, v_ref, rho_ref = ref.get_particles(), ref.get_velocity(), ref.get_density()
    dist = np.sqrt(((pos[:, None, :].astype(np.float64) - pos[bad][None, :, :]) ** 2).sum(2)).min(1)
    # Density sees the NaN particles' absence within h of where they were;
    # force, through the neighbours' densities, within 2h.
    far = dist > 1.01 * h
    far2 = dist > 2.02 * h
    assert far2.sum() > 60000
    assert np.array_equal(rho_nan[far], rho_ref[far])
    assert np.array_equal(x_nan[far2], x_ref[far2]) and np.array_equal(v_nan[far2], v_ref[far2])
    ok = np.ones(len(pos), bool)
    ok[bad] = False
    e.advance_steps(20)
    x = e.get_particles()
    assert np.isfinite(x[ok]).all() and np.isfinite(e.get_velocity()[ok]).all()
    say(f"[nan] ok: {int(far.sum())} / {int(far2.sum())} far particles bit-identical, {len(glst)} neighbour entries")
    e.close()
    ref.close()


def test_overfull_cell_is_flagged_and_radix_mode_orders_it(E):
    """2,000 particles with NaN positions clamp into one cell, past the
    1,024 members the cell-count sort ranks by scanning the cell: the step
    stays in bounds, keeps the arrival order there and reports SPH_E_DEVICE
    saying so; the radix sort has no such limit and equals the oracle."""
    import oracle
    from simplefluid_amd import configs
    p = configs.dambreak_65k()["params"]
    blk = configs.lattice_block((16, 16, 16), (0.4, 0.4, 0.4))
    pile = np.full((2000, 3), np.nan, np.float32)
    pos = np.concatenate([blk, pile])
    e = make_engine(E, p, pos)
    e.advance_steps(1)
    with pytest.raises(E.SPHError) as ei:
        e.synchronize()
    assert ei.value.code == E.SPH_E_DEVICE and "more than 1024" in str(ei.value), str(ei.value)
    e.close()
    e = E.Engine(p)
    e.set_particles(pos)
    e.set_sort_mode(E.SORT_RADIX)
    e.make_ready()
    e.advance_steps(2)
    xs, vs, ids = e.debug_get_state()
    o = oracle.Oracle(to_oracle(p, e.debug_constants()["xsub"]), threads=16)
    o.record_neighbors(True)
    o.set_state(xs.astype(np.float64), vs.astype(np.float64), ids)
    e.advance_steps(1)
    g = e.debug_step_outputs()
    o.step()
    d = o.step_debug()
    assert np.array_equal(g["perm"], d["perm"])
    goff, glst = e.debug_neighbors()
    ooff, olst = o.neighbors()
    assert np.array_equal(goff, ooff) and np.array_equal(glst, olst)
    x = e.get_particles()
    assert np.isfinite(x[:len(blk)]).all() and np.isnan(x[len(blk):]).all()
    e.synchronize()
    e.close()


def test_batched_equals_per_scene(E):
    """Scenes stepped together give bit-identical results to each scene alone."""
    from simplefluid_amd import configs
    ens = configs.ensemble(nscenes=6, seed=7, nmin=2000, nmax=20000)
    params = [s["params"] for s in ens["scenes"]]
    counts = [s["pos"].shape[0] for s in ens["scenes"]]
    b = E.BatchEngine(params, counts)
    b.set_cell_subdivision(2)   # one grid for the batch and the single scenes
    b.set_particles(np.concatenate([s["pos"] for s in ens["scenes"]]))
    b.make_ready()
    b.advance_steps(40)
    xb = b.get_particles()
    off = b.scene_offsets()
    for i, s in enumerate(ens["scenes"]):
        e = E.Engine(s["params"])
        e.set_cell_subdivision(2)
        e.set_particles(s["pos"])
        e.make_ready()
        e.advance_steps(40)
        assert np.array_equal(e.get_particles(), xb[off[i]:off[i + 1]]), f"scene {i}"
        e.close()


def test_batched_with_empty_scenes(E):
    """Scenes with no particles (between others and last) own no tiles and
    change nothing: the other scenes step bit for bit as in a batch without
    them, and the empty ones report empty ranges."""
    from simplefluid_amd import configs
    ens = configs.ensemble(nscenes=2, seed=17, nmin=3000, nmax=9000)
    params = [s["params"] for s in ens["scenes"]]
    pos = [s["pos"] for s in ens["scenes"]]
    out = []
    for layout in ((0, 1), (0, None, 1, None)):
        ps = [params[i] if i is not None else params[0] for i in layout]
        counts = [len(pos[i]) if i is not None else 0 for i in layout]
        b = E.BatchEngine(ps, counts)
        b.set_cell_subdivision(2)
        b.set_particles(np.concatenate([pos[i] for i in layout if i is not None]))
        b.make_ready()
        b.advance_steps(30)
        off = b.scene_offsets()
        assert [int(off[k + 1] - off[k]) for k in range(len(layout))] == counts
        out.append((b.get_particles(), b.get_velocity()))
        b.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_batched_persistent_queues_match(E):
    """A batch of six ~100k-particle scenes (more 64-lane tiles than resident
    blocks, so the per-XCD tile queues are in use) gives the same bits with
    persistent blocks, and with the density and force passes staging their
    rows in LDS, as with one block per tile."""
    from simplefluid_amd import configs
    ens = configs.ensemble(nscenes=6, seed=12, nmin=90000, nmax=110000)
    params = [s["params"] for s in ens["scenes"]]
    counts = [s["pos"].shape[0] for s in ens["scenes"]]
    assert sum(counts) // 64 > 256 * 32
    out = []
    for persistent, staged in ((0, 0), (1, 0), (1, 1)):
        b = E.BatchEngine(params, counts)
        b.set_particles(np.concatenate([s["pos"] for s in ens["scenes"]]))
        b.make_ready()
        b.set_tuning(64, 64, 256, 8, persistent, staged, staged)
        b.advance_steps(15)
        out.append((b.get_particles(), b.get_velocity()))
        b.close()
    for r in out[1:]:
        assert np.array_equal(out[0][0], r[0]) and np.array_equal(out[0][1], r[1])


def test_batched_parity_vs_oracle(E):
    """One batched step of 4 scenes vs the oracle run scene by scene."""
    import oracle
    from simplefluid_amd import configs
    ens = configs.ensemble(nscenes=4, seed=9, nmin=3000, nmax=12000)
    params = [s["params"] for s in ens["scenes"]]
    counts = [s["pos"].shape[0] for s in ens["scenes"]]
    b = E.BatchEngine(params, counts)
    b.set_particles(np.concatenate([s["pos"] for s in ens["scenes"]]))
    b.make_ready()
    b.advance_steps(20)
    xs, vs, ids = b.debug_get_state()
    off = b.scene_offsets()
    b.advance_steps(1)
    xg, _, ig = b.debug_get_state()
    for i, s in enumerate(ens["scenes"]):
        sl = slice(off[i], off[i + 1])
        o = oracle.Oracle(to_oracle(s["params"], b.debug_constants(i)["xsub"]), threads=8)
        o.set_state(xs[sl].astype(np.float64), vs[sl].astype(np.float64), ids[sl] - off[i])
        o.step()
        xo, _, io = o.state()
        assert np.array_equal(ig[sl] - off[i], io)
        assert np.abs(xg[sl] - xo).max() <= POS_ATOL
