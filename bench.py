"""Benchmark of the SPH step engine on one MI355X.

Default workload: the 1,048,576-particle double dam break (two jittered
128x64x64 blocks, box [0,2]x[0,1]x[0,2], r 0.005, h 0.02, dt 1e-4), FP32.
A step = counting sort of the cell keys (keys and per-cell counts come from
the previous step's force pass, so no hash pass runs per step), reorder + cell
table, density + EOS, pressure/viscosity force + integrate + clamp.

Prints ONE JSON line.  `value` is particle-steps/s over K graph-replayed
steps with the state resident in HBM.  `roofline` is the dominant kernel's
algorithmic bytes per launch over its mean duration, measured with HIP events
on the engine's stream in a second run of K steps; `cpu_baseline` times the
oracle's OpenMP C++ step on this host over a bounded sample (all host
threads, and one thread).  Extras: `batched` (256-scene ensemble, one launch
per pass, vs the same scenes launched one engine at a time), `adjoint`
(rollout reverse pass vs forward wall-clock, 16 x 32,768 particles, 200
steps, checkpoint every 20), `dambreak_65k` (the 65,536-particle dam break,
1000 steps), `sphere_4m` (the 4.19M-particle sphere drop: a seeded uniform
sample of a ball with minimum spacing 1.8r over a resting layer, r scaled to
0.0014), `reference_session` (frames per second of the
reference application's default session: DoubleDambreak at resolution 24,
GUI defaults, 334 steps per 1/30 s frame; `cpu_baseline` times one of its
frames on the oracle) and `slab_1_of_8` (one GPU's share of the 1M scene
sharded 8 ways by x slabs: the fullest slab with its 2h ghost bands).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M (plain vs batched); % HBM roofline; adjoint/fwd ratio"
HBM_PEAK_GBS = 8000.0
# MI355X's published FP32 vector rate, 157.3 TFLOP/s, as lane operations (a
# fused multiply-add is one): 78.6 T lane-op/s.  These passes issue a wave64
# VALU instruction every 3.3-3.7 SIMD cycles on average, so their mix runs
# between the 4-cycle rate of a 16-lane SIMD (39.3 T) and this figure.
VALU_PEAK_TOPS = 78.64
# Algorithmic bytes per particle and pass (FP32, unpadded float3): hash 20,
# sort 16 (one ideal pass), reorder 52, cell table 4 + 8 per occupied cell,
# density 16, force 40, integrate + clamp 60.  The engine fuses hash into the
# integrate pass and integrate into the force kernel.
B_ALG = dict(hash=20.0, sort=16.0, reorder=56.0, density=16.0, force=100.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(e, k):
    e.synchronize()
    t0 = time.perf_counter()
    e.advance_steps(k)
    e.synchronize()
    return time.perf_counter() - t0


def plain(args):
    from simplefluid_amd import configs, engine
    cfg = configs.double_dambreak_1m()
    n = cfg["pos"].shape[0]
    e = engine.Engine(cfg["params"])
    e.set_particles(cfg["pos"])
    e.make_ready()
    if args.tuning:
        e.set_tuning(*[int(x) for x in args.tuning.split(",")])
    elif not args.no_autotune:
        e.autotune()
    log("tuning", e.get_tuning())
    e.advance_steps(args.warmup)
    e.synchronize()
    start = e.debug_get_state()
    t = time_steps(e, args.steps)
    value = n * args.steps / t
    log(f"plain: {n} particles, {args.steps} steps in {t:.4f}s -> {value:.4e} particle-steps/s")
    # Per-pass durations over the same K steps again (the state is restored to
    # the start of the timed region), with HIP events between the passes on
    # the engine's stream.
    e.debug_set_state(*start)
    e.profile_enable(True)
    e.advance_steps(args.steps)
    ms, nst = e.profile_read()
    e.profile_enable(False)
    per = {k: v / max(nst, 1) for k, v in ms.items()}
    log("ms per pass", {k: round(v, 5) for k, v in per.items()})
    e.get_particles()  # first call allocates the handle's page-locked buffer
    t0 = time.perf_counter()
    e.get_particles()
    download_ms = (time.perf_counter() - t0) * 1e3
    s, en = e.debug_cells()
    occ = int((en > 0).sum())
    # Per step, the compulsory traffic of each pass: hash (its key is computed
    # in the force pass, but the bytes are the pass's), sort, reorder,
    # density, force + integrate, plus 8 bytes per occupied cell (its start
    # and end in the cell table): 208 + 8 C_occ / N.
    b_step = sum(B_ALG.values()) + 8.0 * occ / n
    dom = max(("sort", "reorder", "density", "force"), key=lambda k: per[k])
    dom_bytes = B_ALG[dom] * n + (8.0 * occ if dom == "reorder" else 0.0)
    achieved = dom_bytes / (per[dom] * 1e-3) / 1e9
    tuning = e.get_tuning()
    traffic = load_traffic(dom, tuning)
    roof = {"bound": "hbm", "kernel": profile_key(dom, tuning), "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic}
    valu = load_valu(dom, tuning)
    if valu:
        # Lane operations per launch (rocprofv3 SQ_INSTS_VALU x 64) over the
        # live mean duration: the issue-rate ceiling these kernels meet first.
        lane_ops = valu * 64.0 / (per[dom] * 1e-3)
        roof["valu"] = {"achieved": round(lane_ops / 1e12, 2), "peak": VALU_PEAK_TOPS, "unit": "T lane-op/s",
                        "frac": round(lane_ops / 1e12 / VALU_PEAK_TOPS, 4)}
    step_ms = t / args.steps * 1e3
    step_roof = b_step * value / 1e9
    extra = {"step_bytes_alg_per_particle": round(b_step, 2),
             "step_achieved_GBs": round(step_roof, 1),
             "step_frac": round(step_roof / HBM_PEAK_GBS, 4),
             "pass_ms": {k: round(v, 5) for k, v in per.items()},
             "occupied_cells": occ, "tuning": e.get_tuning(),
             "download_ms": round(download_ms, 3)}
    e.close()
    return value, step_ms, roof, extra, cfg


def copy_ceiling_gbs():
    """Achieved HBM bandwidth of a plain 256 MiB device-to-device copy (read +
    write bytes over the best of 10 copies timed with HIP events): the
    practical ceiling next to the 8 TB/s peak.  Uses the HIP runtime the
    engine library loaded."""
    import ctypes as C
    from simplefluid_amd import engine
    engine.load_library()
    hip = C.CDLL("libamdhip64.so")
    vp = C.c_void_p
    nbytes = 256 << 20
    a, b = vp(), vp()
    e0, e1 = vp(), vp()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        hip.hipMemset(a, 1, C.c_size_t(nbytes))
        best = None
        for _ in range(12):
            hip.hipEventRecord(e0, None)
            hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)  # hipMemcpyDeviceToDevice
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = C.c_float()
            hip.hipEventElapsedTime(C.byref(ms), e0, e1)
            best = ms.value if best is None else min(best, ms.value)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return round(2 * nbytes / (best * 1e-3) / 1e9, 1)


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    import glob
    sockets = set()
    for f in glob.glob("/sys/devices/system/cpu/cpu[0-9]*/topology/physical_package_id"):
        try:
            sockets.add(open(f).read().strip())
        except OSError:
            pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable,
            "sockets": len(sockets) or None,
            "numa_nodes": len(glob.glob("/sys/devices/system/node/node[0-9]*"))}


def profile_key(kernel, tuning=None):
    """Key of a pass's figures in the committed profile: density and force
    have one entry per form (rows gathered from global memory, or staged in
    LDS: "<pass>_staged")."""
    if tuning and kernel in ("density", "force") and tuning.get(f"{kernel}_staged"):
        return f"{kernel}_staged"
    return kernel


def load_valu(kernel, tuning=None):
    """VALU instructions per launch (wave-instructions) from the committed profile."""
    path = os.path.join(ROOT, "profiles", "valu.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(profile_key(kernel, tuning), {}).get("valu_insts_per_launch")
    except Exception:
        return None


def load_traffic(kernel, tuning=None):
    """HBM bytes per launch from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d.get(profile_key(kernel, tuning), {}).get("bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(cfg, budget_s):
    """The oracle's OpenMP FP32 step on this host: median of three timed
    chunks after warm-up steps, with all host threads and with one thread,
    on the 1M workload, plus a sample of the 65k dam break's 1000 steps."""
    import oracle
    from simplefluid_amd import configs
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)

    def run(c, nthreads, budget, max_steps, warm):
        o = oracle.Oracle(oracle.from_params(c["params"], 4), threads=nthreads)
        o.set_state(c["pos"].astype(np.float64))
        o.step(warm)
        first = o.step(1)  # per-step estimate sizes the chunks
        k = int(max(1, min(max_steps, budget / 3.0 / max(first, 1e-9))))
        rates = sorted(c["pos"].shape[0] * k / o.step(k) for _ in range(3))
        return rates[1], k

    n = cfg["pos"].shape[0]
    v, k = run(cfg, threads, budget_s, 20, 9)
    v1, k1 = run(cfg, 1, budget_s / 3.0, 3, 1)
    small = configs.dambreak_65k()
    vs, ks = run(small, threads, budget_s / 3.0, 300, 10)
    # One frame of the reference's default session (334 steps) after one warm-up step.
    from simplefluid_amd import engine
    rp, rpos = reference_session(engine)
    o = oracle.Oracle(oracle.from_params(rp, 4), threads=threads)
    o.set_state(rpos.astype(np.float64))
    o.step(1)
    frame_s = o.step(334)
    return {"value": round(v, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle C++ FP32 step (OpenMP, {threads} threads), {n} particles of the same workload: "
                      f"median of 3 chunks of {k} steps after 10 warm-up steps",
            "single_thread": {"value": round(v1, 1), "sample": f"median of 3 chunks of {k1} steps"},
            "dambreak_65k": {"value": round(vs, 1), "threads": threads,
                             "sample": f"median of 3 chunks of {ks} of its 1000 steps after 11 warm-up steps"},
            "reference_session": {"ms_per_frame": round(frame_s * 1e3, 1), "threads": threads,
                                  "sample": "one 334-step frame of DoubleDambreak at resolution 24 "
                                            "(GUI defaults) after one warm-up step"},
            "host": host_info()}


def batched(args):
    from simplefluid_amd import configs, engine
    ens = configs.ensemble(nscenes=256, seed=3)
    params = [s["params"] for s in ens["scenes"]]
    counts = [s["pos"].shape[0] for s in ens["scenes"]]
    N = int(sum(counts))
    b = engine.BatchEngine(params, counts)
    b.set_particles(np.concatenate([s["pos"] for s in ens["scenes"]]))
    b.make_ready()
    if args.tuning:
        b.set_tuning(*[int(x) for x in args.tuning.split(",")])
    elif not args.no_autotune:
        b.autotune()
    b.advance_steps(5)
    k = args.batch_steps
    tb = time_steps(b, k)
    vb = N * k / tb
    tuning = b.get_tuning()
    b.close()
    if args.no_per_scene:
        log(f"batched: {N} particles in 256 scenes, {k} steps: {vb:.4e}")
        return {"scenes": 256, "particles": N, "steps": k, "value": round(vb, 1), "unit": "particle-steps/s",
                "tuning": tuning}
    # Per-scene launches: one engine per scene, stepped scene by scene.
    engines = []
    for s in ens["scenes"]:
        e = engine.Engine(s["params"])
        e.set_particles(s["pos"])
        e.make_ready()
        e.advance_steps(2)
        engines.append(e)
    for e in engines:
        e.synchronize()
    t0 = time.perf_counter()
    for e in engines:
        e.advance_steps(k)
    for e in engines:
        e.synchronize()
    ts = time.perf_counter() - t0
    vs = N * k / ts
    for e in engines:
        e.close()
    log(f"batched: {N} particles in 256 scenes, {k} steps: {vb:.4e} vs per-scene {vs:.4e}")
    return {"scenes": 256, "particles": N, "steps": k, "value": round(vb, 1),
            "per_scene_value": round(vs, 1), "speedup": round(vb / vs, 3),
            "unit": "particle-steps/s", "tuning": tuning}


def tune_leg(e, args):
    """--tuning pins the launch shape (profiling runs), else the autotuner picks."""
    if args.tuning:
        e.set_tuning(*[int(x) for x in args.tuning.split(",")])
    elif not args.no_autotune:
        e.autotune()


def dambreak_65k_leg(args):
    """The 65,536-particle dam break over its full 1000 steps."""
    from simplefluid_amd import configs, engine
    cfg = configs.dambreak_65k()
    e = engine.Engine(cfg["params"])
    e.set_particles(cfg["pos"])
    e.make_ready()
    tune_leg(e, args)
    e.advance_steps(5)
    t = time_steps(e, cfg["steps"])
    n = cfg["pos"].shape[0]
    out = {"particles": n, "steps": cfg["steps"], "value": round(n * cfg["steps"] / t, 1),
           "ms_per_step": round(t / cfg["steps"] * 1e3, 5), "unit": "particle-steps/s",
           "tuning": e.get_tuning()}
    e.close()
    log("dambreak_65k", out)
    return out


def sphere_4m_leg(args):
    """The 4.4M-particle sphere drop over a resting layer."""
    from simplefluid_amd import configs, engine
    cfg = configs.sphere_drop_4m()
    n = cfg["pos"].shape[0]
    e = engine.Engine(cfg["params"])
    e.set_particles(cfg["pos"])
    e.make_ready()
    tune_leg(e, args)
    e.advance_steps(5)
    k = min(args.sphere_steps, cfg["steps"])
    t = time_steps(e, k)
    e.profile_enable(True)
    e.advance_steps(20)
    ms, nst = e.profile_read()
    e.profile_enable(False)
    out = {"particles": n, "steps": k, "value": round(n * k / t, 1),
           "ms_per_step": round(t / k * 1e3, 5), "unit": "particle-steps/s",
           "pass_ms": {kk: round(v / max(nst, 1), 5) for kk, v in ms.items()},
           "x_subdivision": int(e.debug_constants()["xsub"]), "tuning": e.get_tuning()}
    e.close()
    log("sphere_4m", out)
    return out


def slab_leg(args):
    """One GPU's share of the 1M double dam break sharded 8 ways by x slabs
    (cuts at particle-count quantiles, a 2h ghost band each inner side): the
    fullest slab, owned particles and ghosts stepped as one scene.  The ghost
    band's exchange between GPUs is not part of it."""
    from simplefluid_amd import configs, engine
    cfg = configs.slab_share(configs.double_dambreak_1m(), 8)
    e = engine.Engine(cfg["params"])
    e.set_particles(cfg["pos"])
    e.make_ready()
    tune_leg(e, args)
    e.advance_steps(args.warmup)
    t = time_steps(e, args.steps)
    e.profile_enable(True)
    e.advance_steps(100)
    ms, nst = e.profile_read()
    e.profile_enable(False)
    out = {"shards": 8, "slab": cfg["shard"], "x_range": [round(v, 5) for v in cfg["x_range"]],
           "owned": cfg["owned"], "ghosts": cfg["ghosts"], "particles": int(cfg["pos"].shape[0]),
           "steps": args.steps, "ms_per_step": round(t / args.steps * 1e3, 5),
           "owned_particle_steps_per_s": round(cfg["owned"] * args.steps / t, 1),
           "pass_ms": {k: round(v / max(nst, 1), 5) for k, v in ms.items()}, "tuning": e.get_tuning()}
    e.close()
    log("slab_1_of_8", out)
    return out


def small_and_large(args):
    """The 65k dam break over its full 1000 steps, the 4.4M sphere drop, the
    reference application's default session, and one GPU's slab of the 1M
    scene sharded 8 ways."""
    return {"dambreak_65k": dambreak_65k_leg(args), "sphere_4m": sphere_4m_leg(args),
            "reference_session": reference_session_leg(args), "slab_1_of_8": slab_leg(args)}


REF_FRAMES = 30


def reference_session(engine):
    """The reference's default session (Source/Simulator.cpp:33-63): the
    DoubleDambreak scene at resolution 24 with the GUI defaults."""
    return engine.default_params(24), engine.scene_generate(engine.SCENE_DOUBLE_DAMBREAK, 24)


def reference_session_leg(args):
    """Frames per second of the reference's default session: each frame is
    the steps whose float sum reaches 1/30 s (334 of 1e-4 s), run on the
    device; 30 frames after 3 untimed ones."""
    from simplefluid_amd import engine
    p, pos = reference_session(engine)
    e = engine.Engine(p)
    e.set_particles(pos)
    e.make_ready()
    t_tune = 0.0
    if not args.no_autotune:
        t0 = time.perf_counter()
        e.autotune()
        t_tune = time.perf_counter() - t0
    for _ in range(3):
        e.run_frame()
    e.synchronize()
    t0 = time.perf_counter()
    steps = 0
    for _ in range(REF_FRAMES):
        steps += e.run_frame()[1]
    e.synchronize()
    t = time.perf_counter() - t0
    e_tuning = e.get_tuning()
    # Per-pass times (HIP events on the engine's stream) over 100 more steps.
    e.profile_enable(True)
    e.advance_steps(100)
    ms, nsteps = e.profile_read()
    e.profile_enable(False)
    pass_us = {k: round(v / max(nsteps, 1) * 1e3, 1) for k, v in ms.items()}
    e.close()
    out = {"scene": "DoubleDambreak", "resolution": 24, "particles": int(pos.shape[0]), "frames": REF_FRAMES,
           "steps_per_frame": steps // REF_FRAMES, "ms_per_frame": round(t / REF_FRAMES * 1e3, 3),
           "frames_per_s": round(REF_FRAMES / t, 1), "pass_us": pass_us, "tuning": e_tuning,
           "autotune_ms": round(t_tune * 1e3, 1)}
    log("reference_session", out)
    return out


def rollout_targets(rc):
    """Final positions of each scene rolled out with its perturbed stiffness,
    viscosity and initial velocities (the fitting targets)."""
    from simplefluid_amd import engine
    sc = rc["scenes"]
    counts = [s["pos"].shape[0] for s in sc]
    t = engine.BatchEngine([s["target_params"] for s in sc], counts)
    t.set_particles(np.concatenate([s["pos"] for s in sc]), np.concatenate([s["target_vel"] for s in sc]))
    t.make_ready()
    t.advance_steps(rc["steps"])
    out = t.get_particles().astype(np.float64)
    t.close()
    return out


def rollout_loss_grad_input(xT, target, offsets):
    """Sum over scenes of the per-scene MSE of final positions, and dL/dxT."""
    L, g = 0.0, np.zeros_like(xT)
    for a, b in zip(offsets[:-1], offsets[1:]):
        d = xT[a:b] - target[a:b]
        s = 1.0 / (3 * (b - a))
        L += s * float((d * d).sum())
        g[a:b] = 2 * s * d
    return L, g.astype(np.float32)


def adjoint(args):
    """The differentiable-rollout workload: 16 dam-break scenes x 32,768
    particles, 200 steps, checkpoints every 20; loss = MSE of final positions
    against rollouts with perturbed stiffness, viscosity and initial
    velocities; gradients for stiffness, viscosity and initial velocity."""
    from simplefluid_amd import configs, engine
    rc = configs.rollout_scenes()
    if args.checkpoint_every:
        rc["checkpoint_every"] = args.checkpoint_every
    sc = rc["scenes"]
    counts = [s["pos"].shape[0] for s in sc]
    target = rollout_targets(rc)
    b = engine.BatchEngine([s["params"] for s in sc], counts)
    pos0 = np.concatenate([s["pos"] for s in sc])
    vel0 = np.concatenate([s["vel"] for s in sc])
    b.set_particles(pos0, vel0)
    b.make_ready()
    tune_leg(b, args)  # launch shape of the forward passes (the state is restored)
    # Warm-up of both passes on the timed shape, so that neither timed region
    # allocates its checkpoints or tape.
    b.rollout_forward(rc["steps"], rc["checkpoint_every"])
    b.rollout_backward(np.zeros((len(pos0), 3), np.float32))
    b.set_particles(pos0, vel0)
    b.make_ready()
    b.synchronize()
    t0 = time.perf_counter()
    b.rollout_forward(rc["steps"], rc["checkpoint_every"])
    b.synchronize()
    tf = time.perf_counter() - t0
    off = b.scene_offsets()
    loss, g = rollout_loss_grad_input(b.get_particles().astype(np.float64), target, off)
    t0 = time.perf_counter()
    dB, dmu, dv0 = b.rollout_backward(g)
    tb = time.perf_counter() - t0
    tuning = b.get_tuning()
    b.close()
    log(f"adjoint: forward {tf:.4f}s backward {tb:.4f}s ratio {tb / tf:.3f} loss {loss:.4e}")
    return {"scenes": len(sc), "particles": int(sum(counts)), "steps": rc["steps"],
            "checkpoint_every": rc["checkpoint_every"], "forward_s": round(tf, 5),
            "backward_s": round(tb, 5), "ratio": round(tb / tf, 3), "loss": loss,
            "dL_dstiffness_scene0": float(dB[0]), "dL_dviscosity_scene0": float(dmu[0]),
            "tuning": tuning}


def calibrate_copy():
    """A plain device copy of 256 MiB (reads 256 MiB, writes 256 MiB): the
    reference point that scales FETCH_SIZE / WRITE_SIZE to HBM bytes."""
    import torch
    x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copy calibration: 5 x 268435456 B read + written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--sphere-steps", type=int, default=100)
    ap.add_argument("--no-per-scene", action="store_true",
                    help="batched leg without the per-scene comparison (profiling)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--only", choices=["plain", "batched", "adjoint", "dambreak65k", "sphere4m", "reference",
                                       "slab"],
                    default=None,
                    help="run one leg alone (profiling)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--checkpoint-every", type=int, default=0,
                    help="adjoint leg: steps between checkpoints (default: the workload's 20; "
                         "200 tapes every step and recomputes nothing)")
    ap.add_argument("--tuning", type=str, default="",
                    help="density,force,reorder blocks, sort items [,persistent 0/1 [,staged force 0/1]], "
                         "e.g. 512,256,256,8,0,1 (skips autotune)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 256 MiB device copy (known bytes) for counter calibration, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported; the engine runs "
              "on one GPU (the batched ensemble would shard by scene with no communication, the "
              "plain path by spatial slabs with a ghost band)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.only == "batched":
        print(json.dumps({"batched": batched(args)}), flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint": adjoint(args)}), flush=True)
        return
    if args.only == "dambreak65k":
        print(json.dumps({"dambreak_65k": dambreak_65k_leg(args)}), flush=True)
        return
    if args.only == "sphere4m":
        print(json.dumps({"sphere_4m": sphere_4m_leg(args)}), flush=True)
        return
    if args.only == "reference":
        print(json.dumps({"reference_session": reference_session_leg(args)}), flush=True)
        return
    if args.only == "slab":
        print(json.dumps({"slab_1_of_8": slab_leg(args)}), flush=True)
        return
    value, step_ms, roof, extra, cfg = plain(args)
    out = {"metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "double_dambreak_1m", "particles": int(cfg["pos"].shape[0]),
                      "box": [2.0, 1.0, 2.0], "particle_radius": 0.005, "kernel_radius": 0.02,
                      "dt": 1e-4, "stiffness": 50.0, "viscosity": 0.01},
           "roofline": roof, "step": extra}
    try:
        out["roofline"]["copy_ceiling_GBs"] = copy_ceiling_gbs()
    except Exception as ex:
        log("copy ceiling not measured:", ex)
    if not args.no_extras:
        out["batched"] = batched(args)
        out["adjoint"] = adjoint(args)
        out.update(small_and_large(args))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
