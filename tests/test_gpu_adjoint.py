"""GPU hand-written adjoint vs FP64 central differences, the CPU reverse-mode
oracle and a torch.autograd wrapper (bar: within 1e-3 relative of FP64 finite
differences at 4,096 particles x 50 steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-3


def say(*a):
    print(*a, flush=True)


def case(n3=(16, 16, 16), seed=0, B=50.0, mu=0.01, tscale=0.003):
    from simplefluid_amd import configs, engine
    rng = np.random.default_rng(seed)
    pos = configs.lattice_block(n3, (0.0, 0.0, 0.0), jitter=1e-6, rng=rng)
    vel = rng.normal(0, 0.1, pos.shape).astype(np.float32)
    target = (pos + rng.normal(0, tscale, pos.shape)).astype(np.float64)
    p = engine.make_params(pressure_stiffness=B, viscosity=mu, dt=1e-4)
    return p, pos, vel, target


def gpu_grads(p, pos, vel, target, steps, every, xsub=0, tuning=None):
    from simplefluid_amd import engine
    e = engine.Engine(p)
    e.set_cell_subdivision(xsub)
    e.set_particles(pos, vel)
    e.make_ready()
    if tuning is not None:
        e.set_tuning(*tuning)
    e.rollout_forward(steps, every)
    xT = e.get_particles().astype(np.float64)
    n = len(pos)
    scale = 1.0 / (3 * n)
    L = scale * ((xT - target) ** 2).sum()
    g = (2 * scale * (xT - target)).astype(np.float32)
    dB, dmu, dv0 = e.rollout_backward(g)
    xT2 = e.get_particles()
    e.close()
    return L, float(dB[0]), float(dmu[0]), dv0.astype(np.float64), xT, xT2


def to_oracle(p):
    import oracle
    return oracle.from_params(p)


def test_adjoint_vs_fp64_central_differences_4096x50():
    import oracle
    p, pos, vel, target = case()
    steps = 50
    L, dB, dmu, dv0, xT, xT2 = gpu_grads(p, pos, vel, target, steps, every=20)
    assert np.array_equal(xT.astype(np.float32), xT2), "state after backward must be the final state"
    op = to_oracle(p)
    B0, mu0 = float(p.pressure_stiffness), float(p.viscosity)
    # Step 1e-6 * max(1, |parameter|): with an absolute 1e-6 on B = 50 the loss
    # moves by ~1e-17 on 9e-6, below FP64 resolution, and the difference
    # quotient is round-off (it converges to the adjoint as the step grows).
    eps = 1e-6
    epsB = eps * max(1.0, abs(B0))

    def loss(**kw):
        return oracle.rollout_loss_grad(op, pos, kw.get("vel", vel), target, steps, grad=False,
                                        threads=16, B=kw.get("B", B0), mu=kw.get("mu", mu0))

    fdB = (loss(B=B0 + epsB) - loss(B=B0 - epsB)) / (2 * epsB)
    fdmu = (loss(mu=mu0 + eps) - loss(mu=mu0 - eps)) / (2 * eps)
    rng = np.random.default_rng(5)
    u = rng.normal(size=vel.shape)
    u /= np.linalg.norm(u)
    fdv = (loss(vel=vel + eps * u) - loss(vel=vel - eps * u)) / (2 * eps)
    gv = float((dv0 * u).sum())
    say(f"dL/dB gpu {dB:.6e} fd {fdB:.6e}; dL/dmu gpu {dmu:.6e} fd {fdmu:.6e}; dv0.u gpu {gv:.6e} fd {fdv:.6e}")
    assert abs(dB - fdB) <= GRAD_RTOL * abs(fdB)
    assert abs(dmu - fdmu) <= GRAD_RTOL * abs(fdmu)
    assert abs(gv - fdv) <= GRAD_RTOL * abs(fdv)
    # the full FP64 reverse-mode oracle agrees on the whole dv0 field
    _, oB, omu, odv0 = oracle.rollout_loss_grad(op, pos, vel, target, steps, threads=16)
    assert abs(oB - fdB) <= 1e-4 * abs(fdB)
    rel = np.linalg.norm(dv0 - odv0) / np.linalg.norm(odv0)
    say(f"dv0 field rel err vs FP64 adjoint {rel:.3e}")
    assert rel <= GRAD_RTOL


def test_adjoint_vs_fp32_oracle_and_checkpoint_invariance():
    import oracle
    p, pos, vel, target = case(n3=(10, 12, 10), seed=3, B=200.0, mu=0.05)
    steps = 24
    res = [gpu_grads(p, pos, vel, target, steps, every) for every in (1, 7, 24)]
    for r in res[1:]:
        assert r[1] == res[0][1] and r[2] == res[0][2] and np.array_equal(r[3], res[0][3])
    _, oB, omu, odv0 = oracle.rollout_loss_grad(to_oracle(p), pos, vel, target, steps, fp64=False)
    L, dB, dmu, dv0 = res[0][:4]
    assert abs(dB - oB) <= GRAD_RTOL * abs(oB)
    assert abs(dmu - omu) <= GRAD_RTOL * abs(omu)
    assert np.linalg.norm(dv0 - odv0) <= GRAD_RTOL * np.linalg.norm(odv0)


def test_adjoint_with_rows_longer_than_the_mask():
    """A block compressed to spacing 1.25r beside a 2r block: rows of more
    than 32 candidates, whose head the adjoint kernels re-test, in the same
    waves as ordinary rows; against the FP32 reverse-mode oracle."""
    import oracle
    from simplefluid_amd import configs, engine
    R = configs.R
    rng = np.random.default_rng(21)
    a = configs.lattice_block((8, 12, 8), (0.1, 0.0, 0.1))
    i, j, k = np.meshgrid(np.arange(10), np.arange(12), np.arange(10), indexing="ij")
    b = np.array([0.1 + 17 * R, R, 0.1 + R]) + 1.25 * R * np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    pos = np.concatenate([a, b.astype(np.float32)])
    vel = rng.normal(0, 0.1, pos.shape).astype(np.float32)
    target = (pos + rng.normal(0, 0.003, pos.shape)).astype(np.float64)
    p = engine.make_params(pressure_stiffness=50.0, viscosity=0.01, dt=1e-4)
    steps = 6
    _, oB, omu, odv0 = oracle.rollout_loss_grad(to_oracle(p), pos, vel, target, steps, fp64=False)
    for xsub in (1, 4):
        L, dB, dmu, dv0 = gpu_grads(p, pos, vel, target, steps, 3, xsub=xsub)[:4]
        say(f"xsub {xsub}: dL/dB gpu {dB:.6e} oracle {oB:.6e}; dL/dmu gpu {dmu:.6e} oracle {omu:.6e}")
        assert abs(dB - oB) <= GRAD_RTOL * abs(oB)
        assert abs(dmu - omu) <= GRAD_RTOL * abs(omu)
        assert np.linalg.norm(dv0 - odv0) <= GRAD_RTOL * np.linalg.norm(odv0)


def test_staged_adjoint_equals_gathered():
    """The adjoint passes with each row's candidates staged in LDS (the form
    taken with the staged forward force pass) give the gathered passes'
    gradients bit for bit: on a dam-break block, and on a compressed block
    whose rows exceed the mask and, at x subdivision 1, the staging slot."""
    from simplefluid_amd import configs, engine
    R = configs.R
    rng = np.random.default_rng(22)
    a = configs.lattice_block((8, 12, 8), (0.1, 0.0, 0.1))
    i, j, k = np.meshgrid(np.arange(10), np.arange(12), np.arange(10), indexing="ij")
    b = np.array([0.1 + 17 * R, R, 0.1 + R]) + 1.25 * R * np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    pos = np.concatenate([a, b.astype(np.float32)])
    vel = rng.normal(0, 0.1, pos.shape).astype(np.float32)
    target = (pos + rng.normal(0, 0.003, pos.shape)).astype(np.float64)
    p = engine.make_params(pressure_stiffness=50.0, viscosity=0.01, dt=1e-4)
    cases = [(p, pos, vel, target, 6, 3, 1), (p, pos, vel, target, 6, 3, 4)]
    p2, pos2, vel2, target2 = case(n3=(20, 24, 16), seed=5, B=200.0, mu=0.02)
    cases.append((p2, pos2, vel2, target2, 10, 5, 0))
    for p_, x, v, t, steps, every, xsub in cases:
        say(f"staged vs gathered adjoint: n {len(x)} steps {steps} xsub {xsub}")
        gat = gpu_grads(p_, x, v, t, steps, every, xsub=xsub, tuning=(256, 256, 256, 8, 0, 0, 0))
        stg = gpu_grads(p_, x, v, t, steps, every, xsub=xsub, tuning=(256, 256, 256, 8, 0, 1, 1))
        assert stg[1] == gat[1] and stg[2] == gat[2]
        assert np.array_equal(stg[3], gat[3])
        assert np.array_equal(stg[4], gat[4])


def test_adjoint_attractive_pressure_and_zero_viscosity():
    """Adjoint through the attractive branch of the EOS (p < 0 scaled by the
    ratio) and with zero viscosity, against the FP32 reverse-mode oracle."""
    import oracle
    from simplefluid_amd import engine
    for kw in (dict(use_attractive_pressure=1, attractive_pressure_ratio=0.3, viscosity=0.02),
               dict(viscosity=0.0)):
        p, pos, vel, target = case(n3=(10, 12, 10), seed=9, B=400.0)
        for k, v in kw.items():
            setattr(p, k, v)
        steps = 12
        _, oB, omu, odv0 = oracle.rollout_loss_grad(to_oracle(p), pos, vel, target, steps, fp64=False)
        L, dB, dmu, dv0 = gpu_grads(p, pos, vel, target, steps, 4)[:4]
        say(f"{kw}: dL/dB gpu {dB:.6e} oracle {oB:.6e}; dL/dmu gpu {dmu:.6e} oracle {omu:.6e}")
        assert abs(dB - oB) <= GRAD_RTOL * abs(oB)
        assert abs(dmu - omu) <= GRAD_RTOL * abs(omu)
        assert np.linalg.norm(dv0 - odv0) <= GRAD_RTOL * np.linalg.norm(odv0)


def test_backward_refuses_a_stale_rollout():
    """After new particles or new parameters the recorded rollout no longer
    describes the handle: the reverse pass refuses it (before, a changed
    particle count made it read past its checkpoints), and a new forward
    rollout can be reversed again."""
    from simplefluid_amd import engine
    p, pos, vel, target = case(n3=(8, 8, 8), seed=2)
    e = engine.Engine(p)
    e.set_particles(pos, vel)
    e.make_ready()
    g = np.zeros((len(pos), 3), np.float32)
    e.rollout_forward(6, 3)
    bigger = np.concatenate([pos, pos[:100] + np.float32(0.3)])
    e.set_particles(bigger)
    e.make_ready()
    with pytest.raises(engine.SPHError) as ei:
        e.rollout_backward(np.zeros((len(bigger), 3), np.float32))
    assert ei.value.code == engine.SPH_E_NOT_READY and "changed" in str(ei.value)
    e.set_particles(pos, vel)
    e.make_ready()
    e.rollout_forward(6, 3)
    p2 = engine.make_params(pressure_stiffness=80.0, viscosity=0.01, dt=1e-4)
    e.set_params(p2)
    e.make_ready()
    with pytest.raises(engine.SPHError):
        e.rollout_backward(g)
    e.rollout_forward(6, 3)
    e.rollout_backward(g)
    e.close()


@pytest.mark.parametrize("staged", [0, 1])
def test_batched_adjoint_equals_single_scene(staged):
    """A batch's gradients equal each scene's own handle's bit for bit, with
    the rows gathered and with them staged in LDS (forward and adjoint)."""
    from simplefluid_amd import engine
    cases = [case(n3=(8, 10, 8), seed=s, B=40.0 + 20 * s, mu=0.01 * (s + 1)) for s in range(3)]
    steps, every = 16, 5
    tuning = (256, 256, 256, 8, 0, staged, staged)
    b = engine.BatchEngine([c[0] for c in cases], [len(c[1]) for c in cases])
    b.set_cell_subdivision(4)   # same grid as the single-scene handles below
    b.set_particles(np.concatenate([c[1] for c in cases]), np.concatenate([c[2] for c in cases]))
    b.make_ready()
    b.set_tuning(*tuning)
    b.rollout_forward(steps, every)
    xT = b.get_particles().astype(np.float64)
    tgt = np.concatenate([c[3] for c in cases])
    off = b.scene_offsets()
    g = np.zeros_like(xT)
    for i in range(3):
        sl = slice(off[i], off[i + 1])
        g[sl] = 2 * (1.0 / (3 * (off[i + 1] - off[i]))) * (xT[sl] - tgt[sl])   # as gpu_grads forms it
    dB, dmu, dv0 = b.rollout_backward(g.astype(np.float32))
    for i, (p, pos, vel, target) in enumerate(cases):
        _, sB, smu, sdv0, _, _ = gpu_grads(p, pos, vel, target, steps, every, xsub=4, tuning=tuning)
        assert dB[i] == np.float32(sB) and dmu[i] == np.float32(smu)
        assert np.array_equal(dv0[off[i]:off[i + 1]], sdv0.astype(np.float32))


def test_batched_adjoint_with_an_empty_scene():
    """An empty scene in a batch gets zero parameter gradients and leaves the
    other scenes' gradients bit for bit as without it."""
    from simplefluid_amd import engine
    cases = [case(n3=(8, 10, 8), seed=s, B=40.0 + 20 * s, mu=0.01 * (s + 1)) for s in range(2)]
    steps, every = 12, 5
    res = []
    for layout in ((0, 1), (0, None, 1)):
        ps = [cases[i][0] if i is not None else cases[0][0] for i in layout]
        counts = [len(cases[i][1]) if i is not None else 0 for i in layout]
        b = engine.BatchEngine(ps, counts)
        b.set_cell_subdivision(4)
        b.set_particles(np.concatenate([cases[i][1] for i in layout if i is not None]),
                        np.concatenate([cases[i][2] for i in layout if i is not None]))
        b.make_ready()
        b.rollout_forward(steps, every)
        xT = b.get_particles().astype(np.float64)
        tgt = np.concatenate([cases[i][3] for i in layout if i is not None])
        dB, dmu, dv0 = b.rollout_backward((2e-3 * (xT - tgt)).astype(np.float32))
        res.append((np.asarray(dB), np.asarray(dmu), dv0, xT))
        b.close()
    (dB0, dmu0, dv00, x0), (dB1, dmu1, dv01, x1) = res
    assert np.array_equal(x0, x1) and np.array_equal(dv00, dv01)
    assert dB1[1] == 0.0 and dmu1[1] == 0.0
    assert np.array_equal(dB0, dB1[[0, 2]]) and np.array_equal(dmu0, dmu1[[0, 2]])


def test_torch_autograd_function():
    import torch
    import oracle
    from tests import sph_torch
    from simplefluid_amd import engine
    p, pos, vel, target = case(n3=(8, 8, 8), seed=6, B=80.0, mu=0.02)
    steps = 20
    e = engine.Engine(p)
    v0 = torch.tensor(vel, dtype=torch.float64, requires_grad=True)
    B = torch.tensor([80.0], dtype=torch.float64, requires_grad=True)
    mu = torch.tensor([0.02], dtype=torch.float64, requires_grad=True)
    xT = sph_torch.rollout(v0, B, mu, e, [p], pos, steps, every=8)
    L = ((xT - torch.tensor(target)) ** 2).sum() / (3 * len(pos))
    L.backward()
    _, oB, omu, odv0 = oracle.rollout_loss_grad(to_oracle(p), pos, vel, target, steps)
    assert abs(B.grad.item() - oB) <= GRAD_RTOL * abs(oB)
    assert abs(mu.grad.item() - omu) <= GRAD_RTOL * abs(omu)
    assert np.linalg.norm(v0.grad.numpy() - odv0) <= GRAD_RTOL * np.linalg.norm(odv0)


def test_rollout_fit_16x32k_descends():
    """The full differentiable-rollout workload: 16 scenes x 32,768 particles,
    200 steps, checkpoints every 20, loss = per-scene MSE of final positions
    against rollouts with perturbed stiffness, viscosity and initial
    velocities.  A small step of the initial velocities against the adjoint
    gradient lowers the loss of the GPU forward rollout."""
    from simplefluid_amd import configs, engine
    rc = configs.rollout_scenes()
    sc = rc["scenes"]
    counts = [s["pos"].shape[0] for s in sc]
    pos0 = np.concatenate([s["pos"] for s in sc])
    t = engine.BatchEngine([s["target_params"] for s in sc], counts)
    t.set_particles(pos0, np.concatenate([s["target_vel"] for s in sc]))
    t.make_ready()
    t.advance_steps(rc["steps"])
    target = t.get_particles().astype(np.float64)
    t.close()
    b = engine.BatchEngine([s["params"] for s in sc], counts)
    off = b.scene_offsets()

    def loss_of(v0, grad=False):
        b.set_particles(pos0, v0)
        b.make_ready()
        b.rollout_forward(rc["steps"], rc["checkpoint_every"])
        xT = b.get_particles().astype(np.float64)
        L, g = 0.0, np.zeros_like(xT)
        for a, e in zip(off[:-1], off[1:]):
            d = xT[a:e] - target[a:e]
            s = 1.0 / (3 * (e - a))
            L += s * float((d * d).sum())
            g[a:e] = 2 * s * d
        if not grad:
            return L
        return L, b.rollout_backward(g.astype(np.float32))

    v0 = np.concatenate([s["vel"] for s in sc])
    L0, (dB, dmu, dv0) = loss_of(v0, grad=True)
    assert np.isfinite(dB).all() and np.isfinite(dmu).all() and np.isfinite(dv0).all()
    step = 0.01 / np.abs(dv0).max()          # at most 0.01 m/s per component
    L1 = loss_of((v0 - step * dv0).astype(np.float32))
    say(f"fit 16x32k: loss {L0:.6e} -> {L1:.6e} after one gradient step on v0")
    assert L1 < L0
    b.close()
