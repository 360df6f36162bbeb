"""Synthetic benchmark scenes (seeded numpy), the engine's standard workloads.

  dambreak_65k        32x64x32 = 65,536 particles on a 2r lattice (r = 0.005)
                      in the corner of the unit box, 1000 steps of 1e-4 s
  double_dambreak_1m  two 128x64x64 blocks = 1,048,576 particles with +-0.05r
                      uniform jitter; the unit box cannot hold them at this
                      spacing, so the box is [0,2] x [0,1] x [0,2]; 500 steps
  sphere_drop_4m      4,194,304 points sampled uniformly in a ball of radius
                      0.3 with minimum spacing 1.8r (seeded random sequential
                      addition) over a one-cell-thick resting layer; r is
                      scaled to 0.0014 so the ball holds that many at that
                      spacing; 300 steps of 5e-5 s
  ensemble_256        256 dam-break scenes, counts log-uniform in
                      [4096, 65536], stiffness U[20,100], viscosity
                      U[0.001, 0.05], block aspect U[0.5, 2]; own unit box each
  rollout_16x32k      16 dam-break scenes of 32,768 particles, 200 steps,
                      targets from perturbed stiffness / viscosity / v0

Kernel radius is 4r everywhere (Source/Controller.cpp:55 with the radius
rule of the reference's scenes); rest density 1000, Tait exponent 7,
gravity -9.81 y, restitution 0.1.
"""
from __future__ import annotations

import math

import numpy as np

from .engine import make_params

R = 0.005


def lattice_block(n3, origin, r=R, jitter=0.0, rng=None):
    """(nx*ny*nz, 3) float32 lattice at spacing 2r starting at origin + r."""
    nx, ny, nz = n3
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    pos = np.asarray(origin, np.float64) + r + 2.0 * r * ijk
    if jitter > 0:
        pos += rng.uniform(-jitter, jitter, size=pos.shape)
    return pos.astype(np.float32)


def dambreak_65k(jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    pos = lattice_block((32, 64, 32), (0.0, 0.0, 0.0), jitter=jitter, rng=rng)
    p = make_params(pressure_stiffness=50.0, viscosity=0.01, dt=1e-4)
    return dict(name="dambreak_65k", params=p, pos=pos, steps=1000)


def double_dambreak_1m(seed=1):
    rng = np.random.default_rng(seed)
    a = lattice_block((128, 64, 64), (0.0, 0.0, 0.0), jitter=0.05 * R, rng=rng)
    b = lattice_block((128, 64, 64), (2.0 - 1.28, 0.0, 2.0 - 0.64), jitter=0.05 * R, rng=rng)
    pos = np.concatenate([a, b])
    p = make_params(pressure_stiffness=50.0, viscosity=0.01, dt=1e-4,
                    box_min=(0.0, 0.0, 0.0), box_max=(2.0, 1.0, 2.0))
    return dict(name="double_dambreak_1m", params=p, pos=pos, steps=500)


_BALLS = {}


def sphere_drop_4m(seed=2, n_sphere=4_194_304, r=0.0014):
    """The sphere drop: n_sphere points sampled uniformly inside a ball of
    radius 0.3 at the box centre with a minimum spacing of 1.8r (seeded
    random sequential addition, sph_scene_sample_ball), over a resting layer
    one cell (h = 4r) thick: two 2r lattice layers covering the floor.  At
    r = 0.005 the ball holds only ~0.1M points at that spacing, so r is
    scaled to 0.0014 (h = 0.0056); the sample then fills 31 % of the ball's
    volume with exclusion spheres (the saturation of random sequential
    addition is 38 %), about 0.81 of rest density, with the clumps and gaps
    of a random packing."""
    from .engine import sample_ball
    key = (seed, n_sphere, r)
    if key not in _BALLS:
        _BALLS[key] = sample_ball(seed, (0.5, 0.5, 0.5), 0.3, 1.8 * r, n_sphere)
    sphere = _BALLS[key]
    sp = 2.0 * r
    g = int(1.0 / sp)
    i, k = np.meshgrid(np.arange(g), np.arange(g), indexing="ij")
    layer = []
    for j in range(2):  # one cell (h = 4r) thick = two lattice layers
        layer.append(np.stack([r + sp * i.ravel(), np.full(i.size, r + sp * j), r + sp * k.ravel()], 1))
    pos = np.concatenate([sphere] + [a.astype(np.float32) for a in layer]).astype(np.float32)
    p = make_params(particle_radius=r, kernel_radius=4 * r, pressure_stiffness=50.0,
                    viscosity=0.01, dt=5e-5)
    return dict(name="sphere_drop_4m", params=p, pos=pos, steps=300, n_sphere=n_sphere)


def slab_share(cfg, nshards=8, shard=None):
    """One GPU's share of a scene sharded by slabs along x: cuts at
    particle-count quantiles, so each slab owns N / nshards particles, plus a
    ghost band 2h wide on each inner side (ghosts recompute their own
    density, so the band must hold every neighbour of an owned particle's
    neighbours).  shard=None picks the slab with the most owned + ghost
    particles.  Returned as a standalone scene whose box spans the slab and
    its bands (2r wider, so no particle starts clamped), x extent only."""
    pos = cfg["pos"]
    p = cfg["params"]
    h = float(p.kernel_radius)
    r = float(p.particle_radius)
    bmin = [float(v) for v in p.box_min]
    bmax = [float(v) for v in p.box_max]
    n = pos.shape[0]
    xs = np.sort(pos[:, 0])
    cuts = [bmin[0]] + [float(xs[(n * s) // nshards]) for s in range(1, nshards)] + [math.inf]
    shares = []
    for s in range(nshards):
        x0, x1 = cuts[s], cuts[s + 1]
        own = (pos[:, 0] >= x0) & (pos[:, 0] < x1)
        band = (pos[:, 0] >= x0 - 2 * h) & (pos[:, 0] < x1 + 2 * h) & ~own
        shares.append((int(own.sum()) + int(band.sum()), s, x0, x1, own, band))
    if shard is None:
        _, shard, x0, x1, own, band = max(shares, key=lambda t: t[0])
    else:
        _, shard, x0, x1, own, band = shares[shard]
    # Owned particles first, then the ghosts.
    sel = np.concatenate([pos[own], pos[band]])
    lo = max(bmin[0], x0 - 2 * h - 2 * r)
    hi = min(bmax[0], x1 + 2 * h + 2 * r)
    q = type(p).from_buffer_copy(p)
    q.box_min[0], q.box_max[0] = lo, hi
    return dict(name=f"{cfg['name']}_slab{shard}of{nshards}", params=q, pos=np.ascontiguousarray(sel),
                owned=int(own.sum()), ghosts=int(band.sum()), shard=shard, x_range=(x0, min(x1, bmax[0])),
                steps=cfg["steps"])


def _block_dims(n, aspect):
    nx = max(1, int(round((n / aspect) ** (1.0 / 3.0))))
    ny = max(1, int(round(n / (nx * nx))))
    return nx, ny, nx


def ensemble(nscenes=256, seed=3, nmin=4096, nmax=65536, steps=1000):
    rng = np.random.default_rng(seed)
    scenes = []
    for s in range(nscenes):
        n = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        aspect = rng.uniform(0.5, 2.0)
        B = rng.uniform(20.0, 100.0)
        mu = rng.uniform(0.001, 0.05)
        nx, ny, nz = _block_dims(n, aspect)
        pos = lattice_block((nx, ny, nz), (0.0, 0.0, 0.0))
        p = make_params(pressure_stiffness=B, viscosity=mu, dt=1e-4)
        scenes.append(dict(params=p, pos=pos))
    return dict(name=f"ensemble_{nscenes}", scenes=scenes, steps=steps)


def rollout_scenes(nscenes=16, n3=(32, 32, 32), seed=4, steps=200, B=50.0, mu=0.01):
    """Nominal scenes plus the perturbed parameters of their target rollouts."""
    rng = np.random.default_rng(seed)
    scenes = []
    for s in range(nscenes):
        pos = lattice_block(n3, (0.0, 0.0, 0.0))
        n = pos.shape[0]
        scenes.append(dict(
            params=make_params(pressure_stiffness=B, viscosity=mu, dt=1e-4),
            pos=pos, vel=np.zeros((n, 3), np.float32),
            target_params=make_params(pressure_stiffness=B * rng.uniform(0.8, 1.2),
                                      viscosity=mu * rng.uniform(0.5, 2.0), dt=1e-4),
            target_vel=rng.normal(0.0, 0.1, size=(n, 3)).astype(np.float32)))
    return dict(name=f"rollout_{nscenes}x{n3[0] * n3[1] * n3[2]}", scenes=scenes, steps=steps,
                checkpoint_every=20)
