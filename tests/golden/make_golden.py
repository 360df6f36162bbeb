"""Writes the golden fixtures of tests/:
  tiny_step.npz  one FP64 oracle step of a small jittered lattice (inputs,
                 permutation, neighbour CSR, density, acceleration);
  frame_cubedrop16_step40.npy  the reference CubeDrop at resolution 16 with
                 the GUI defaults after 40 FP64 oracle steps, in the engine's
                 frame-export format (float32 (n, 6): x, y, z, vx, vy, vz);
  wall_step.npz  one FP64 oracle step of a jittered lattice in the box corner
                 (three faces within h: the walls' density and pressure terms
                 act), with inputs, permutation, density and acceleration.

The reference (ttnghia/SimpleFluid) holds no vectors for its solver, which is
not vendored; these fixtures freeze the restatement once it passed the
closed-form, brute-force and differentiation checks in tests/test_oracle.py.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def main():
    rng = np.random.default_rng(2024)
    r = 0.005
    i, j, k = np.meshgrid(np.arange(7), np.arange(9), np.arange(6), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    pos = (0.2 + r + 2 * r * ijk + rng.uniform(-0.3 * r, 0.3 * r, ijk.shape)).astype(np.float32).astype(np.float64)
    vel = rng.normal(0, 0.3, pos.shape).astype(np.float32).astype(np.float64)
    B, mu = 400.0, 0.05
    o = oracle.Oracle(oracle.make_params(pressure_stiffness=B, viscosity=mu), fp64=True)
    o.record_neighbors(True)
    o.set_state(pos, vel)
    o.step()
    d = o.step_debug()
    off, lst = o.neighbors()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_step.npz")
    np.savez(out, pos=pos, vel=vel, B=B, mu=mu, perm=d["perm"], nbr_off=off, nbr_list=lst,
             rho=d["rho"], acc=d["acc"])
    print("wrote", out, len(pos), "particles,", len(lst), "neighbour entries")

    # Frame fixture: host-only calls of the engine library (parameters and the
    # scene generator need no GPU), stepped by the oracle.
    from simplefluid_amd import engine
    p = engine.default_params(16)
    x0 = engine.scene_generate(engine.SCENE_CUBE_DROP, 16)
    o = oracle.Oracle(oracle.from_params(p, 4), fp64=True, threads=8)
    o.set_state(x0.astype(np.float64))
    o.step(40)
    x, v = o.state_by_id()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "frame_cubedrop16_step40.npy")
    np.save(out, np.concatenate([x, v], 1).astype(np.float32))
    print("wrote", out, x.shape[0], "particles")
    wall_step()


def wall_step():
    rng = np.random.default_rng(7)
    r = 0.005
    i, j, k = np.meshgrid(np.arange(6), np.arange(7), np.arange(5), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    pos = (r + 2 * r * ijk + rng.uniform(-0.2 * r, 0.2 * r, ijk.shape)).astype(np.float32).astype(np.float64)
    pos = np.maximum(pos, r)  # inside the clamp bounds
    vel = rng.normal(0, 0.2, pos.shape).astype(np.float32).astype(np.float64)
    B, mu = 400.0, 0.05
    o = oracle.Oracle(oracle.make_params(pressure_stiffness=B, viscosity=mu), fp64=True)
    o.set_state(pos, vel)
    o.step()
    d = o.step_debug()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wall_step.npz")
    np.savez(out, pos=pos, vel=vel, B=B, mu=mu, perm=d["perm"], rho=d["rho"], acc=d["acc"])
    print("wrote", out, len(pos), "particles")


if __name__ == "__main__":
    main()
