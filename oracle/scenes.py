"""TEST INFRASTRUCTURE ONLY: float32 restatement of the reference's scene
generators, used to pin the engine's own generator (bit-exact positions) and
the screenshot particle counts.

Reference: Source/SceneManager.cpp
  * CubeDrop        :41-64   lattice over [-0.5, 0.5]^3
  * SphereDrop      :67-91   lattice over the sphere's bounding cube, reject
                             points with |p - c| > radius (glm::length, :83)
  * Dambreak        :94-119  block [-1+r, 0.4] x [-1+r, 0.4] x [-1+r, -0.5]
  * DoubleDambreak  :122-173 that block plus a mirrored block laid out from
                             its max corner (ppos = bMax - spacing*(i,j,k))
Loop order i (x) outer, j (y), k (z) inner, as in the reference; the grid
extents are float quotients truncated to int (Vec3<int>(float...)).
Parameter mapping: kernelRadius = 2/res (Source/Controller.cpp:55),
particleRadius = kernelRadius/4 (the only ratio that reproduces the particle
counts in Captured/1.png, 2.png and 3.png).
"""
from __future__ import annotations

import numpy as np

f32 = np.float32

SCENES = ("SphereDrop", "CubeDrop", "Dambreak", "DoubleDambreak")  # Include/Common.h:49-60


def radii(res: int):
    kr = f32(2.0) / f32(res)
    return kr, kr / f32(4.0)


def _lattice(bmin, spacing, grid, sign=1.0):
    bmin = np.asarray(bmin, f32)
    i, j, k = np.meshgrid(np.arange(grid[0]), np.arange(grid[1]), np.arange(grid[2]), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(f32)
    off = (spacing * ijk).astype(f32)
    return (bmin + off).astype(f32) if sign > 0 else (bmin - off).astype(f32)


def _grid(bmin, bmax, spacing):
    bmin = np.asarray(bmin, f32); bmax = np.asarray(bmax, f32)
    q = ((bmax - bmin) / spacing).astype(f32)
    return [int(x) for x in q]


def cube_drop(r):
    spacing = f32(2.0) * f32(r)
    bmin = np.array([-0.5] * 3, f32); bmax = np.array([0.5] * 3, f32)
    return _lattice(bmin, spacing, _grid(bmin, bmax, spacing))


def sphere_drop(r):
    spacing = f32(2.0) * f32(r)
    center = np.zeros(3, f32); radius = f32(0.5)
    bmin = center - radius
    g = int(f32(f32(2.0) * radius) / spacing)
    p = _lattice(bmin, spacing, [g, g, g])
    d = p - center
    ln = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(f32)
    return p[~(ln > radius)]


def dambreak(r):
    spacing = f32(2.0) * f32(r)
    bmin = np.array([-1.0] * 3, f32) + f32(r)
    bmax = np.array([0.4, 0.4, -0.5], f32)
    return _lattice(bmin, spacing, _grid(bmin, bmax, spacing))


def double_dambreak(r):
    r = f32(r)
    spacing = f32(2.0) * r
    a = dambreak(r)
    bmin = np.array([-0.4, -1.0, 0.5], f32) + np.array([0, r, 0], f32)
    bmax = np.array([1.0, 0.4, 1.0], f32) - np.array([r, 0, r], f32)
    b = _lattice(bmax, spacing, _grid(bmin, bmax, spacing), sign=-1.0)
    return np.concatenate([a, b])


def reference_scene(name: str, res: int = 24):
    _, r = radii(res)
    fn = {"SphereDrop": sphere_drop, "CubeDrop": cube_drop, "Dambreak": dambreak,
          "DoubleDambreak": double_dambreak}[name]
    return fn(r)
