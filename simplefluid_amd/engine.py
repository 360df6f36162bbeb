"""ctypes binding of libsph_engine.so (include/sph_engine.h).

The classes mirror the reference's solver surface:

* ``SPHParameters`` - the fields set in Source/Controller.cpp:52-64
  (kernelRadius, pressureStiffness, viscosity, boundaryRestitution,
  bUseAttractivePressure, stopTime, numThreads) plus particleRadius
  (Source/SceneManager.cpp:45) and the engine's formulation constants.
* ``SPHSolver`` - makeReady() (Source/Simulator.cpp:42), advanceFrame()
  (Source/Simulator.cpp:49) and the QtSPHSolver accessors getNumParticles(),
  getParticles(), getVelocity() (Include/QtSPHSolver.h:33-35).
* ``Engine`` / ``BatchEngine`` - the full C ABI (batched scenes, rollout
  adjoint, profiling, autotuning and the test-only views).

There is no CPU fallback: loading fails loudly when the HIP library is
missing, and every call raises ``SPHError`` on a negative status.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsph_engine.so")

SPH_OK = 0
SPH_E_INVALID = -1
SPH_E_OOM = -2
SPH_E_HIP = -3
SPH_E_NOT_READY = -4
SPH_E_DEVICE = -5
SPH_E_CAPACITY = -6

SCENE_SPHERE_DROP = 0
SCENE_CUBE_DROP = 1
SCENE_DAMBREAK = 2
SCENE_DOUBLE_DAMBREAK = 3
SCENE_NAMES = ("SphereDrop", "CubeDrop", "Dambreak", "DoubleDambreak")

PASS_NAMES = ("hash", "sort", "reorder", "density", "force")

# Every symbol include/sph_engine.h declares.
EXPORTS = (
    "sph_params_default", "sph_params_set_resolution", "sph_create", "sph_destroy",
    "sph_set_params", "sph_get_params", "sph_set_particles", "sph_make_ready",
    "sph_advance_frame", "sph_advance_steps", "sph_run_frame", "sph_num_particles",
    "sph_get_particles", "sph_get_velocity", "sph_get_density", "sph_export_frame", "sph_snapshot_begin", "sph_snapshot_wait", "sph_synchronize",
    "sph_last_error", "sph_set_use_graph", "sph_scene_generate", "sph_scene_sample_ball", "sph_batch_create",
    "sph_num_scenes", "sph_scene_offsets", "sph_set_scene_params", "sph_rollout_forward",
    "sph_rollout_backward", "sph_profile_enable", "sph_profile_read", "sph_autotune",
    "sph_get_tuning", "sph_set_tuning", "sph_debug_constants", "sph_debug_get_state",
    "sph_debug_set_state", "sph_debug_keep_acc", "sph_debug_step_outputs",
    "sph_debug_num_cells", "sph_debug_cells", "sph_debug_neighbors", "sph_set_sort_mode",
    "sph_get_sort_mode", "sph_set_cell_subdivision", "sph_set_device_memory_limit",
    "sph_device_memory_used", "sph_step_noalloc",
)
SORT_CELL_COUNT = 0
SORT_RADIX = 1


class SPHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sph error {code}: {msg}")
        self.code = code


class sph_params(C.Structure):
    _fields_ = [
        ("particle_radius", C.c_float),
        ("kernel_radius", C.c_float),
        ("rest_density", C.c_float),
        ("pressure_stiffness", C.c_float),
        ("viscosity", C.c_float),
        ("boundary_restitution", C.c_float),
        ("use_attractive_pressure", C.c_int),
        ("attractive_pressure_ratio", C.c_float),
        ("gravity", C.c_float * 3),
        ("dt", C.c_float),
        ("box_min", C.c_float * 3),
        ("box_max", C.c_float * 3),
        ("stop_time", C.c_float),
        ("num_threads", C.c_int),
        ("cfl_number", C.c_float),
    ]


class sph_scene_desc(C.Structure):
    _fields_ = [("params", sph_params), ("num_particles", C.c_uint32)]


_LIB = None


def load_library(path: str = LIB_PATH):
    """Load the HIP engine.  Raises when the library has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m simplefluid_amd.build`")
    L = C.CDLL(path)
    vp, u32, i32, f32p = C.c_void_p, C.c_uint32, C.c_int, C.c_void_p
    P = C.POINTER
    H = C.c_void_p
    sig = {
        "sph_params_default": (i32, [P(sph_params)]),
        "sph_params_set_resolution": (i32, [P(sph_params), i32]),
        "sph_create": (i32, [P(sph_params), P(H)]),
        "sph_destroy": (None, [H]),
        "sph_set_params": (i32, [H, P(sph_params)]),
        "sph_get_params": (i32, [H, P(sph_params)]),
        "sph_set_particles": (i32, [H, f32p, f32p, u32]),
        "sph_make_ready": (i32, [H]),
        "sph_advance_frame": (i32, [H, P(C.c_float)]),
        "sph_advance_steps": (i32, [H, u32]),
        "sph_run_frame": (i32, [H, P(C.c_float), P(u32)]),
        "sph_num_particles": (u32, [H]),
        "sph_get_particles": (i32, [H, f32p]),
        "sph_get_velocity": (i32, [H, f32p]),
        "sph_get_density": (i32, [H, f32p]),
        "sph_snapshot_begin": (i32, [H]),
        "sph_snapshot_wait": (i32, [H, f32p]),
        "sph_export_frame": (i32, [H, C.c_char_p]),
        "sph_synchronize": (i32, [H]),
        "sph_last_error": (C.c_char_p, [H]),
        "sph_set_use_graph": (i32, [H, i32]),
        "sph_scene_generate": (i32, [i32, i32, f32p, u32, P(u32)]),
        "sph_scene_sample_ball": (i32, [C.c_uint64, f32p, C.c_float, C.c_float, u32, f32p, P(u32)]),
        "sph_batch_create": (i32, [P(sph_scene_desc), u32, P(H)]),
        "sph_num_scenes": (u32, [H]),
        "sph_scene_offsets": (i32, [H, vp]),
        "sph_set_scene_params": (i32, [H, u32, P(sph_params)]),
        "sph_rollout_forward": (i32, [H, u32, u32]),
        "sph_rollout_backward": (i32, [H, f32p, f32p, f32p, f32p]),
        "sph_profile_enable": (i32, [H, i32]),
        "sph_profile_read": (i32, [H, vp, P(u32)]),
        "sph_autotune": (i32, [H]),
        "sph_get_tuning": (i32, [H, vp]),
        "sph_set_tuning": (i32, [H, vp]),
        "sph_debug_constants": (i32, [H, u32, f32p, vp]),
        "sph_debug_get_state": (i32, [H, f32p, f32p, vp]),
        "sph_debug_set_state": (i32, [H, f32p, f32p, vp]),
        "sph_debug_keep_acc": (i32, [H, i32]),
        "sph_debug_step_outputs": (i32, [H, vp, vp, vp, f32p, f32p]),
        "sph_debug_num_cells": (u32, [H]),
        "sph_debug_cells": (i32, [H, vp, vp]),
        "sph_debug_neighbors": (i32, [H, vp, vp, C.c_uint64, P(C.c_uint64)]),
        "sph_set_sort_mode": (i32, [H, i32]),
        "sph_get_sort_mode": (i32, [H]),
        "sph_set_cell_subdivision": (i32, [H, i32]),
        "sph_set_device_memory_limit": (i32, [H, C.c_uint64]),
        "sph_device_memory_used": (i32, [H, P(C.c_uint64)]),
        "sph_step_noalloc": (i32, [P(sph_params), vp, vp, vp, vp, u32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f3(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 3)
    if n is not None and a.shape[0] != n:
        raise ValueError(f"expected {n} rows, got {a.shape[0]}")
    return a


def default_params(resolution: int = 24, **overrides) -> sph_params:
    """Reference defaults (res 24, stiffness 50000, viscosity 0.05, ...)."""
    L = load_library()
    p = sph_params()
    L.sph_params_default(C.byref(p))
    L.sph_params_set_resolution(C.byref(p), int(resolution))
    return set_fields(p, **overrides)


def set_fields(p: sph_params, **kw) -> sph_params:
    for k, v in kw.items():
        if k in ("gravity", "box_min", "box_max"):
            getattr(p, k)[:] = [float(x) for x in v]
        elif not hasattr(p, k):
            raise KeyError(k)
        else:
            setattr(p, k, v)
    return p


def clone_params(p: sph_params, **overrides) -> sph_params:
    q = sph_params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(sph_params))
    return set_fields(q, **overrides)


def make_params(**kw) -> sph_params:
    """Params of the synthetic benchmark scenes: r 0.005, h 4r, unit box."""
    p = sph_params()
    base = dict(particle_radius=0.005, kernel_radius=0.02, rest_density=1000.0,
                pressure_stiffness=50.0, viscosity=0.01, boundary_restitution=0.1,
                use_attractive_pressure=0, attractive_pressure_ratio=0.1,
                gravity=(0.0, -9.81, 0.0), dt=1e-4, box_min=(0.0, 0.0, 0.0),
                box_max=(1.0, 1.0, 1.0), stop_time=5.0, num_threads=0, cfl_number=0.0)
    base.update(kw)
    return set_fields(p, **base)


def load_frame(path: str):
    """(positions, velocities) of a frame written by export_frame."""
    a = np.load(path, allow_pickle=False)
    if a.ndim != 2 or a.shape[1] != 6 or a.dtype != np.float32:
        raise ValueError(f"{path}: not an engine frame (shape {a.shape}, {a.dtype})")
    return np.ascontiguousarray(a[:, :3]), np.ascontiguousarray(a[:, 3:])


def scene_generate(scene: int, resolution: int = 24) -> np.ndarray:
    """Reference scene lattice (Source/SceneManager.cpp) as (n, 3) float32."""
    L = load_library()
    n = C.c_uint32(0)
    rc = L.sph_scene_generate(int(scene), int(resolution), None, 0, C.byref(n))
    if rc != SPH_OK:
        raise SPHError(rc, "scene_generate")
    out = np.empty((n.value, 3), np.float32)
    rc = L.sph_scene_generate(int(scene), int(resolution), _p(out), n.value, C.byref(n))
    if rc != SPH_OK:
        raise SPHError(rc, "scene_generate")
    return out


def sample_ball(seed: int, centre, radius: float, min_spacing: float, count: int) -> np.ndarray:
    """`count` seeded uniform points in a ball, no two closer than
    min_spacing (random sequential addition), as (count, 3) float32."""
    L = load_library()
    c = np.ascontiguousarray(centre, np.float32).reshape(3)
    out = np.empty((int(count), 3), np.float32)
    n = C.c_uint32(0)
    rc = L.sph_scene_sample_ball(C.c_uint64(int(seed)), _p(c), float(radius), float(min_spacing), int(count),
                                 _p(out), C.byref(n))
    if rc != SPH_OK:
        raise SPHError(rc, f"sample_ball: placed {n.value} of {count} points (the ball is too full at this spacing)")
    return out


class Engine:
    """One handle of the engine (a single scene unless built by BatchEngine)."""

    def __init__(self, params: sph_params | None = None, _handle=None):
        self._L = load_library()
        if _handle is not None:
            self._h = _handle
        else:
            h = C.c_void_p()
            rc = self._L.sph_create(C.byref(params if params is not None else make_params()), C.byref(h))
            if rc != SPH_OK:
                raise SPHError(rc, "sph_create: " + self._L.sph_last_error(None).decode())
            self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int):
        if rc != SPH_OK:
            msg = self._L.sph_last_error(self._h)
            raise SPHError(rc, msg.decode() if msg else "")

    # -- reference-shaped surface --------------------------------------------
    def set_params(self, p: sph_params):
        self._chk(self._L.sph_set_params(self._h, C.byref(p)))

    def get_params(self) -> sph_params:
        p = sph_params()
        self._chk(self._L.sph_get_params(self._h, C.byref(p)))
        return p

    def set_particles(self, xyz, vel=None):
        xyz = _f3(xyz)
        n = xyz.shape[0]
        vel = None if vel is None else _f3(vel, n)
        self._chk(self._L.sph_set_particles(self._h, _p(xyz), _p(vel), n))

    def make_ready(self):
        self._chk(self._L.sph_make_ready(self._h))

    def advance_frame(self) -> float:
        dt = C.c_float(0)
        self._chk(self._L.sph_advance_frame(self._h, C.byref(dt)))
        return dt.value

    def advance_steps(self, k: int):
        self._chk(self._L.sph_advance_steps(self._h, int(k)))

    def run_frame(self):
        t = C.c_float(0)
        s = C.c_uint32(0)
        self._chk(self._L.sph_run_frame(self._h, C.byref(t), C.byref(s)))
        return t.value, s.value

    @property
    def num_particles(self) -> int:
        return int(self._L.sph_num_particles(self._h))

    def get_particles(self) -> np.ndarray:
        out = np.empty((self.num_particles, 3), np.float32)
        self._chk(self._L.sph_get_particles(self._h, _p(out)))
        return out

    def get_velocity(self) -> np.ndarray:
        out = np.empty((self.num_particles, 3), np.float32)
        self._chk(self._L.sph_get_velocity(self._h, _p(out)))
        return out

    def snapshot_begin(self):
        """Capture the positions now; the copy to the host overlaps later steps."""
        self._chk(self._L.sph_snapshot_begin(self._h))

    def snapshot_wait(self) -> np.ndarray:
        """Positions (original order) captured by the last snapshot_begin."""
        out = np.empty((self.num_particles, 3), np.float32)
        self._chk(self._L.sph_snapshot_wait(self._h, _p(out)))
        return out

    def export_frame(self, path: str):
        """Writes the state as a (n, 6) float32 .npy: x, y, z, vx, vy, vz in original order."""
        self._chk(self._L.sph_export_frame(self._h, os.fsencode(path)))

    def get_density(self) -> np.ndarray:
        out = np.empty(self.num_particles, np.float32)
        self._chk(self._L.sph_get_density(self._h, _p(out)))
        return out

    def synchronize(self):
        self._chk(self._L.sph_synchronize(self._h))

    def set_use_graph(self, on: bool):
        self._chk(self._L.sph_set_use_graph(self._h, 1 if on else 0))

    # -- device memory ----------------------------------------------------------
    def set_device_memory_limit(self, nbytes: int):
        """Cap on the device bytes this handle's buffers may hold (0: none)."""
        self._chk(self._L.sph_set_device_memory_limit(self._h, int(nbytes)))

    def device_memory_used(self) -> int:
        v = C.c_uint64(0)
        self._chk(self._L.sph_device_memory_used(self._h, C.byref(v)))
        return int(v.value)

    # -- measurement ----------------------------------------------------------
    def profile_enable(self, on: bool = True):
        self._chk(self._L.sph_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        ms = np.zeros(5, np.float64)
        steps = C.c_uint32(0)
        self._chk(self._L.sph_profile_read(self._h, _p(ms), C.byref(steps)))
        return dict(zip(PASS_NAMES, ms.tolist())), steps.value

    def autotune(self):
        self._chk(self._L.sph_autotune(self._h))

    def get_tuning(self):
        cfg = np.zeros(5, np.int32)
        self._chk(self._L.sph_get_tuning(self._h, _p(cfg)))
        return dict(density_block=int(cfg[0]), force_block=int(cfg[1]),
                    reorder_block=int(cfg[2]), sort_items=int(cfg[3]), persistent=int(cfg[4]) & 1,
                    force_staged=(int(cfg[4]) >> 1) & 1, density_staged=(int(cfg[4]) >> 2) & 1)

    def set_tuning(self, density_block=256, force_block=256, reorder_block=256, sort_items=8, persistent=0,
                   force_staged=0, density_staged=0):
        """persistent: density and force as persistent blocks over tile
        queues; force_staged / density_staged: that pass reads each row's
        candidates from LDS.  Every choice gives the same results bit for
        bit."""
        cfg = np.array([density_block, force_block, reorder_block, sort_items,
                        (persistent & 1) | ((force_staged & 1) << 1) | ((density_staged & 1) << 2)], np.int32)
        self._chk(self._L.sph_set_tuning(self._h, _p(cfg)))

    def set_sort_mode(self, mode: int):
        self._chk(self._L.sph_set_sort_mode(self._h, int(mode)))

    def get_sort_mode(self) -> int:
        return int(self._L.sph_get_sort_mode(self._h))

    def set_cell_subdivision(self, xsub: int):
        self._chk(self._L.sph_set_cell_subdivision(self._h, int(xsub)))

    # -- rollout adjoint ------------------------------------------------------
    def rollout_forward(self, steps: int, checkpoint_every: int = 20):
        self._chk(self._L.sph_rollout_forward(self._h, int(steps), int(checkpoint_every)))

    def rollout_backward(self, dL_dx_final):
        n = self.num_particles
        g = _f3(dL_dx_final, n)
        ns = self.num_scenes
        dB = np.zeros(ns, np.float32)
        dmu = np.zeros(ns, np.float32)
        dv0 = np.zeros((n, 3), np.float32)
        self._chk(self._L.sph_rollout_backward(self._h, _p(g), _p(dB), _p(dmu), _p(dv0)))
        return dB, dmu, dv0

    @property
    def num_scenes(self) -> int:
        return int(self._L.sph_num_scenes(self._h))

    # -- test-only views ------------------------------------------------------
    def debug_constants(self, scene: int = 0):
        out = np.zeros(14, np.float32)
        G = np.zeros(4, np.int32)
        self._chk(self._L.sph_debug_constants(self._h, int(scene), _p(out), _p(G)))
        names = ["h", "h2", "inv_cs", "mass", "cD", "cP", "cV", "inv_rho0"]
        d = {k: out[i] for i, k in enumerate(names)}
        d["lo"] = out[8:11].copy()
        d["hi"] = out[11:14].copy()
        d["G"] = G[:3].copy()
        d["xsub"] = int(G[3])
        return d

    def debug_get_state(self):
        n = self.num_particles
        p = np.empty((n, 3), np.float32)
        v = np.empty((n, 3), np.float32)
        i = np.empty(n, np.uint32)
        self._chk(self._L.sph_debug_get_state(self._h, _p(p), _p(v), _p(i)))
        return p, v, i

    def debug_set_state(self, xyz, vel, ids):
        n = self.num_particles
        xyz = _f3(xyz, n)
        vel = _f3(vel, n)
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        self._chk(self._L.sph_debug_set_state(self._h, _p(xyz), _p(vel), _p(ids)))

    def debug_keep_acc(self, on: bool = True):
        self._chk(self._L.sph_debug_keep_acc(self._h, 1 if on else 0))

    def debug_step_outputs(self, acc: bool = False):
        n = self.num_particles
        key = np.empty(n, np.uint32)
        perm = np.empty(n, np.uint32)
        skey = np.empty(n, np.uint32)
        rho = np.empty(n, np.float32)
        a = np.empty((n, 3), np.float32) if acc else None
        self._chk(self._L.sph_debug_step_outputs(self._h, _p(key), _p(perm), _p(skey), _p(rho), _p(a)))
        d = dict(key=key, perm=perm, sorted_key=skey, rho=rho)
        if acc:
            d["acc"] = a
        return d

    def debug_cells(self):
        c = int(self._L.sph_debug_num_cells(self._h))
        s = np.empty(c, np.uint32)
        e = np.empty(c, np.uint32)
        self._chk(self._L.sph_debug_cells(self._h, _p(s), _p(e)))
        return s, e

    def debug_neighbors(self):
        n = self.num_particles
        off = np.empty(n + 1, np.uint32)
        cnt = C.c_uint64(0)
        self._chk(self._L.sph_debug_neighbors(self._h, _p(off), None, 0, C.byref(cnt)))
        lst = np.empty(max(cnt.value, 1), np.uint32)
        self._chk(self._L.sph_debug_neighbors(self._h, _p(off), _p(lst), cnt.value, C.byref(cnt)))
        return off, lst[:cnt.value]


class BatchEngine(Engine):
    """Many independent scenes stepped together (one launch per pass)."""

    def __init__(self, params_list, counts):
        L = load_library()
        descs = (sph_scene_desc * len(params_list))()
        for i, (p, n) in enumerate(zip(params_list, counts)):
            descs[i].params = p
            descs[i].num_particles = int(n)
        h = C.c_void_p()
        rc = L.sph_batch_create(descs, len(params_list), C.byref(h))
        if rc != SPH_OK:
            raise SPHError(rc, "sph_batch_create: " + L.sph_last_error(None).decode())
        super().__init__(_handle=h)

    def scene_offsets(self) -> np.ndarray:
        off = np.zeros(self.num_scenes + 1, np.uint32)
        self._chk(self._L.sph_scene_offsets(self._h, _p(off)))
        return off

    def set_scene_params(self, scene: int, p: sph_params):
        self._chk(self._L.sph_set_scene_params(self._h, int(scene), C.byref(p)))


class SPHSolver:
    """QtSPHSolver-shaped wrapper (Include/QtSPHSolver.h:27-36) over Engine.

    Method names follow the reference: makeReady (Source/Simulator.cpp:42),
    advanceFrame (:49), getNumParticles / getParticles / getVelocity
    (Include/QtSPHSolver.h:33-35).  setupScene mirrors
    Source/Simulator.cpp:85-103 with the engine's scene generator.
    """

    def __init__(self, params: sph_params):
        self.params = params
        self.engine = Engine(params)
        self._dirty = True

    def setupScene(self, scene: int, resolution: int | None = None):
        if resolution is not None:
            L = load_library()
            L.sph_params_set_resolution(C.byref(self.params), int(resolution))
            self.engine.set_params(self.params)
        res = int(round(2.0 / self.params.kernel_radius))
        self.engine.set_particles(scene_generate(scene, res))

    def setParticles(self, xyz, vel=None):
        self.engine.set_particles(xyz, vel)

    def makeReady(self):
        """Uploads and builds the grid, then picks block sizes for this
        handle's size (sph_autotune; cached per process, results unchanged).
        A failed autotune leaves the state as it was and keeps the defaults."""
        self.engine.make_ready()
        try:
            self.engine.autotune()
        except SPHError:
            pass

    def advanceFrame(self) -> float:
        return self.engine.advance_frame()

    def getNumParticles(self) -> int:
        return self.engine.num_particles

    def getParticles(self) -> np.ndarray:
        return self.engine.get_particles()

    def getVelocity(self) -> np.ndarray:
        return self.engine.get_velocity()
