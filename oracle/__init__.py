"""TEST INFRASTRUCTURE ONLY: the CPU oracle for the SPH step.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg use this package, and only as the checker.  The engine never imports it.

``liboracle.so`` is the C++ restatement in ``sph_oracle.hpp`` (forward step)
and ``sph_adjoint_oracle.hpp`` (reverse step); ``scenes.py`` restates the
reference's scene generators (Source/SceneManager.cpp:21-173) in float32.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Params(C.Structure):
    """Mirror of orc::Params (same field meaning as the engine's sph_params)."""
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
        ("x_subdivision", C.c_int),
        ("cfl_number", C.c_float),
    ]


def make_params(**kw) -> Params:
    p = Params()
    d = dict(particle_radius=0.005, kernel_radius=0.02, rest_density=1000.0,
             pressure_stiffness=50.0, viscosity=0.01, boundary_restitution=0.1,
             use_attractive_pressure=0, attractive_pressure_ratio=0.1,
             gravity=(0.0, -9.81, 0.0), dt=1e-4, box_min=(0.0, 0.0, 0.0),
             box_max=(1.0, 1.0, 1.0), x_subdivision=1, cfl_number=0.0)
    d.update(kw)
    for k, v in d.items():
        if k in ("gravity", "box_min", "box_max"):
            getattr(p, k)[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


def from_params(p, x_subdivision: int = 1) -> Params:
    """Oracle params from any object with the same field names (e.g. the
    engine's sph_params) plus the grid's x subdivision."""
    kw = {}
    for f, _ in Params._fields_:
        if f == "x_subdivision":
            continue
        v = getattr(p, f)
        kw[f] = list(v) if f in ("gravity", "box_min", "box_max") else v
    return make_params(x_subdivision=x_subdivision, **kw)


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        vp, u32, dbl = C.c_void_p, C.c_uint32, C.c_double
        P = C.POINTER
        L.orc_new.restype = vp
        L.orc_new.argtypes = [C.c_int, P(Params)]
        L.orc_free.argtypes = [vp]
        L.orc_set_threads.argtypes = [vp, C.c_int]
        L.orc_record_neighbors.argtypes = [vp, C.c_int]
        L.orc_set_state.argtypes = [vp, u32, vp, vp, vp]
        L.orc_step.restype = dbl
        L.orc_step.argtypes = [vp, C.c_int]
        L.orc_num_particles.restype = u32
        L.orc_num_particles.argtypes = [vp]
        L.orc_num_cells.restype = u32
        L.orc_num_cells.argtypes = [vp]
        L.orc_get_state.argtypes = [vp, vp, vp, vp]
        L.orc_last_dt.restype = dbl
        L.orc_last_dt.argtypes = [vp]
        L.orc_get_constants.argtypes = [vp, vp, vp]
        L.orc_get_step_debug.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_get_cells.argtypes = [vp, vp, vp]
        L.orc_num_neighbor_entries.restype = C.c_uint64
        L.orc_num_neighbor_entries.argtypes = [vp]
        L.orc_get_neighbors.argtypes = [vp, vp, vp]
        L.orc_bruteforce_neighbors.restype = C.c_uint64
        L.orc_bruteforce_neighbors.argtypes = [u32, vp, C.c_float, vp, vp]
        L.orc_cell_keys.argtypes = [vp, u32, vp, vp]
        L.orc_rollout_loss_grad.restype = dbl
        L.orc_rollout_loss_grad.argtypes = [C.c_int, P(Params), u32, vp, vp, vp, C.c_int,
                                            dbl, vp, vp, vp, C.c_int, dbl, dbl]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One scene stepped on the CPU, float32 (fp64=False) or float64."""

    def __init__(self, params: Params, fp64: bool = False, threads: int = 1):
        self._L = lib()
        self.params = params
        self.fp64 = fp64
        self._h = self._L.orc_new(1 if fp64 else 0, C.byref(params))
        self._L.orc_set_threads(self._h, int(threads))

    def __del__(self):
        try:
            self._L.orc_free(self._h)
        except Exception:
            pass

    def record_neighbors(self, on: bool = True):
        self._L.orc_record_neighbors(self._h, 1 if on else 0)

    def set_state(self, pos, vel=None, ids=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        self._L.orc_set_state(self._h, n, _ptr(pos), _ptr(vel), _ptr(ids))

    def step(self, k: int = 1) -> float:
        return self._L.orc_step(self._h, int(k))

    def last_dt(self) -> float:
        """Step length of the last step (dt, or the CFL step)."""
        return self._L.orc_last_dt(self._h)

    @property
    def n(self) -> int:
        return self._L.orc_num_particles(self._h)

    @property
    def ncells(self) -> int:
        return self._L.orc_num_cells(self._h)

    def state(self):
        n = self.n
        p = np.empty((n, 3)); v = np.empty((n, 3)); i = np.empty(n, np.uint32)
        self._L.orc_get_state(self._h, _ptr(p), _ptr(v), _ptr(i))
        return p, v, i

    def state_by_id(self):
        p, v, i = self.state()
        P = np.empty_like(p); V = np.empty_like(v)
        P[i] = p; V[i] = v
        return P, V

    def constants(self):
        out = np.empty(14); G = np.empty(4, np.int32)
        self._L.orc_get_constants(self._h, _ptr(out), _ptr(G))
        names = ["h", "h2", "inv_cs", "mass", "cD", "cP", "cV", "inv_rho0"]
        d = {k: out[i] for i, k in enumerate(names)}
        d["lo"] = out[8:11].copy(); d["hi"] = out[11:14].copy(); d["G"] = G[:3].copy()
        d["xsub"] = int(G[3])
        return d

    def step_debug(self):
        n = self.n
        key = np.empty(n, np.uint32); perm = np.empty(n, np.uint32); skey = np.empty(n, np.uint32)
        rho = np.empty(n); acc = np.empty((n, 3)); sid = np.empty(n, np.uint32)
        self._L.orc_get_step_debug(self._h, _ptr(key), _ptr(perm), _ptr(skey), _ptr(rho),
                                   _ptr(acc), _ptr(sid))
        return dict(key=key, perm=perm, sorted_key=skey, rho=rho, acc=acc, sorted_id=sid)

    def cells(self):
        c = self.ncells
        s = np.empty(c, np.uint32); e = np.empty(c, np.uint32)
        self._L.orc_get_cells(self._h, _ptr(s), _ptr(e))
        return s, e

    def neighbors(self):
        n = self.n
        m = self._L.orc_num_neighbor_entries(self._h)
        off = np.empty(n + 1, np.uint32); lst = np.empty(max(int(m), 1), np.uint32)
        self._L.orc_get_neighbors(self._h, _ptr(off), _ptr(lst))
        return off, lst[:m]

    def cell_keys(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        k = np.empty(pos.shape[0], np.uint32)
        self._L.orc_cell_keys(self._h, pos.shape[0], _ptr(pos), _ptr(k))
        return k


def bruteforce_neighbors(pos, h):
    """All-pairs neighbour CSR with the engine's float predicate (self included)."""
    L = lib()
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    n = pos.shape[0]
    h2 = np.float32(h) * np.float32(h)
    m = L.orc_bruteforce_neighbors(n, _ptr(pos), float(h2), None, None)
    off = np.empty(n + 1, np.uint32); lst = np.empty(max(int(m), 1), np.uint32)
    L.orc_bruteforce_neighbors(n, _ptr(pos), float(h2), _ptr(off), _ptr(lst))
    return off, lst[:m]


def rollout_loss_grad(params: Params, pos0, vel0, target, steps: int, fp64: bool = True,
                      grad: bool = True, loss_scale: float | None = None, threads: int = 1,
                      B: float | None = None, mu: float | None = None):
    """Loss (and dL/dB, dL/dmu, dL/dv0) of an n-particle rollout on the CPU.

    B / mu, when given, override the float parameters with double values
    (for central differences below float resolution)."""
    L = lib()
    pos0 = np.ascontiguousarray(pos0, dtype=np.float64).reshape(-1, 3)
    n = pos0.shape[0]
    vel0 = np.ascontiguousarray(vel0, dtype=np.float64).reshape(-1, 3)
    target = np.ascontiguousarray(target, dtype=np.float64).reshape(-1, 3)
    scale = 1.0 / (3 * n) if loss_scale is None else float(loss_scale)
    dB = np.zeros(1); dmu = np.zeros(1); dv0 = np.zeros((n, 3))
    loss = L.orc_rollout_loss_grad(1 if fp64 else 0, C.byref(params), n, _ptr(pos0), _ptr(vel0),
                                   _ptr(target), int(steps), scale,
                                   _ptr(dB) if grad else None, _ptr(dmu) if grad else None,
                                   _ptr(dv0) if grad else None, int(threads),
                                   float("nan") if B is None else float(B),
                                   float("nan") if mu is None else float(mu))
    if not grad:
        return loss
    return loss, float(dB[0]), float(dmu[0]), dv0
