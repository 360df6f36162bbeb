"""GPU check of sph_step_noalloc, the step that allocates nothing, through
the C ABI: two calls in a row, the second on fewer particles (its scratch
then holds the first call's contents) and on a stream of its own, each equal
bit for bit to the first step of a handle given the same particles and
parameters; and the limits.  The caller's device arrays come from the HIP
runtime the engine library links (hipMalloc through ctypes)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _cases():
    from simplefluid_amd import configs, engine
    rng = np.random.default_rng(21)
    p0 = engine.default_params(24)
    x0 = engine.scene_generate(engine.SCENE_DOUBLE_DAMBREAK, 24)
    v0 = rng.normal(0.0, 0.2, x0.shape).astype(np.float32)
    cfg = configs.dambreak_65k(jitter=1e-6, seed=5)
    x1 = cfg["pos"][:8192].copy()
    v1 = rng.normal(0.0, 0.1, x1.shape).astype(np.float32)
    return [("DoubleDambreak res 24, GUI defaults", p0, x0, v0), ("dam-break corner, 8192", cfg["params"], x1, v1)]


class Hip:
    """The few HIP runtime calls the test needs, resolved through the engine
    library (its dependency), so they use the same runtime instance."""

    H2D, D2H = 1, 2

    def __init__(self, L):
        self.L = L
        sig = {"hipMalloc": [C.POINTER(C.c_void_p), C.c_size_t], "hipFree": [C.c_void_p],
               "hipMemcpy": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int], "hipDeviceSynchronize": [],
               "hipStreamCreateWithFlags": [C.POINTER(C.c_void_p), C.c_uint], "hipStreamDestroy": [C.c_void_p],
               "hipStreamSynchronize": [C.c_void_p]}
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args

    def ok(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.ok(self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D), "hipMemcpy")
        return p

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p

    def download(self, p, shape):
        out = np.empty(shape, np.float32)
        self.ok(self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, self.D2H), "hipMemcpy")
        return out


def test_noalloc_step_equals_the_handle_step_twice():
    from simplefluid_amd import engine
    L = engine.load_library()
    hip = Hip(L)
    own = C.c_void_p()
    hip.ok(L.hipStreamCreateWithFlags(C.byref(own), 1), "hipStreamCreateWithFlags")
    for (label, p, x, v), stream in zip(_cases(), (None, own)):
        n = x.shape[0]
        say(f"[noalloc] {label}: n={n}, {'default' if stream is None else 'own'} stream")
        xin, vin = hip.upload(x), hip.upload(v)
        xout, vout = hip.upload(np.full_like(x, np.nan)), hip.upload(np.full_like(v, np.nan))
        rc = L.sph_step_noalloc(C.byref(p), xin, vin, xout, vout, n, stream)
        assert rc == engine.SPH_OK, L.sph_last_error(None)
        hip.ok(L.hipDeviceSynchronize(), "hipDeviceSynchronize")
        e = engine.Engine(p)
        e.set_particles(x, v)
        e.make_ready()
        e.advance_steps(1)
        xg, vg = e.get_particles(), e.get_velocity()
        e.close()
        xa, va = hip.download(xout, x.shape), hip.download(vout, v.shape)
        for ptr in (xin, vin, xout, vout):
            hip.ok(L.hipFree(ptr), "hipFree")
        assert np.isfinite(xa).all() and np.isfinite(va).all()
        assert np.array_equal(xa, xg), f"{label}: positions differ, max {np.abs(xa - xg).max()}"
        assert np.array_equal(va, vg), f"{label}: velocities differ, max {np.abs(va - vg).max()}"
        assert not np.array_equal(xa, x)   # the step moved the particles
        say(f"[noalloc] {label}: equal to the handle's step")
    hip.ok(L.hipStreamDestroy(own), "hipStreamDestroy")


def test_noalloc_step_limits():
    from simplefluid_amd import engine
    L = engine.load_library()
    hip = Hip(L)
    p = engine.default_params(24)
    ptr = hip.alloc(65537 * 3 * 4)
    rc = L.sph_step_noalloc(C.byref(p), ptr, None, ptr, ptr, 65537, None)
    assert rc == engine.SPH_E_INVALID and b"65536" in L.sph_last_error(None)
    q = engine.clone_params(p, cfl_number=0.5)
    rc = L.sph_step_noalloc(C.byref(q), ptr, None, ptr, ptr, 100, None)
    assert rc == engine.SPH_E_INVALID and b"fixed dt" in L.sph_last_error(None)
    q = engine.clone_params(p, kernel_radius=0.004, particle_radius=0.001)   # 500^3 cells
    rc = L.sph_step_noalloc(C.byref(q), ptr, None, ptr, ptr, 100, None)
    assert rc == engine.SPH_E_INVALID and b"2^20" in L.sph_last_error(None)
    hip.ok(L.hipFree(ptr), "hipFree")
