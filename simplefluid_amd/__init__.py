"""MI355X (gfx950) SPH step engine: a drop-in for SimpleFluid's CPU solver.

The compute lives in ``libsph_engine.so`` (hand-written HIP, C ABI in
``include/sph_engine.h``); this package holds its ctypes binding, the build
script and the synthetic benchmark scenes.
"""
from .engine import (  # noqa: F401
    BatchEngine, Engine, SPHError, SPHSolver, default_params, load_library, make_params,
    scene_generate, sph_params,
)
