"""nbodygo_amd -- MI355X-native per-cycle compute path of aceeric/nbodygo.

The compute lives in libnbgpu.so (hand-written HIP for gfx950) behind the C-ABI declared in
include/nbody_gpu.h; this package is its Python driver:

  sim      Bodies (SoA body arrays) and NBodyGPU (one device context)
  runner   Body / BodyCollection / ComputationRunner: the reference's host-side API
           (cmd/body, cmd/runner) driving the C-ABI once per cycle
  gen      seeded initial conditions (uniform sphere, Plummer, two Plummer spheres, batches)
  csvio    the reference's body CSV format (cmd/sim/fromcsv.go)

Importing this package raises if libnbgpu.so has not been built: there is no CPU fallback.
"""
from . import _lib
from ._lib import NBodyError
from .sim import Bodies, NBodyBatch, NBodyGPU

__all__ = ["Bodies", "NBodyBatch", "NBodyGPU", "NBodyError", "_lib"]
