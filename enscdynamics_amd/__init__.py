"""EnSC-MI355X: an explicit-dynamics step engine for C3D8R hexahedra on
AMD MI355X (gfx950), drop-in for the central-difference step loop of
ssadz/EnSCDynamics' exDyna3D (Explicit3D/source/exDyna3D.cpp:54-102).

The compute path is libensc_hip.so (HIP kernels behind the C ABI declared in
include/ensc_engine.h); this package is the host-side Python mirror used by
tests and the benchmark.
"""
from .model import Bc, Model, Step  # noqa: F401

__all__ = ["Bc", "Model", "Step"]
