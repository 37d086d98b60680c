"""zeno_amd -- MI355X-native (gfx950) MLS/APIC-MPM substep for Zeno's CuLagrange nodes.

The compute path is libzmpm.so (hand-written HIP kernels behind the C ABI of
include/zmpm.h).  This package holds its Python binding (ctypes, the
counterpart of the PyZpc `zpy` handle wrapper, projects/PyZpc/Lib/zpy) and the
host-side scene builders used by the benchmark and tests.
"""
from .zmpm import (  # noqa: F401
    ZmpmError, Context, Material, StepParams, load_library,
    MODEL_FCR, MODEL_JFLUID, BOUNDARY_NONE, BOUNDARY_STICKY_BOX,
    LAYOUT_SOA, LAYOUT_AOS, LAYOUT_AOSOA32,
)
