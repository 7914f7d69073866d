"""MI355X (gfx950) implementation of FlexRT's per-timestep radiative-transfer core.

The compute path is libflexrt_gpu.so (hand-written HIP kernels behind the C ABI
in include/flexrt_gpu.h).  ``rt.FlexRT`` is the Python host mirror of the
reference routines; ``inputs`` builds the synthetic workloads.
"""
from . import inputs  # noqa: F401

__all__ = ["inputs", "rt"]
