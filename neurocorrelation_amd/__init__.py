"""MI355X-native simulation step of NeuroCorrelation.

The product is a C ABI (include/neucor_amd.h) over hand-written HIP kernels
for gfx950, built in-tree into libneucor_amd.so.  This package adds a thin
Python view for tests, the benchmark, the batched form and the serving loop.
"""
from ._lib import NCError, lib  # noqa: F401
from .engine import NEVER, Network, Simulation, default_params, footprint  # noqa: F401
