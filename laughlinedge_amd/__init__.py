"""MI355X-native edge-state Monte-Carlo engine for Laughlin droplets.

The compute path is liblaughlinedge.so (hand-written HIP kernels for gfx950 behind the
C ABI in include/laughlinedge.h).  `laughlinedge_amd.engine` binds that ABI for Python
callers; the drop-in C++ interface is include/edge_montecarlo.h.
"""
from .engine import Engine, geneig, LEError, lib  # noqa: F401
