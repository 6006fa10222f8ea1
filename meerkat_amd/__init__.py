"""meerkat_amd: MI355X (gfx950) dynamic-graph hot path behind a C ABI.

The compute lives in libmeerkat_hip.so (hand-written HIP kernels); this
package is the host-side mirror of the reference's C++ API for scripts and
tests.  Importing it does not load the library; the first call does, and
fails loudly if it is not built.
"""
from ._lib import MeerkatError, header_symbols, lib
from .algorithms import BFS, CC, INF, SSSP, TC, MultiSource, PageRank
from .graph import DynamicGraph, SlabPool, device_u32
from . import autotune, gen, serving

__all__ = [
    "BFS",
    "CC",
    "DynamicGraph",
    "INF",
    "MeerkatError",
    "MultiSource",
    "PageRank",
    "SSSP",
    "SlabPool",
    "TC",
    "device_u32",
    "gen",
    "header_symbols",
    "lib",
    "serving",
]
