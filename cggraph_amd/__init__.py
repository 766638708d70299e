"""cggraph_amd: MI355X-native frontier traversal with CGgraph's surface.

The compute path is libcgmi.so (hand-written gfx950 HIP kernels behind a C ABI,
include/cgmi.h); this package is its Python binding plus a mirror of the
reference's driver-level names (Algorithm_type, Engine_type, OrderMethod,
CGgraphEngine, TwoLevelPartition, main_partition, ...).
"""
from ._lib import CgError, CG_UNREACHED, default_options, lib  # noqa: F401
from .core import Engine, Graph  # noqa: F401
