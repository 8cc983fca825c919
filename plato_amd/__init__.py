"""plato_amd: MI355X-native dual-mode push/pull edge engine.

The product is libplato_hip.so (HIP kernels for gfx950 behind the C-ABI in
include/plato_hip.h) and the host C++ headers under plato_amd/include/plato/.
This Python package is a thin ctypes binding used by the tests and the bench;
it has no CPU compute path.
"""
from . import _lib  # noqa: F401
from .engine import (Bitmap, Context, DenseState, Graph, dualmode_engine_t, graph_info_t,  # noqa: F401
                     spread_message)

__all__ = ["Context", "Graph", "DenseState", "Bitmap", "dualmode_engine_t", "graph_info_t", "spread_message"]
