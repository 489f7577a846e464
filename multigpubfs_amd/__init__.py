"""MI355X-native exact BFS engine for implicit Cayley graphs.

Native code (HIP kernels, engine, C ABI) lives in csrc/ and is built in-tree
into lib/libmgbfs_hip.so and bin/mgbfs; there is no CPU execution path.
The public API mirrors the reference package multigpubfs
(multigpubfs/__init__.py:2-9).
"""
from .graph_definition import GraphDefinition, from_cayleypy, lrx
from ._native import NativeEngine, NativeUnavailable, device_info
from .launch import NativeError, run_graph
from .cayleypy_backend import NativeBfsResult, NativeOptions, bfs, disable_native, enable_native
from .catalog import from_catalog
from .session import persistent_graph_workers

__all__ = ["GraphDefinition", "from_cayleypy", "from_catalog", "lrx", "run_graph", "NativeError", "NativeOptions",
           "NativeBfsResult", "bfs", "enable_native", "disable_native", "persistent_graph_workers", "NativeEngine",
           "NativeUnavailable", "device_info"]
