"""MI355X-native exact BFS engine for implicit Cayley graphs.

Native code (HIP kernels, engine, C ABI) lives in csrc/ and is built in-tree
into lib/libmgbfs_hip.so; there is no CPU execution path.
"""
from .graph_definition import GraphDefinition, from_cayleypy, lrx
from ._native import NativeEngine, NativeUnavailable, device_info

__all__ = ["GraphDefinition", "from_cayleypy", "lrx", "NativeEngine", "NativeUnavailable", "device_info"]
