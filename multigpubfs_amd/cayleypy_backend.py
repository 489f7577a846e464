"""CayleyPy integration: graph.bfs() on the MI355X engine, explicit and reversible.

The public contract is the reference's (multigpubfs/cayleypy_backend.py:
15-204): the names NativeOptions, NativeBfsResult, bfs, enable_native and
disable_native, their fields and keywords, the validation codes, the
backend choice native / torch / auto, the CayleyPy keyword defaults the
native path accepts, and the compact-sample result.  Importing the module
changes nothing; enable_native() puts one wrapper on CayleyGraph.bfs and
disable_native() takes it off again, refusing when something else replaced
the method in between.

Every native search goes through run_graph (the `mgbfs graph` process
boundary) and reads its report.json and states.json.
"""
from __future__ import annotations

import json
import os
import tempfile
import threading
from dataclasses import dataclass, field
from functools import wraps
from pathlib import Path

from .graph_definition import GraphDefinition, _flatten, from_cayleypy
from .launch import _MEMORY_BACKENDS, run_graph

# CayleyGraph.bfs keywords the native path accepts, at CayleyPy's defaults
# only (cayleypy_backend.py:80-90); start_states is handled separately.
CAYLEYPY_DEFAULTS = {"max_layer_size_to_store": 1000, "max_layer_size_to_explore": 10 ** 12,
                     "max_diameter": 1000000, "return_all_edges": False, "return_all_hashes": False,
                     "stop_condition": None, "disable_batching": False}
BACKEND_CHOICES = ("native", "torch", "auto")
SEGMENT_STRATEGIES = ("atomic", "warp", "batch")


def _owner_page_error(page, memory_backend):
    """Validation code for bitmap_owner_page, or None (0 or a power of two
    up to 2^20, and only with the permutation bitmap)."""
    if page is None:
        return None
    if type(page) is not int or not 0 <= page <= 1 << 20 or page & (page - 1):
        return "INVALID_BITMAP_OWNER_PAGE"
    if page and memory_backend != "permutation_bitmap":
        return "BITMAP_OWNER_PAGE_REQUIRES_BITMAP"
    return None


@dataclass(frozen=True)
class NativeOptions:
    output_root: str | Path | None = None
    devices: tuple[int, ...] | None = None
    max_seconds: int | None = None
    autotune: bool = True
    memory_backend: str = "legacy"
    memory_preset: str = "balanced"
    segment_rows: int = 1 << 32
    segment_strategy: str = field(default_factory=lambda: os.environ.get("MGBFS_SEGMENT_STRATEGY", "atomic"))
    bitmap_owner_page: int | None = None

    def __post_init__(self):
        problems = (
            (self.segment_strategy not in SEGMENT_STRATEGIES, "INVALID_SEGMENT_STRATEGY"),
            (self.memory_backend not in _MEMORY_BACKENDS, "INVALID_MEMORY_BACKEND"),
            (self.memory_preset not in ("balanced", "max_capacity"), "INVALID_MEMORY_PRESET"),
            (type(self.segment_rows) is not int or not 1 <= self.segment_rows <= 1 << 32, "INVALID_SEGMENT_ROWS"),
        )
        for failed, code in problems:
            if failed:
                raise ValueError(code)
        code = _owner_page_error(self.bitmap_owner_page, self.memory_backend)
        if code:
            raise ValueError(code)

    def run_keywords(self) -> dict:
        """The run_graph keywords these options stand for."""
        return {"devices": None if self.devices is None else list(self.devices), "max_seconds": self.max_seconds,
                "autotune": self.autotune, "memory_backend": self.memory_backend,
                "segment_rows": self.segment_rows, "memory_preset": self.memory_preset,
                "segment_strategy": self.segment_strategy, "bitmap_owner_page": self.bitmap_owner_page}


@dataclass(frozen=True)
class NativeBfsResult:
    """Exact layer counts with compact state samples.  Not CayleyPy's full
    BfsResult: small layers, edges, hashes and paths are not retained; a
    layer is available in full only when its sample holds all of it."""
    bfs_completed: bool
    layer_sizes: list
    graph: object
    report: dict
    output: Path
    layers: dict = field(default_factory=dict)
    layer_samples: dict = field(default_factory=dict)
    backend: str = field(default="native", init=False)

    def num_vertices(self):
        return sum(self.layer_sizes)

    def diameter(self):
        if not self.bfs_completed:
            raise RuntimeError("BFS_INCOMPLETE: only completed_depth is known")
        return self.completed_depth

    @property
    def completed_depth(self):
        return len(self.layer_sizes) - 1

    def get_layer(self, layer_id):
        if layer_id not in self.layers:
            raise KeyError("Full layer not retained; use layer_samples for compact samples")
        return self.layers[layer_id].cpu().numpy()

    def last_layer(self):
        return self.get_layer(self.completed_depth)

    @classmethod
    def from_run(cls, report, states, definition, graph, output):
        """Build the result from a run's report and states.json: the last
        layer's sample and the previous layer's (present only when small),
        as int64 tensors shaped like CayleyPy states."""
        import torch  # CayleyPy hands states back as tensors
        sizes = report["layer_sizes"]
        action = definition.action
        shape = (-1, action["rows"], action["cols"]) if action["kind"] == "matrix" else None
        samples, full = {}, {}
        for key, depth in (("current", len(sizes) - 1), ("previous_small", len(sizes) - 2)):
            rows = states.get(key) or []
            if depth < 0 or not rows:
                continue
            tensor = torch.tensor(rows, dtype=torch.int64)
            samples[depth] = tensor.reshape(shape) if shape else tensor
            if len(rows) == sizes[depth]:
                full[depth] = samples[depth]
        return cls(report["status"] == "COMPLETE", sizes, graph, report, output, full, samples)


def _first_unsupported(kwargs):
    """Name of the first keyword the native path cannot honour, else None."""
    return next((key for key, value in kwargs.items()
                 if key != "start_states" and (key not in CAYLEYPY_DEFAULTS or value != CAYLEYPY_DEFAULTS[key])),
                None)


def _definition_for(graph, start_states):
    """The graph's definition, with its start replaced by one decoded
    start state when given; None when start_states holds more than one."""
    definition = from_cayleypy(graph)
    if start_states is None:
        return definition
    start = _flatten(start_states)
    if len(start) != len(definition.start):
        return None
    value = definition.to_dict()
    value["start"] = start
    return GraphDefinition.from_dict(value)


def _fresh_output(options):
    """A new run directory under output_root (default ./bfs-results); the
    parent stays for the receipt and run_graph creates the child."""
    root = Path(options.output_root or Path.cwd() / "bfs-results")
    root.mkdir(parents=True, exist_ok=True)
    return Path(tempfile.mkdtemp(prefix="run-", dir=root)) / "result"


def _dispatch(original, graph, backend, options, output, kwargs):
    """One CayleyGraph.bfs call: the original method (torch, or auto when
    the native path cannot serve the call) or a native search."""
    if backend not in BACKEND_CHOICES:
        raise ValueError("backend must be native, torch or auto")
    if not isinstance(options, NativeOptions):
        raise TypeError("native_options must be NativeOptions")
    if backend == "torch":
        if output is not None:
            raise ValueError("output belongs to the native backend")
        return original(graph, **kwargs)
    may_fall_back = backend == "auto" and output is None
    blocker = _first_unsupported(kwargs)
    definition = None if blocker else _definition_for(graph, kwargs.get("start_states"))
    if blocker or definition is None:
        if may_fall_back:
            return original(graph, **kwargs)
        if blocker:
            raise NotImplementedError(f"Native BFS does not support {blocker}; use backend='torch'")
        raise NotImplementedError("Native BFS supports a single decoded start state")
    target = Path(output if output is not None else _fresh_output(options)).absolute()
    report = run_graph(definition, target, **options.run_keywords())
    states = json.loads((target / "states.json").read_text())
    return NativeBfsResult.from_run(report, states, definition, getattr(graph, "definition", graph), target)


class _Hook:
    """The single wrapper this module may put on CayleyGraph.bfs, and the
    method it replaced."""

    def __init__(self):
        self.lock = threading.RLock()
        self.replaced = None   # CayleyPy's own method while the hook is on
        self.installed = None  # our wrapper while the hook is on

    def original(self, cls):
        return self.replaced or cls.bfs

    def check_ownership(self, cls):
        if self.replaced is not None and cls.bfs is not self.installed:
            raise RuntimeError("CayleyGraph.bfs was replaced by another integration")

    def put_on(self, cls, options, default_backend):
        with self.lock:
            self.check_ownership(cls)
            base = self.original(cls)

            @wraps(base)
            def bfs_wrapper(graph, **kwargs):
                backend = kwargs.pop("backend", default_backend)
                chosen = kwargs.pop("native_options", options)
                return _dispatch(base, graph, backend, chosen, kwargs.pop("output", None), kwargs)

            self.replaced, self.installed = base, bfs_wrapper
            cls.bfs = bfs_wrapper

    def take_off(self, cls):
        with self.lock:
            if self.replaced is None:
                return
            self.check_ownership(cls)
            cls.bfs = self.replaced
            self.replaced = self.installed = None


_HOOK = _Hook()


def bfs(graph, *, backend="native", native_options=None, output=None, **kwargs):
    """One search without touching CayleyGraph."""
    from cayleypy import CayleyGraph
    with _HOOK.lock:
        original = _HOOK.original(CayleyGraph)
    return _dispatch(original, graph, backend, native_options or NativeOptions(), output, kwargs)


def enable_native(options=None, *, default_backend="native", recovery=True):
    """Route CayleyGraph.bfs through the native engine (explicit opt-in).

    `recovery` is accepted for the reference's signature: this build keeps no
    recovery registry (a failed run leaves its receipt in the output
    directory and nothing to restore)."""
    from cayleypy import CayleyGraph
    if default_backend not in BACKEND_CHOICES:
        raise ValueError("Invalid default_backend")
    chosen = NativeOptions() if options is None else options
    if not isinstance(chosen, NativeOptions):
        raise TypeError("options must be NativeOptions")
    if default_backend == "native":
        from .launch import native_runtime
        native_runtime()  # a missing executable fails here, not on the first search
    _HOOK.put_on(CayleyGraph, chosen, default_backend)


def disable_native():
    """Restore CayleyPy's method unless another integration replaced ours."""
    from cayleypy import CayleyGraph
    _HOOK.take_off(CayleyGraph)
