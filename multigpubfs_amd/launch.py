"""run_graph: launch the exact BFS through the `mgbfs` executable.

Mirrors multigpubfs/launch.py:151-185 (argument names, validation error codes,
report dict) for the single-GPU build: one process owns one GPU; the process
boundary is `mgbfs graph <graph.json> <out-dir> --device N [--capacity N]
[--seconds N] --backend NAME`, which writes report.json and states.json.
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile
from pathlib import Path

from ._native import CLI_PATH
from .graph_definition import GraphDefinition, from_cayleypy

# memory_backend names of the reference mapped onto this build's backends.
_MEMORY_BACKENDS = {
    "legacy": "auto",                  # best exact backend for the graph
    "auto": "auto",
    "segmented": "sorted",             # sort-unique + three rotating banks
    "segmented24": "sorted",
    "sorted": "sorted",
    "permutation_bitmap": "permutation_bitmap",
    "generic": "generic",
}


class NativeError(RuntimeError):
    """The native executable failed; .code is its ERROR code."""

    def __init__(self, code: str, stderr: str = ""):
        super().__init__(code)
        self.code = code
        self.stderr = stderr


def native_runtime(executable=None) -> Path:
    """The executable: explicit argument, MGBFS_EXECUTABLE, or the in-tree build
    (native_distribution.py:40-44 resolves MGBFS_EXECUTABLE the same way)."""
    path = Path(executable or os.environ.get("MGBFS_EXECUTABLE") or CLI_PATH)
    if not path.exists():
        raise FileNotFoundError(f"NATIVE_EXECUTABLE_MISSING: {path} (run python -m multigpubfs_amd.build)")
    return path


def _error_code(stderr: str) -> str:
    for line in reversed(stderr.strip().splitlines()):
        try:
            doc = json.loads(line)
        except ValueError:
            continue
        if isinstance(doc, dict) and doc.get("status") == "ERROR":
            return str(doc.get("error"))
    return "NATIVE_FAILED"


def run_graph(graph, output, *, device=None, devices=None, capacity=None, max_seconds=None, executable=None,
              autotune=True, backend="auto", memory_backend="legacy", timeout=None):
    """Run one exact BFS on one GPU and return the parsed report.json."""
    if not isinstance(graph, GraphDefinition):
        graph = from_cayleypy(graph)
    if device is not None and devices is not None:
        raise ValueError("DEVICE_SELECTION_CONFLICT")
    if device is not None:
        if type(device) is not int or device < 0:
            raise ValueError("INVALID_DEVICE")
        devices = [device]
    if devices is not None:
        if (not isinstance(devices, (list, tuple)) or not devices or any(type(v) is not int or v < 0 for v in devices)
                or len(set(devices)) != len(devices)):
            raise ValueError("INVALID_DEVICES")
        if len(devices) != 1:
            raise ValueError("SINGLE_DEVICE_BUILD: one BFS runs on one GPU")
    device = devices[0] if devices else 0
    if max_seconds is not None and (type(max_seconds) is not int or max_seconds < 0):
        raise ValueError("INVALID_MAX_SECONDS")
    if type(autotune) is not bool:
        raise ValueError("INVALID_AUTOTUNE")
    if backend not in ("auto", "generic"):
        raise ValueError("INVALID_BACKEND")
    if memory_backend not in _MEMORY_BACKENDS:
        raise ValueError("INVALID_MEMORY_BACKEND")
    engine_backend = "generic" if backend == "generic" else _MEMORY_BACKENDS[memory_backend]
    if engine_backend == "permutation_bitmap" and graph.action["kind"] != "permutation":
        raise ValueError("BITMAP_REQUIRES_PERMUTATION_ACTION")
    if capacity is not None and (type(capacity) is not int or not 1 <= capacity <= 0xFFFFFFFFFFFFFFFF):
        raise ValueError("INVALID_CAPACITY")
    output = Path(output)
    exe = native_runtime(executable)
    with tempfile.TemporaryDirectory(prefix="mgbfs-graph-") as tmp:
        path = Path(tmp) / "graph.json"
        path.write_text(graph.to_json())
        cmd = [str(exe), "graph", str(path), str(output), "--device", str(device), "--backend", engine_backend]
        if capacity is not None:
            cmd += ["--capacity", str(capacity)]
        if max_seconds:
            cmd += ["--seconds", str(max_seconds)]
        from . import session
        pool = session.active_pool()
        if pool is not None:
            # Resident worker: same arguments, one job file instead of a process.
            job = pool.worker_for(exe, device).submit("graph", cmd[2:])
            if job.wait(timeout):
                raise NativeError(job.error or "GENERIC_SESSION_JOB_FAILED", json.dumps(job.response))
        else:
            proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
            if proc.returncode:
                raise NativeError(_error_code(proc.stderr), proc.stderr)
    return json.loads((output / "report.json").read_text())
