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
    "segmented": "retained",           # three rotating banks: sorted when packable, else segmented hash
    "segmented24": "retained",
    "sorted": "sorted",
    "segmented_hash": "segmented",     # segmented open-addressing banks, any state width
    "permutation_bitmap": "permutation_bitmap",
    "generic": "generic",
}


def _retained_backend(graph) -> str:
    """The reference's segmented (three-bank) memory backend on this build:
    the sorted backend for inverse-closed permutations of degree <= 15 on
    labels < 16, the segmented hash for every other graph."""
    start = graph.start
    packable = (graph.action["kind"] == "permutation" and len(start) <= 15
                and all(0 <= v < 16 for v in start))
    return "sorted" if packable and graph.inverse_closed() else "segmented"


# Same file name as the reference's rank archives (specialized.py:164).
ARCHIVE_NAME = "archive-rank-0.mgbfsar1"


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
              shards=None, autotune=True, backend="auto", memory_backend="legacy", segment_rows=None,
              memory_preset="balanced", segment_strategy=None, peer_transport="auto", transport="auto",
              candidate_order="auto", bitmap_owner_page=None, timeout=None, archive=False):
    """Run one exact BFS on one GPU and return the parsed report.json.

    Keyword arguments and their validation codes follow the reference
    (launch.py:151-200).  Multi-rank and reference-internal tuning knobs
    (shards, segment_rows, segment_strategy, transport, candidate_order,
    bitmap_owner_page) are validated and otherwise have no effect: one GPU
    runs the whole BFS and the engines size their own tables.
    memory_preset="max_capacity" lowers the VRAM reserve and tightens slabs.
    The specialized SHARD_AB backends are not part of this build.

    archive=True also writes output/archive-rank-0.mgbfsar1, the layer
    archive (MGBFSAS2: up to 1000 states of every layer, one byte per
    element, sha256-chained frames; permutation graphs only), checks it the
    way the reference checks its rank archives (specialized.py:187-190:
    config digest = graph digest, one archived layer per reported layer) and
    returns its receipt under report["archive"]."""
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
    if shards is not None and (type(shards) is not int or not 1 <= shards <= 4096):
        raise ValueError("INVALID_SHARDS")
    if type(autotune) is not bool:
        raise ValueError("INVALID_AUTOTUNE")
    if segment_rows is not None and (type(segment_rows) is not int or not 1 <= segment_rows <= 1 << 32):
        raise ValueError("INVALID_SEGMENT_ROWS")
    if bitmap_owner_page is not None and (type(bitmap_owner_page) is not int or bitmap_owner_page < 0
                                          or bitmap_owner_page > 1 << 20
                                          or (bitmap_owner_page and bitmap_owner_page & (bitmap_owner_page - 1))):
        raise ValueError("INVALID_BITMAP_OWNER_PAGE")
    if backend not in ("auto", "generic", "shard_ab_hash", "shard_ab_sort_merge"):
        raise ValueError("INVALID_BACKEND")
    if backend.startswith("shard_ab_"):
        raise ValueError("SPECIALIZED_BACKEND_NOT_PROVIDED")
    if memory_backend not in _MEMORY_BACKENDS:
        raise ValueError("INVALID_MEMORY_BACKEND")
    if bitmap_owner_page and memory_backend != "permutation_bitmap":
        raise ValueError("BITMAP_OWNER_PAGE_REQUIRES_BITMAP")
    if segment_strategy is None:
        segment_strategy = os.environ.get("MGBFS_SEGMENT_STRATEGY", "atomic")
    if segment_strategy not in ("atomic", "warp", "batch"):
        raise ValueError("INVALID_SEGMENT_STRATEGY")
    if memory_backend == "legacy" and segment_strategy != "atomic":
        raise ValueError("SEGMENT_STRATEGY_REQUIRES_SEGMENTED_BACKEND")
    if memory_preset not in ("balanced", "max_capacity"):
        raise ValueError("INVALID_MEMORY_PRESET")
    if memory_preset == "max_capacity" and memory_backend not in ("segmented", "segmented24", "permutation_bitmap"):
        raise ValueError("MEMORY_PRESET_UNSUPPORTED_FOR_LEGACY")
    engine_backend = "generic" if backend == "generic" else _MEMORY_BACKENDS[memory_backend]
    if engine_backend == "retained":
        engine_backend = _retained_backend(graph)
    if engine_backend == "permutation_bitmap":
        if graph.action["kind"] != "permutation":
            raise ValueError("BITMAP_REQUIRES_PERMUTATION_ACTION")
        if not 1 <= len(graph.start) <= 64:
            raise ValueError("BITMAP_UNSUPPORTED_DEGREE")
    if peer_transport not in ("auto", "host", "lsa"):
        raise ValueError("INVALID_PEER_TRANSPORT")
    if peer_transport == "lsa":
        raise ValueError("LSA_REQUIRES_SPECIALIZED_BACKEND")
    if transport not in ("auto", "full", "parent"):
        raise ValueError("INVALID_TRANSPORT")
    if candidate_order not in ("auto", "none", "radix"):
        raise ValueError("INVALID_CANDIDATE_ORDER")
    wide = memory_backend in ("segmented", "segmented24", "permutation_bitmap")
    if capacity is not None and (type(capacity) is not int
                                 or not 1 <= capacity <= (0xFFFFFFFFFFFFFFFF if wide else 0x7FFFFFFE)):
        raise ValueError("INVALID_CAPACITY")
    if type(archive) is not bool:
        raise ValueError("INVALID_ARCHIVE")
    if archive and graph.action["kind"] != "permutation":
        raise ValueError("ARCHIVE_REQUIRES_PERMUTATION_ACTION")
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
        if memory_preset == "max_capacity":
            cmd += ["--memory-preset", "max_capacity"]
        if archive:
            cmd += ["--archive", str(output / ARCHIVE_NAME)]
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
    report = json.loads((output / "report.json").read_text())
    if archive:
        from .archive import read_archive
        receipt = read_archive(output / ARCHIVE_NAME, len(graph.start))
        if receipt["config_digest"] != report["graph_digest"]:
            raise RuntimeError("SPECIALIZED_ARCHIVE_IDENTITY_MISMATCH")
        if receipt["depths"] != len(report["layer_sizes"]):
            raise RuntimeError("SPECIALIZED_LAYER_CONSENSUS")
        report["archive"] = dict(path=str(output / ARCHIVE_NAME), format="MGBFSAS2", depths=receipt["depths"],
                                 terminal_replaced=receipt["terminal_replaced"])
    return report
