"""Measured autotuner for the batched engine.

Candidate launch profiles (persistent expand grid, layers per hipGraph
launch, visited-set choice per instance) are timed on a bounded pilot: a
deterministic sample of the workload's instances.  A profile is accepted only
if every pilot instance's layer table is identical to the default profile's,
and the default is replaced only on a >= 5% improvement — the rules of the
reference's prefix autotuner (multigpubfs/autotune.py:276-691,
docs/unified-launch.md:30).  Results are cached by workload signature and
device (MGBFS_AUTOTUNE_CACHE, default ~/.cache/multigpubfs_amd/autotune.json).

Per-instance bitmap-vs-hash is decided inside the engine by planned bytes
(mgbfs_batch_plan: the Lehmer bitmap of n!/8 bytes vs an open-addressing
table of next_pow2(2 * state bound) u64 slots); the tuner can override it
for the whole batch when a pilot shows the hash faster.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path

DEFAULT = {"grid_per_cu": 8, "layers_per_graph": 4, "visited": "auto"}
CANDIDATES = [
    {"grid_per_cu": 4, "layers_per_graph": 4, "visited": "auto"},
    {"grid_per_cu": 16, "layers_per_graph": 4, "visited": "auto"},
    {"grid_per_cu": 8, "layers_per_graph": 8, "visited": "auto"},
    {"grid_per_cu": 8, "layers_per_graph": 2, "visited": "auto"},
    {"grid_per_cu": 8, "layers_per_graph": 4, "visited": "hash"},
]
MIN_GAIN = 0.05


def cache_path() -> Path:
    return Path(os.environ.get("MGBFS_AUTOTUNE_CACHE",
                               Path.home() / ".cache" / "multigpubfs_amd" / "autotune.json"))


def signature(instances) -> str:
    """Shape of the workload: (kind, degree/rows, generators, inverse-closed) counts."""
    from .graph_definition import GraphDefinition
    shapes = {}
    for g in instances:
        gd = GraphDefinition(g)
        key = (gd.action["kind"], gd.state_elements, gd.generator_count, gd.inverse_closed())
        shapes[str(key)] = shapes.get(str(key), 0) + 1
    return hashlib.sha256(json.dumps(sorted(shapes.items())).encode()).hexdigest()[:16]


def _run(instances, profile, device, cus):
    from ._native import BatchEngine
    tables = [None] * len(instances)
    with BatchEngine(device=device, max_slots=len(instances), grid=profile["grid_per_cu"] * cus,
                     layers_per_graph=profile["layers_per_graph"]) as b:
        owner = {}
        t0 = time.perf_counter()
        for i, g in enumerate(instances):
            slot = b.admit(g, visited=profile["visited"])
            if slot is None:
                raise RuntimeError("AUTOTUNE_PILOT_DOES_NOT_FIT")
            owner[slot] = i
        b.run()
        for slot in b.poll():
            sizes, status, error = b.result(slot)
            if error:
                raise RuntimeError(f"AUTOTUNE_PILOT_ERROR {error}")
            tables[owner[slot]] = sizes
        return tables, time.perf_counter() - t0


def choose_profile(instances, *, device=0, pilot=32, use_cache=True) -> dict:
    """Best measured profile for this workload (cached)."""
    from ._native import device_info
    info = device_info(device)
    cus = info["compute_units"]
    sig = f"{signature(instances)}-cu{cus}"
    path = cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except ValueError:
            cache = {}
        if sig in cache:
            return dict(cache[sig], cached=True)
    step = max(1, len(instances) // pilot)
    sample = list(instances)[::step][:pilot]
    _run(sample, DEFAULT, device, cus)  # warm kernels and allocator
    base_tables, base_t = _run(sample, DEFAULT, device, cus)
    best, best_t = dict(DEFAULT), base_t
    trials = [{"profile": dict(DEFAULT), "seconds": base_t}]
    for cand in CANDIDATES:
        tables, t = _run(sample, cand, device, cus)
        trials.append({"profile": cand, "seconds": t, "identical": tables == base_tables})
        if tables == base_tables and t < best_t * (1 - MIN_GAIN):
            best, best_t = dict(cand), t
    result = dict(best, pilot_instances=len(sample), pilot_seconds=best_t, default_seconds=base_t, trials=trials)
    if use_cache:
        cache[sig] = result
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(cache, indent=1))
    return dict(result, cached=False)
