"""Measured autotuner for the batched engine.

A profile sets every knob of the batched engine: the expand workgroup size
(threads), the persistent expand grid (grid_per_cu), the layers per hipGraph
launch, the visited-set policy (the automatic per-instance choice, or hash
forced), the per-instance bitmap-vs-hash rule (bitmap_bias_permille: an
instance takes the Lehmer bitmap when its bytes are at most that many
thousandths of its hash tables'), the hash tables' load factor
(hash_load_pct), and the retained history of inverse-closed hash instances
(the full orbit when it fits, or always three rolling tables sized at
layer_pct % of the state bound).

Profiles are timed on a bounded pilot, a deterministic sample of the
workload's instances, by coordinate descent: each knob in turn tries its
alternatives with the others at the best values so far.  A value is taken
only if every pilot instance's layer table is identical to the default
profile's and the pilot runs >= 5% faster -- the rules of the reference's
prefix autotuner (multigpubfs/autotune.py:276-691, docs/unified-launch.md:30).
The hash knobs are tried only when some pilot instance runs on a hash table.
Results are cached by workload signature and device (MGBFS_AUTOTUNE_CACHE,
default ~/.cache/multigpubfs_amd/autotune.json).
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path

DEFAULT = {"threads": 256, "grid_per_cu": 8, "layers_per_graph": 4, "visited": "auto",
           "bitmap_bias_permille": 1000, "hash_load_pct": 50, "history": "auto", "layer_pct": 30}
# Alternatives per knob, in the order the descent visits them.
DIMENSIONS = [
    ("threads", [128, 512]),
    ("grid_per_cu", [4, 16]),
    ("layers_per_graph", [2, 8]),
    ("visited", ["hash"]),
    ("bitmap_bias_permille", [5, 4000]),
    ("hash_load_pct", [30, 70]),
    ("history", ["rolling"]),
    ("layer_pct", [50]),
]
HASH_KNOBS = {"hash_load_pct", "history", "layer_pct"}
MIN_GAIN = 0.05
VERSION = 2


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
    return hashlib.sha256(json.dumps([VERSION, sorted(shapes.items())]).encode()).hexdigest()[:16]


def _run(instances, profile, device, cus):
    """Pilot run: (layer tables, seconds, visited sets used)."""
    from ._native import BatchEngine
    from .sweep import ENGINE_KEYS
    tables = [None] * len(instances)
    used = set()
    options = {k: profile[k] for k in ENGINE_KEYS}
    with BatchEngine(device=device, max_slots=len(instances), grid=profile["grid_per_cu"] * cus,
                     layers_per_graph=profile["layers_per_graph"], **options) as b:
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
            used.add(b.instance_info(slot)["visited"])
        return tables, time.perf_counter() - t0, used


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
    base_tables, base_t, used = _run(sample, DEFAULT, device, cus)
    best, best_t, best_used = dict(DEFAULT), base_t, used
    trials = [{"profile": dict(DEFAULT), "seconds": base_t}]
    for knob, values in DIMENSIONS:
        if knob in HASH_KNOBS and "hash" not in best_used:
            continue
        current = dict(best)
        for value in values:
            cand = dict(current, **{knob: value})
            tables, t, used = _run(sample, cand, device, cus)
            same = tables == base_tables
            trials.append({"profile": cand, "seconds": t, "identical": same})
            if same and t < best_t * (1 - MIN_GAIN):
                best, best_t, best_used = cand, t, used
    result = dict(best, pilot_instances=len(sample), pilot_seconds=best_t, default_seconds=base_t, trials=trials)
    if use_cache:
        cache[sig] = result
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(cache, indent=1))
    return dict(result, cached=False)
