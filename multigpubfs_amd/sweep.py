"""Generator-set sweep: many independent BFS instances on one GPU.

`make_sweep` builds the 256-instance workload deterministically: instance i
draws from a SplitMix64 stream seeded 0x4D47424600000000 + i; n is uniform
in {9, 10}, k uniform in 2..5 generators, each a uniform random permutation
(Fisher-Yates); even instances add every generator's inverse (undirected,
inverse-closed), odd ones do not (directed, full visited history).

`run_batched` advances all instances together through the batch engine
(hipGraph-captured layer loop); `run_sequential` runs the same instances one
at a time through the plain single-instance engine.  Both return per-instance
layer tables, which must be identical.  Every leg validates its instances
and builds their library descriptors before its clock starts, so the timed
region holds the same work in each: engine set-up, BFS and reading the
tables.

The reference has no batched kernel; its sweeps run instances one by one
through resident workers (docs/reproduce-pair-sweeps.md:3,
multigpubfs/generic_session.py:139-144).
"""
from __future__ import annotations

import json
import time

from .graph_definition import GraphDefinition

SWEEP_SEED = 0x4D47424600000000
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) -- the public reference stream."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, bound: int) -> int:
        return self.next() % bound


def random_permutation(rng: SplitMix64, n: int) -> list[int]:
    p = list(range(n))
    for i in range(n - 1, 0, -1):
        j = rng.below(i + 1)
        p[i], p[j] = p[j], p[i]
    return p


def inverse(p: list[int]) -> list[int]:
    q = [0] * len(p)
    for i, v in enumerate(p):
        q[v] = i
    return q


def make_instance(i: int) -> dict:
    rng = SplitMix64(SWEEP_SEED + i)
    n = 9 + rng.below(2)
    k = 2 + rng.below(4)
    gens = [random_permutation(rng, n) for _ in range(k)]
    undirected = i % 2 == 0
    if undirected:
        gens = gens + [inverse(g) for g in gens]
    g = GraphDefinition.permutation(gens, list(range(n)), name=f"sweep-{i}")
    return g.to_dict()


def make_sweep(count: int = 256) -> list[dict]:
    return [make_instance(i) for i in range(count)]


def prepare(instances) -> list[GraphDefinition]:
    """Validated GraphDefinitions with their library descriptors built."""
    from ._native import graph_desc
    graphs = [g if isinstance(g, GraphDefinition) else GraphDefinition(g) for g in instances]
    for g in graphs:
        graph_desc(g)
    return graphs


# BatchEngine options an autotuned profile may set (mgbfs_batch_options).
ENGINE_KEYS = ("threads", "hash_load_pct", "history", "layer_pct", "bitmap_bias_permille")


def run_batched(instances, *, device=0, max_slots=None, visited="auto", layers=16, profile=None, **engine):
    """All instances through one BatchEngine; returns (layer tables, seconds, info).
    `profile` (from autotune.choose_profile) sets the launch shape and the
    engine options; keyword arguments (ENGINE_KEYS) override it.  info counts
    the visited sets and retained-layer modes chosen and the restarts."""
    from ._native import BatchEngine, device_info
    slots = max_slots or len(instances)
    tables = [None] * len(instances)
    grid = layers_per_graph = 0
    options = {}
    if profile:
        grid = profile["grid_per_cu"] * device_info(device)["compute_units"]
        layers_per_graph = profile["layers_per_graph"]
        visited = profile["visited"]
        options = {k: profile[k] for k in ENGINE_KEYS if k in profile}
    options.update(engine)
    modes = {}
    restarts = 0
    graphs = prepare(instances)
    with BatchEngine(device=device, max_slots=slots, grid=grid, layers_per_graph=layers_per_graph, **options) as b:
        t0 = time.perf_counter()
        pending = list(range(len(graphs)))
        owner = {}
        while pending or owner:
            while pending:
                slot = b.admit(graphs[pending[0]], visited=visited)
                if slot is None:
                    break
                owner[slot] = pending.pop(0)
            b.step(layers)
            for slot in b.poll():
                sizes, status, error = b.result(slot)
                if error:
                    raise RuntimeError(f"BATCH_INSTANCE_ERROR {owner[slot]}: {error}")
                info = b.instance_info(slot)
                key = f"{info['visited']}/{'rolling' if info['history_layers'] == 3 else 'full'}"
                modes[key] = modes.get(key, 0) + 1
                restarts += info["restarts"]
                tables[owner.pop(slot)] = sizes
                b.release(slot)
        seconds = time.perf_counter() - t0
    return tables, seconds, {"modes": modes, "restarts": restarts}


def run_sequential(instances, *, device=0, backend="permutation_bitmap"):
    """Plain engine, one instance at a time (create, run to completion,
    destroy).  The Lehmer-bitmap backend is the plain engine's fastest exact
    choice for these S9/S10 instances, directed or not."""
    from ._native import NativeEngine
    graphs = prepare(instances)
    tables = []
    t0 = time.perf_counter()
    for gd in graphs:
        with NativeEngine(gd, device=device, backend=backend) as e:
            status = e.run()
            if status != 1:
                raise RuntimeError(f"SEQUENTIAL_INSTANCE_STATUS {status}")
            tables.append(e.layer_sizes())
    return tables, time.perf_counter() - t0


def run_sequential_reusing(instances, *, device=0, backend="permutation_bitmap"):
    """Plain engine, one instance at a time, without per-instance device
    allocation: one engine per distinct degree is created up front (outside
    the timed region), and each instance swaps its generators into that
    engine, which resets its bitmap and frontiers in place."""
    from ._native import NativeEngine
    graphs = prepare(instances)
    engines = {}
    for g in graphs:
        if g.state_elements not in engines:
            engines[g.state_elements] = NativeEngine(g, device=device, backend=backend)
    tables = []
    try:
        t0 = time.perf_counter()
        for g in graphs:
            e = engines[g.state_elements]
            e.rebind(g)
            status = e.run()
            if status != 1:
                raise RuntimeError(f"SEQUENTIAL_INSTANCE_STATUS {status}")
            tables.append(e.layer_sizes())
        seconds = time.perf_counter() - t0
    finally:
        for e in engines.values():
            e.close()
    return tables, seconds


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="batched vs one-at-a-time generator-set sweep")
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--write-instances", help="write the instance list as JSON and exit")
    args = ap.parse_args(argv)
    inst = make_sweep(args.count)
    if args.write_instances:
        with open(args.write_instances, "w") as f:
            json.dump(inst, f)
        return 0
    batched, tb, info = run_batched(inst)
    seq, ts = run_sequential(inst)
    print(json.dumps({"instances": len(inst), "batched_s": tb, "sequential_s": ts, "speedup": ts / tb,
                      "identical": batched == seq, **info}))
    return 0 if batched == seq else 1


if __name__ == "__main__":
    raise SystemExit(main())
