"""Continuous-batching BFS serving loop on one GPU.

Jobs arrive on a seeded Poisson stream; each is admitted into the batch
engine as soon as a slot and its planned VRAM footprint (the batch engine's
hipMemGetInfo-based budget) are free, joins the running layer loop at the
next layer boundary, keeps its frontier and visited set resident until it
completes, and is released right after.  Reported: jobs/s and p50/p99 of
arrival-to-result latency.

Workload (default): 2,000 jobs at 40 jobs/s; 70% permutation jobs on S_8..S_12
with 2-4 random inverse-closed generators ({a, a^-1}, {a, a^-1, t} with t a
random involution, or {a, a^-1, b, b^-1}); 30% SL(2, Z_p) jobs, 2x2 int64
matrices mod a random prime p in 101..521 with generators [[1,1],[0,1]],
[[1,0],[1,1]] and their inverses.

The reference serves jobs strictly one after another per resident worker
(multigpubfs/generic_session.py:56-144, GENERIC_SESSION_CONCURRENT_JOB); this
loop is new.
"""
from __future__ import annotations

import math
import time

from .graph_definition import GraphDefinition
from .sweep import SplitMix64, inverse, random_permutation

SERVE_SEED = 0x5345525645000000


def primes_between(lo: int, hi: int) -> list[int]:
    return [p for p in range(lo, hi + 1) if p > 1 and all(p % d for d in range(2, int(p ** 0.5) + 1))]


def random_involution(rng: SplitMix64, n: int) -> list[int]:
    """A uniformly shuffled product of disjoint transpositions (>= 1)."""
    order = random_permutation(rng, n)
    pairs = 1 + rng.below(n // 2)
    t = list(range(n))
    for k in range(pairs):
        a, b = order[2 * k], order[2 * k + 1]
        t[a], t[b] = b, a
    return t


def sl2_graph(p: int) -> GraphDefinition:
    gens = [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p), ([1, 0, p - 1, 1], p)]
    return GraphDefinition.matrix(2, 2, gens, [1, 0, 0, 1], name=f"sl2-{p}",
                                  expected_max_unique_states=p * (p * p - 1))


def make_jobs(count: int = 2000, rate: float = 40.0, seed: int = SERVE_SEED) -> list[dict]:
    rng = SplitMix64(seed)
    primes = primes_between(101, 521)
    jobs, t = [], 0.0
    for i in range(count):
        u = (rng.next() >> 11) / float(1 << 53)
        t += -math.log(1.0 - u) / rate
        if rng.below(10) < 7:
            n = 8 + rng.below(5)
            m = 2 + rng.below(3)
            a = random_permutation(rng, n)
            gens = [a, inverse(a)]
            if m == 3:
                gens.append(random_involution(rng, n))
            elif m == 4:
                b = random_permutation(rng, n)
                gens += [b, inverse(b)]
            graph = GraphDefinition.permutation(gens, list(range(n)), name=f"job-{i}")
        else:
            graph = sl2_graph(primes[rng.below(len(primes))])
        jobs.append({"id": i, "arrival": t, "graph": graph.to_dict()})
    return jobs


def percentile(values, q):
    if not values:
        return None
    s = sorted(values)
    k = min(len(s) - 1, max(0, int(math.ceil(q / 100.0 * len(s))) - 1))
    return s[k]


def serve(jobs, *, device=0, vram_budget=200 * 10 ** 9, max_slots=256, layers_per_step=4, time_scale=1.0):
    """Run the arrival stream through one batch engine.  Returns a dict with
    per-job results (layer tables, latency) and the aggregate metrics."""
    from ._native import BatchEngine
    results = {}
    with BatchEngine(device=device, max_slots=max_slots, vram_budget=vram_budget) as b:
        pending = sorted(jobs, key=lambda j: j["arrival"])
        owner = {}
        t0 = time.perf_counter()
        now = lambda: (time.perf_counter() - t0) / time_scale
        peak_active = 0
        while pending or owner:
            clock = now()
            # FIFO admission of arrived jobs while a slot and VRAM are free.
            while pending and pending[0]["arrival"] <= clock:
                slot = b.admit(pending[0]["graph"])
                if slot is None:
                    break
                job = pending.pop(0)
                owner[slot] = job
                results[job["id"]] = {"arrival": job["arrival"], "admitted": now()}
            peak_active = max(peak_active, len(owner))
            if owner:
                b.step(layers_per_step)
                for slot in b.poll():
                    sizes, status, error = b.result(slot)
                    job = owner.pop(slot)
                    b.release(slot)
                    r = results[job["id"]]
                    r.update(finished=now(), layer_sizes=sizes, error=error)
                    r["latency"] = r["finished"] - r["arrival"]
            elif pending:
                time.sleep(max(0.0, min(0.005, (pending[0]["arrival"] - now()) * time_scale)))
        makespan = now()
    lat = [r["latency"] for r in results.values()]
    return {
        "jobs": len(results),
        "jobs_per_s": len(results) / makespan if makespan > 0 else 0.0,
        "makespan_s": makespan,
        "p50_latency_s": percentile(lat, 50),
        "p99_latency_s": percentile(lat, 99),
        "max_latency_s": max(lat) if lat else None,
        "peak_concurrent_jobs": peak_active,
        "errors": sum(1 for r in results.values() if r.get("error")),
        "results": results,
    }


def run_alone(job, device=0):
    """The same job with nothing else on the GPU (its own batch engine)."""
    from ._native import BatchEngine
    with BatchEngine(device=device, max_slots=1) as b:
        slot = b.admit(job["graph"])
        if slot is None:
            raise RuntimeError("SERVE_JOB_DOES_NOT_FIT")
        b.run()
        sizes, status, error = b.result(slot)
    return sizes, error
