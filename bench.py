"""Benchmark: expanded edges/s of a full exact LRX BFS on S_13 on one MI355X.

A step is one complete BFS of the LRX Cayley graph of S_13 from the identity
(6,227,020,800 states, 79 layers, 3 generators -> 18,681,062,400 expanded
edges) through libmgbfs_hip.so, with the engine's device buffers allocated
beforehand (engine reset + layer loop are inside the timed region).  Every
step's layer table is checked against the reference's golden S13 table.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 is refused: this build
runs one BFS per GPU and has no multi-GPU path.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE per kernel for one full BFS,
# committed beside the text summaries (profiles/*_s13_lrx.md).
TRAFFIC_FILES = {"sorted": ROOT / "profiles" / "sorted_engine_s13_lrx.json",
                 "bitmap": ROOT / "profiles" / "bitmap_expand_kernel_s13_lrx.json"}
REFERENCE_EDGES_PER_S = 18681062400 / 11.716894  # 8x H200 median, docs/validation/lrx13-eight-h200.md:26


def host_cores() -> tuple[int, int]:
    """(threads this job may use, processors the host reports).  A GPU job's
    CPU share is the processors it may run on, capped by the thread count
    the job environment sets (OMP_NUM_THREADS, 16 per GPU on the MI355X
    boxes), whatever os.cpu_count() reports for the whole machine."""
    nproc = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    limit = os.environ.get("OMP_NUM_THREADS")
    if limit and limit.isdigit() and int(limit) > 0:
        usable = min(usable, int(limit))
    return max(1, usable), nproc


def cpu_baseline(n: int, seconds_target: float) -> dict:
    """The oracle's packed Lehmer-bitmap BFS on this job's host cores, on a
    bounded prefix of the same S_n LRX BFS (stops after the layer at which
    the edge budget is reached)."""
    from oracle import oracle as O
    cores, nproc = host_cores()
    budget = int(seconds_target * 10e6 * cores)  # ~10 M edges/s per thread
    t0 = time.perf_counter()
    r = O.packed_bfs(n, threads=cores, stop_after_edges=budget)
    dt = time.perf_counter() - t0
    return {"value": r["edges"] / dt, "unit": "edges/s", "cores": cores, "kind": "port", "host_nproc": nproc,
            "sample": f"S{n} LRX BFS prefix: first {len(r['layers'])} of {79 if n == 13 else '?'} layers, "
                      f"{r['edges']} edges expanded in {dt:.2f} s (oracle/mgbfs_oracle.c orc_packed_bfs, "
                      f"{cores} threads: the job's CPU share of a {nproc}-processor host)"}


def sweep_section() -> dict:
    """256 S9/S10 instances: batched (one hipGraph layer loop) vs one at a time."""
    from multigpubfs_amd.sweep import run_batched, run_sequential, run_sequential_reusing
    golden = json.loads((ROOT / "tests" / "golden" / "sweep256.json").read_text())["instances"]
    inst = [x["graph"] for x in golden]
    from multigpubfs_amd.autotune import DEFAULT as DEFAULT_PROFILE, choose_profile
    t0 = time.perf_counter()
    profile = choose_profile(inst, use_cache=False)
    tune_s = time.perf_counter() - t0
    run_batched(inst[:8], profile=profile)  # warm the kernels and the allocator
    batched, tb, info = run_batched(inst, profile=profile)
    seq, ts = run_sequential(inst)
    reuse, tr = run_sequential_reusing(inst)
    want = [x["layer_sizes"] for x in golden]
    return {"instances": len(inst), "batched_s": tb, "sequential_s": ts, "sequential_reusing_s": tr,
            "batched_bfs_per_s": len(inst) / tb, "sequential_bfs_per_s": len(inst) / ts,
            "speedup": ts / tb, "speedup_vs_reusing": tr / tb,
            "sequential_legs": {"create_per_instance": "plain engine created, run and destroyed per instance",
                                "reusing": "plain engine per distinct degree, kept; one device allocation "
                                           "reused by every instance of that degree"},
            "identical_tables": batched == seq == reuse, "golden_exact": batched == want,
            "states": sum(sum(w) for w in want), "autotune_seconds": tune_s,
            "modes": info["modes"], "restarts": info["restarts"],
            "profile": {k: profile[k] for k in DEFAULT_PROFILE}}


def serving_checks(jobs, results) -> dict:
    """Every job's table from the serving loop against an independent one:
    permutation jobs of degree <= 10 against the oracle's committed tables
    (tests/golden/serve2000_small.json, same seeded stream), degree 11 and 12
    against the plain engine run alone on the sorted backend (one engine per
    degree, given each job's generators in turn), and SL(2, Z_p) jobs
    against |SL(2,p)| = p(p^2 - 1) and the plain engine's generic backend
    (one run per distinct p)."""
    from multigpubfs_amd import GraphDefinition, NativeEngine
    golden = json.loads((ROOT / "tests" / "golden" / "serve2000_small.json").read_text())
    small = golden["tables"] if len(jobs) == golden["count"] else {}
    plain_sl2, plain_perm = {}, {}
    perm = {"jobs": 0, "mismatches": 0, "vs_golden": 0, "vs_plain_engine": 0, "max_degree": 0}
    sl2 = {"jobs": 0, "mismatches": 0, "max_p": 0}
    t0 = time.perf_counter()
    for done, job in enumerate(jobs):
        if done % 250 == 0:
            print(f"[bench] serving checks: {done}/{len(jobs)} jobs, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
        got = results[job["id"]]["layer_sizes"]
        graph = job["graph"]
        a = graph["action"]
        if a["kind"] == "permutation":
            perm["jobs"] += 1
            perm["max_degree"] = max(perm["max_degree"], a["degree"])
            if str(job["id"]) in small:
                want = small[str(job["id"])]
                perm["vs_golden"] += 1
            else:
                gd = GraphDefinition(graph)
                e = plain_perm.get(a["degree"])
                if e is None:
                    e = plain_perm[a["degree"]] = NativeEngine(gd, backend="sorted")
                else:
                    e.rebind(gd)
                e.run()
                want = e.layer_sizes()
                perm["vs_plain_engine"] += 1
            perm["mismatches"] += got != want
        else:
            p = a["generators"][0]["modulo"]
            sl2["jobs"] += 1
            sl2["max_p"] = max(sl2["max_p"], p)
            if p not in plain_sl2:
                with NativeEngine(GraphDefinition(graph), backend="generic") as e:
                    e.run()
                    plain_sl2[p] = e.layer_sizes()
            sl2["mismatches"] += got != plain_sl2[p] or sum(got) != p * (p * p - 1)
    for e in plain_perm.values():
        e.close()
    return {"perm_tables_exact": perm["mismatches"] == 0, "sl2_tables_exact": sl2["mismatches"] == 0,
            "permutation": perm, "sl2": sl2, "distinct_sl2_primes": len(plain_sl2),
            "check_seconds": time.perf_counter() - t0}


def serving_section(count: int) -> dict:
    """Poisson stream (40 jobs/s) through the continuous-batching loop, every
    table checked; then the same jobs all arriving at once (arrival far above
    service), whose jobs/s is the loop's capacity."""
    from multigpubfs_amd.serving import make_jobs, serve
    jobs = make_jobs(count)
    print(f"[bench] serving {count} jobs at 40 jobs/s", file=sys.stderr, flush=True)
    report = serve(jobs, vram_budget=200 * 10 ** 9)
    out = {k: v for k, v in report.items() if k != "results"}
    out.update(vram_budget_bytes=200 * 10 ** 9, arrival_rate_per_s=40)
    out.update(serving_checks(jobs, report["results"]))
    print(f"[bench] serving the same {count} jobs arriving at once", file=sys.stderr, flush=True)
    burst = [dict(job, arrival=0.0) for job in jobs]
    saturated = serve(burst, vram_budget=200 * 10 ** 9)
    same = all(saturated["results"][j["id"]]["layer_sizes"] == report["results"][j["id"]]["layer_sizes"]
               for j in jobs)
    out["saturation"] = {"jobs": saturated["jobs"], "capacity_jobs_per_s": saturated["jobs_per_s"],
                         "makespan_s": saturated["makespan_s"], "p50_latency_s": saturated["p50_latency_s"],
                         "p99_latency_s": saturated["p99_latency_s"],
                         "peak_concurrent_jobs": saturated["peak_concurrent_jobs"],
                         "errors": saturated["errors"], "tables_same_as_paced_run": same}
    return out


def s11_bitmap_section(golden_all, steps=5):
    """S11 LRX on the Lehmer-rank permutation bitmap (39,916,800 states):
    mean of `steps` full BFS after one warmup.  Each layer table is checked
    against the golden table when the golden file holds one, else against
    the sorted backend's table for the same graph (an independent visited
    set: hash + sort-unique instead of the rank bitmap) and 11! states."""
    from multigpubfs_amd import NativeEngine, lrx
    n = 11
    states = math.factorial(n)
    golden = golden_all.get(str(n))
    if golden is None:
        with NativeEngine(lrx(n), backend="sorted") as ref:
            ref.run()
            golden = ref.layer_sizes() if sum(ref.layer_sizes()) == states else None
    with NativeEngine(lrx(n), backend="permutation_bitmap") as e:
        e.run()
        exact = (e.layer_sizes() == golden) if golden else sum(e.layer_sizes()) == states
        t0 = time.perf_counter()
        for _ in range(steps):
            e.reset()
            e.run()
            sizes = e.layer_sizes()
            exact &= (sizes == golden) if golden else sum(sizes) == states
        dt = (time.perf_counter() - t0) / steps
    return {"workload": "S11 LRX full exact BFS, permutation bitmap", "states": states, "layers": len(sizes),
            "ms_per_bfs": 1e3 * dt, "edges_per_s": 3 * states / dt, "layer_table_exact": bool(exact),
            "checked_against": "golden table" if str(n) in golden_all else "sorted backend's table"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=13)
    ap.add_argument("--backend", default="sorted")
    ap.add_argument("--engine-flags", type=int, default=0,
                    help="mgbfs_engine_options.flags for the timed engine (MGBFS_FLAG_* of include/mgbfs_hip.h)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the 256-instance batched sweep")
    ap.add_argument("--no-bitmap", action="store_true", help="skip the S11 permutation-bitmap run")
    ap.add_argument("--no-serve", action="store_true", help="skip the 2,000-job serving run")
    ap.add_argument("--serve-jobs", type=int, default=2000)
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 1 GiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args(argv)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "single-GPU build: one exact BFS runs on one GPU; "
                                   "multi-GPU sharding is not part of this build", "n_gpus": args.gpus}))
        return 2

    if args.calibrate:
        import torch
        src = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 1073741824 bytes read + written", file=sys.stderr, flush=True)
        del src, dst
    from multigpubfs_amd import NativeEngine, lrx
    golden_all = json.loads((ROOT / "tests" / "golden" / "lrx_layers.json").read_text())["layers"]
    golden = golden_all.get(str(args.n))
    n = args.n
    states = math.factorial(n)
    edges = 3 * states
    engine = NativeEngine(lrx(n), backend=args.backend, flags=args.engine_flags)
    print(f"[bench] engine {engine.backend_name} S{n} LRX; warmup {args.warmup}", file=sys.stderr, flush=True)
    exact = True
    for _ in range(args.warmup):
        engine.reset()
        engine.run()
        sizes = engine.layer_sizes()
        exact &= (sizes == golden) if golden else sum(sizes) == states
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None
    sync()
    kernel_s = 0.0
    launches = 0
    algo_bytes = 0
    phase_s = [0.0] * 4
    phase_b = [0] * 4
    t0 = time.perf_counter()
    for _ in range(args.steps):
        engine.reset()
        status = engine.run()
        info = engine.info()
        kernel_s += info["kernel_seconds"]
        launches += info["kernel_launches"]
        algo_bytes += info["algorithmic_bytes"]
        for i, (ps, pb) in enumerate(engine.phases(4)):
            phase_s[i] += ps
            phase_b[i] += pb
        sizes = engine.layer_sizes()
        exact &= status == 1 and ((sizes == golden) if golden else sum(sizes) == states)
    sync()
    elapsed = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    value = args.steps * edges / elapsed
    layers = len(sizes)
    if info["backend"] == 1:
        names = ["bitmap_expand_kernel"]
        models = ["8F frontier read + 8N next-layer write (bitmap probes excluded)"]
    else:
        names = ["expand_scatter_kernel", "partition_scatter_kernel", "unique_diff_kernel", "compact_kernel"]
        models = ["8F + 8C", "16C", "8C + 8F + 8N (back-edge masks: current bank only)", "16N"]
    phases = []
    for i, name in enumerate(names):
        if phase_s[i] <= 0 or phase_b[i] == 0:  # compaction fused into unique_diff: no bytes of its own
            continue
        gbs = phase_b[i] / phase_s[i] / 1e9
        phases.append({"kernel": name, "ms_per_step": 1e3 * phase_s[i] / args.steps,
                       "bytes_per_step": phase_b[i] // args.steps, "achieved_gbs": gbs, "bytes_model": models[i]})
    if not phases:  # backends without per-phase events: all their kernels as one phase
        phases.append({"kernel": f"{engine.backend_name} kernels", "ms_per_step": 1e3 * kernel_s / args.steps,
                       "bytes_per_step": algo_bytes // args.steps,
                       "achieved_gbs": algo_bytes / kernel_s / 1e9 if kernel_s > 0 else 0.0,
                       "bytes_model": "the engine's streaming-bytes model"})
    dom = max(phases, key=lambda p: p["ms_per_step"])
    traffic, traffic_source = None, None
    tf = TRAFFIC_FILES.get("bitmap" if info["backend"] == 1 else "sorted")
    if n == 13 and tf is not None and tf.exists():
        prof = json.loads(tf.read_text())
        k = prof["kernels"].get(dom["kernel"])
        if k:
            traffic = (k["fetch_bytes"] + k["write_bytes"]) / k["dispatches"]
            traffic_source = f"{tf.relative_to(ROOT)} ({prof['library']})"
    per_launch_layers = layers * args.steps
    result = {
        "metric": f"expanded edges/s on S{n} LRX BFS",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REFERENCE_EDGES_PER_S if n == 13 else None,
        "dtype": "u64",
        "data": "synthetic: implicit Cayley graph (LRX generators on S_n from the identity), no dataset",
        "config": {"workload": f"S{n} LRX full exact BFS, one GPU", "backend": engine.backend_name,
                   "engine_flags": args.engine_flags,
                   "states": states, "layers": layers, "edges_per_step": edges,
                   "peak_layer": max(sizes), "layer_table_exact": bool(exact),
                   "unique_candidates_per_step": info["unique_candidates"],
                   "generated_candidates_per_step": info["candidates"],
                   "device_bytes": info["device_bytes"],
                   "baseline": "reference 8x H200, 11.716894 s (docs/validation/lrx13-eight-h200.md:26)"},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": dom["achieved_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_source,
                     "avg_launch_ms": dom["ms_per_step"] / layers,
                     "bytes_per_launch": dom["bytes_per_step"] / layers,
                     "bytes_model": dom["bytes_model"],
                     "all_hot_kernels_gbs": algo_bytes / kernel_s / 1e9 if kernel_s else 0.0,
                     "phases": phases},
    }
    engine.close()
    if not args.no_bitmap:
        result["s11_bitmap"] = s11_bitmap_section(golden_all)
        print(f"[bench] s11 bitmap {result['s11_bitmap']}", file=sys.stderr, flush=True)
    if not args.no_sweep:
        result["sweep"] = sweep_section()
        print(f"[bench] sweep {result['sweep']}", file=sys.stderr, flush=True)
    if not args.no_serve:
        result["serving"] = serving_section(args.serve_jobs)
        print(f"[bench] serving {result['serving']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    serving = result.get("serving", {})
    ok = (exact and result.get("sweep", {}).get("golden_exact", True) and not serving.get("errors")
          and serving.get("perm_tables_exact", True) and serving.get("sl2_tables_exact", True)
          and serving.get("saturation", {}).get("tables_same_as_paced_run", True)
          and result.get("s11_bitmap", {}).get("layer_table_exact", True))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
