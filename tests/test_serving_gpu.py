"""Continuous-batching serving loop: every job's layer table is the same in
the mix as alone (and as the oracle's, where the oracle is quick)."""
import math

import pytest

from multigpubfs_amd.serving import make_jobs, run_alone, serve
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_jobs(count):
    jobs = []
    for j in make_jobs(count=400, rate=400.0):
        a = j["graph"]["action"]
        if a["kind"] == "permutation" and a["degree"] <= 10:
            jobs.append(j)
        elif a["kind"] == "matrix" and a["generators"][0]["modulo"] <= 150:
            jobs.append(j)
        if len(jobs) == count:
            break
    return jobs


def test_results_identical_alone_and_in_the_mix():
    jobs = small_jobs(48)
    report = serve(jobs, max_slots=16, vram_budget=64 * 10 ** 9, layers_per_step=2)
    assert report["jobs"] == len(jobs) and report["errors"] == 0
    assert report["peak_concurrent_jobs"] > 1  # jobs really overlapped
    for job in jobs:
        mixed = report["results"][job["id"]]["layer_sizes"]
        alone, error = run_alone(job)
        assert error == 0 and mixed == alone
        a = job["graph"]["action"]
        if a["kind"] == "permutation" and a["degree"] <= 9:
            assert mixed == O.packed_bfs(a["degree"], a["generators"])["layers"]
        if a["kind"] == "matrix":
            p = a["generators"][0]["modulo"]
            assert sum(mixed) == p * (p * p - 1)


def test_large_jobs_identical_alone_and_in_the_mix():
    # The stream's first S12 jobs with three and four generators, S11 jobs,
    # and the first SL(2, Z_p) jobs with p > 500, served together beside
    # small jobs; each table equals the same job run alone, S12/S11 tables
    # also the plain engine's, SL(2,p) its order p(p^2 - 1).
    import json
    from pathlib import Path
    from multigpubfs_amd import GraphDefinition, NativeEngine
    stream = make_jobs(400, rate=400.0)
    picked, seen = [], set()
    for j in stream:
        a = j["graph"]["action"]
        if a["kind"] == "permutation":
            key = (a["degree"], len(a["generators"]))
            if (a["degree"] >= 11 and key not in seen) or (a["degree"] <= 9 and len(picked) < 12):
                seen.add(key)
                picked.append(j)
        elif a["generators"][0]["modulo"] > 500 and ("p", a["generators"][0]["modulo"]) not in seen:
            seen.add(("p", a["generators"][0]["modulo"]))
            picked.append(j)
    degrees = {(j["graph"]["action"].get("degree"), len(j["graph"]["action"]["generators"])) for j in picked}
    assert {(12, 4), (12, 3), (11, 4)} <= degrees
    primes = [j["graph"]["action"]["generators"][0]["modulo"] for j in picked
              if j["graph"]["action"]["kind"] == "matrix"]
    assert max(primes) > 500
    print(f"serving {len(picked)} jobs: degrees {sorted(degrees, key=str)}, primes {primes}", flush=True)
    report = serve(picked, max_slots=32, vram_budget=160 * 10 ** 9, layers_per_step=4)
    assert report["jobs"] == len(picked) and report["errors"] == 0 and report["peak_concurrent_jobs"] > 1
    golden = json.loads((Path(__file__).parent / "golden" / "serve2000_small.json").read_text())["tables"]
    for job in picked:
        mixed = report["results"][job["id"]]["layer_sizes"]
        alone, error = run_alone(job)
        assert error == 0 and mixed == alone, job["id"]
        a = job["graph"]["action"]
        if a["kind"] == "matrix":
            p = a["generators"][0]["modulo"]
            assert sum(mixed) == p * (p * p - 1)
        elif a["degree"] <= 10:
            assert mixed == golden[str(job["id"])]
        else:
            with NativeEngine(GraphDefinition(job["graph"]), backend="sorted") as e:
                e.run()
                assert mixed == e.layer_sizes()


def test_latency_metrics_reported():
    jobs = small_jobs(20)
    report = serve(jobs, max_slots=8, vram_budget=32 * 10 ** 9)
    assert report["p50_latency_s"] <= report["p99_latency_s"] <= report["max_latency_s"]
    assert report["jobs_per_s"] > 0
