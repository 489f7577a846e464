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


def test_latency_metrics_reported():
    jobs = small_jobs(20)
    report = serve(jobs, max_slots=8, vram_budget=32 * 10 ** 9)
    assert report["p50_latency_s"] <= report["p99_latency_s"] <= report["max_latency_s"]
    assert report["jobs_per_s"] > 0
