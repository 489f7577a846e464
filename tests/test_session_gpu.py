"""Resident graph workers on the GPU: run_graph inside persistent_graph_workers
gives the same reports as the one-shot executable, reuses the engine for a
repeated graph, answers inventory / release-memory, and a cancel file stops a
job at a layer boundary."""
import json
from pathlib import Path

import pytest

from multigpubfs_amd import lrx, persistent_graph_workers, run_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parents[1] / "multigpubfs_amd" / "bin" / "mgbfs"


def test_resident_runs_match_one_shot_and_reuse_the_engine(tmp_path):
    ref = O.packed_bfs(9)["layers"]
    one_shot = run_graph(lrx(9), tmp_path / "one", device=0)
    with persistent_graph_workers() as pool:
        a = run_graph(lrx(9), tmp_path / "a", device=0)
        first = pool.workers[0].job.response
        b = run_graph(lrx(9), tmp_path / "b", device=0)
        second = pool.workers[0].job.response
        c = run_graph(lrx(8), tmp_path / "c", device=0)
        inv = pool.inventory(EXE, [0])
        pool.release_memory()
        d = run_graph(lrx(9), tmp_path / "d", device=0)
        after = pool.workers[0].job.response
    for r in (one_shot, a, b, d):
        assert r["layer_sizes"] == ref and r["status"] == "COMPLETE"
    assert c["layer_sizes"] == O.packed_bfs(8)["layers"]
    # Sample order may differ between runs; the sampled states may not.
    current = [sorted(json.loads((tmp_path / x / "states.json").read_text())["current"]) for x in ("a", "b")]
    assert current[0] == current[1]
    assert first["result"]["engine_reused"] is False and second["result"]["engine_reused"] is True
    assert after["result"]["engine_reused"] is False and after["cache"]["cached_engines"] == 1
    assert inv[0]["total_bytes"] > 0 and inv[0]["sm_count"] > 0


def test_cancel_file_stops_a_resident_job(tmp_path):
    graph = tmp_path / "g.json"
    graph.write_text(lrx(11).to_json())
    with persistent_graph_workers() as pool:
        worker = pool.worker_for(EXE, 0)
        (worker.root / "cancel-00000000").touch()  # present before the job starts
        job = worker.submit("graph", [str(graph), str(tmp_path / "out")])
        assert job.wait(60) == 0
        report = json.loads((tmp_path / "out" / "report.json").read_text())
    assert report["status"] == "INCOMPLETE" and report["reason"] == "CANCELLED"
    assert report["layer_sizes"] == [1]
