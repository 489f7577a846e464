"""End-to-end drop-in path: run_graph -> `mgbfs graph` -> report.json /
states.json (crates/mgbfs-runtime/src/generic_run.rs:35-42 output contract)."""
import hashlib
import json
import math

import pytest

from multigpubfs_amd import GraphDefinition, lrx, run_graph
from multigpubfs_amd.launch import NativeError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("memory_backend", ["legacy", "segmented", "permutation_bitmap", "generic"])
def test_lrx8_report_and_states_contract(tmp_path, memory_backend):
    g = lrx(8)
    out = tmp_path / "run"
    report = run_graph(g, out, device=0, memory_backend=memory_backend)
    ref = O.packed_bfs(8, keep_keys=True)
    assert report["status"] == "COMPLETE" and report["reason"] == "FRONTIER_EXHAUSTED"
    assert report["layer_sizes"] == ref["layers"]
    assert report["graph_digest"] == g.digest()
    raw = (out / "states.json").read_bytes()
    assert report["states_sha256"] == hashlib.sha256(raw).hexdigest()
    states = json.loads(raw)
    assert states["schema"] == 2 and states["state_encoding"] == "signed_int64_vectors"
    assert states["current_sample_limit"] == 1000
    # last layer (1 state) complete; previous layer < 1000 rows is retained
    assert sorted(O.pack(r) for r in states["current"]) == sorted(ref["keys"][-1])
    assert sorted(O.pack(r) for r in states["previous_small"]) == sorted(ref["keys"][-2])


def test_matrix_graph_through_cli(tmp_path):
    p = 7
    gens = [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p), ([1, 0, p - 1, 1], p)]
    g = GraphDefinition.matrix(2, 2, gens, [1, 0, 0, 1])
    report = run_graph(g, tmp_path / "m", device=0)
    expected = O.c_exact_layers(g.to_dict(), 400)
    assert report["layer_sizes"] == [len(x) for x in expected]
    assert sum(report["layer_sizes"]) == p * (p * p - 1)
    states = json.loads((tmp_path / "m" / "states.json").read_text())
    assert sorted(map(tuple, states["current"])) == sorted(map(tuple, expected[-1]))


def test_large_last_layer_sample_is_subset(tmp_path):
    g = lrx(10)
    report = run_graph(g, tmp_path / "s10", device=0)
    states = json.loads((tmp_path / "s10" / "states.json").read_text())
    ref = O.packed_bfs(10, keep_keys=True)
    assert report["layer_sizes"] == ref["layers"]
    assert len(states["current"]) == min(1000, ref["layers"][-1])
    assert set(O.pack(r) for r in states["current"]) <= set(ref["keys"][-1])
    assert states["previous_small"] == [] or ref["layers"][-2] < 1000


@pytest.mark.parametrize("memory_backend", ["generic", "permutation_bitmap", "segmented"])
def test_resource_stop_is_incomplete_with_prefix(tmp_path, memory_backend):
    # capacity=50 stops LRX on S8 mid-run.  The bitmap backend writes the
    # failed layer's first chunk over the buffer that held the previous
    # layer; states.json must still hold the oracle's last two finished layers.
    out = tmp_path / "r"
    report = run_graph(lrx(8), out, device=0, capacity=50, memory_backend=memory_backend)
    assert report["status"] == "INCOMPLETE" and report["reason"].startswith("RESOURCE_")
    ref = O.packed_bfs(8, keep_keys=True)
    sizes = report["layer_sizes"]
    assert len(sizes) >= 3 and sizes == ref["layers"][:len(sizes)]
    states = json.loads((out / "states.json").read_text())
    depth = len(sizes) - 1
    assert sorted(O.pack(r) for r in states["current"]) == sorted(ref["keys"][depth])
    assert sorted(O.pack(r) for r in states["previous_small"]) == sorted(ref["keys"][depth - 1])


def test_output_directory_must_not_exist(tmp_path):
    (tmp_path / "x").mkdir()
    with pytest.raises(NativeError) as err:
        run_graph(lrx(5), tmp_path / "x", device=0)
    assert err.value.code == "CLI_GRAPH_OUTPUT_EXISTS"


@pytest.mark.parametrize("memory_backend", ["segmented", "permutation_bitmap"])
def test_max_capacity_preset_runs_exact(tmp_path, memory_backend):
    # memory_preset="max_capacity" keeps a 256 MiB VRAM reserve instead of 2 GiB.
    report = run_graph(lrx(9), tmp_path / "out", device=0, memory_backend=memory_backend,
                       memory_preset="max_capacity")
    assert report["layer_sizes"] == O.packed_bfs(9)["layers"] and report["status"] == "COMPLETE"


@pytest.mark.parametrize("memory_backend", ["segmented", "permutation_bitmap", "generic"])
def test_archive_samples_every_layer(tmp_path, memory_backend):
    from multigpubfs_amd import archive as A
    from multigpubfs_amd.launch import ARCHIVE_NAME
    n = 9
    out = tmp_path / "a"
    report = run_graph(lrx(n), out, device=0, memory_backend=memory_backend, archive=True)
    ref = O.packed_bfs(n, keep_keys=True)
    assert report["layer_sizes"] == ref["layers"]
    assert report["archive"]["depths"] == len(ref["layers"]) and not report["archive"]["terminal_replaced"]
    data = (out / ARCHIVE_NAME).read_bytes()
    assert A.verify(data)["depths"] == len(ref["layers"])
    walked = list(A.layers(data))
    assert [d for d, _ in walked] == list(range(len(ref["layers"])))
    for d, raw in walked:
        keys = [O.pack(list(raw[i:i + n])) for i in range(0, len(raw), n)]
        assert len(keys) == min(1000, ref["layers"][d]) and len(set(keys)) == len(keys)
        assert set(keys) <= set(ref["keys"][d])


def test_archive_of_resource_stop_commits_finished_layers(tmp_path):
    report = run_graph(lrx(8), tmp_path / "r", device=0, capacity=50, memory_backend="generic", archive=True)
    assert report["status"] == "INCOMPLETE"
    assert report["archive"]["depths"] == len(report["layer_sizes"])
