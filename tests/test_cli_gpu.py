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


# ------------------------------------------------------------------ admission + rank protocol

import subprocess  # noqa: E402

from multigpubfs_amd._native import CLI_PATH  # noqa: E402

INVENTORY_KEYS = {"device", "free_bytes", "total_bytes", "sm_count", "compute_major", "compute_minor", "l2_bytes",
                  "warp_size", "threads_per_sm", "threads_per_block", "shared_bytes_per_sm"}
PLAN_KEYS = {"generator_backend", "history_algorithm", "owner_lanes", "sorted_owner_bytes",
             "driver_graph_reserve_bytes", "parent_transport", "sort_candidates", "history_layers", "state_bytes",
             "elements", "world", "shards", "capacity", "batch", "table_layout", "queue_capacity", "table_slots",
             "generator_bytes", "device_bytes"}


def mgbfs(*args, check=True):
    proc = subprocess.run([str(CLI_PATH), *map(str, args)], capture_output=True, text=True, timeout=300)
    if check and proc.returncode:
        raise AssertionError(proc.stderr)
    return proc


def error_code(proc):
    assert proc.returncode != 0
    return json.loads(proc.stderr.strip().splitlines()[-1])["error"]


def directed_graph():
    # Two random directed permutations on S7 (no inverses): full history.
    return GraphDefinition.permutation([[1, 2, 3, 4, 5, 6, 0], [1, 0, 3, 2, 5, 4, 6]], list(range(7)))


def sl2(p):
    return GraphDefinition.matrix(2, 2, [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p),
                                         ([1, 0, p - 1, 1], p)], [1, 0, 0, 1])


@pytest.mark.parametrize("name,graph,history,backend", [
    ("lrx8", lrx(8), 3, "sorted"), ("directed7", directed_graph(), 1, "permutation_bitmap"),
    ("sl2_13", sl2(13), 3, "segmented"),
    ("shears_13", GraphDefinition.matrix(2, 2, [([1, 1, 0, 1], 13), ([1, 0, 1, 1], 13)], [1, 0, 0, 1]), 1, "generic")])
def test_graph_info_admission_contract(tmp_path, name, graph, history, backend):
    # generic_distributed_run.rs:23-73: inventory per device, a rank plan per
    # device, the smallest as `plan`, owner cut points, reserve policy.
    path = tmp_path / "graph.json"
    path.write_text(graph.to_json())
    info = json.loads(mgbfs("graph-info", path, "0", "1").stdout)
    assert info["graph_digest"] == graph.digest() and info["devices"] == [0]
    assert set(info["inventory"][0]) >= INVENTORY_KEYS
    inv = info["inventory"][0]
    assert inv["warp_size"] == 64 and (inv["compute_major"], inv["compute_minor"]) == (9, 5)
    plan = info["plan"]
    assert set(plan) >= PLAN_KEYS and info["rank_plans"] == [plan]
    assert plan["history_layers"] == history and plan["backend"] == backend
    assert plan["world"] == 1 and plan["shards"] == 1 and plan["elements"] == len(graph.to_dict()["start"])
    assert 0 < plan["batch"] <= plan["capacity"]
    assert plan["queue_capacity"] == plan["batch"] * len(graph.to_dict()["action"]["generators"])
    assert 0 < plan["device_bytes"] <= inv["free_bytes"]
    if backend == "segmented":
        assert plan["history_algorithm"] == "SEGMENTED_HASH" and plan["retained_banks"] == 3
    if backend == "generic":
        assert plan["table_slots"] >= 2 * plan["capacity"] and plan["history_algorithm"] == "HASH"
        # Capped by the state-space bound p^(rows*cols) of a single modulus
        # (generic_memory.rs:82-108).
        assert plan["capacity"] == plan["state_bound"] == 13 ** 4
    assert info["owner_cuts"] == [0, 1 << 32]
    assert info["additional_reserve_bytes"] == 2 << 30


def test_graph_info_requested_capacity_and_batch(tmp_path):
    path = tmp_path / "graph.json"
    path.write_text(sl2(13).to_json())
    info = json.loads(mgbfs("graph-info", path, "auto", "2", "5000", "100").stdout)
    assert info["plan"]["capacity"] == 5000 and info["plan"]["batch"] == 100 and info["plan"]["shards"] == 2
    assert info["owner_cuts"] is None
    assert error_code(mgbfs("graph-info", path, "0", "1", "auto", "0", check=False)) == "CLI_GRAPH_BATCH"
    assert error_code(mgbfs("graph-info", path, "0", "1", "100", "101", check=False)) == \
        "REQUESTED_BATCH_EXCEEDS_ADMISSION"
    assert error_code(mgbfs("graph-info", path, "0", "1", str(1 << 40), check=False)) == \
        "REQUESTED_CAPACITY_EXCEEDS_ADMISSION"
    assert error_code(mgbfs("graph-info", path, "0,0", "1", check=False)) == "CLI_GRAPH_DEVICE_SELECTION"
    assert error_code(mgbfs("graph-info", path, "0", check=False)) == "CLI_GRAPH_INFO_ARGUMENTS"


def reference_launch(graph, output, native=CLI_PATH, max_seconds=600):
    """The one-device path of the reference launcher (multigpubfs/launch.py:
    311-413) as a foreign caller: graph-info admission, then either the
    direct `graph` command (history_layers 1, HASH) or one graph-rank worker
    and the launcher's receipt checks and aggregation."""
    output.parent.mkdir(parents=True, exist_ok=True)
    definition = output.parent / f"{output.name}-graph.json"
    definition.write_text(graph.to_json())
    admission = json.loads(mgbfs("graph-info", definition, "auto", "1").stdout)
    devices, plan = admission["devices"], admission["plan"]
    assert admission["graph_digest"] == graph.digest()
    if plan.get("history_layers", 1) == 1 and plan.get("history_algorithm", "HASH") == "HASH" and len(devices) == 1:
        mgbfs("graph", definition, output, "--device", devices[0], "--seconds", max_seconds)
        return "direct", json.loads((output / "report.json").read_text())
    output.mkdir()
    (output / "launch.json").write_text(json.dumps(admission))
    env_phase = output / "rank-0.phase.json"
    proc = subprocess.run([str(native), "graph-rank", str(definition), str(output / "launch.json"), "0",
                           str(output.parent / "nccl-id"), str(output / "rank-0"), str(max_seconds)],
                          capture_output=True, text=True, timeout=600,
                          env=dict(__import__("os").environ, RANK="0", WORLD_SIZE="1", LOCAL_RANK=str(devices[0]),
                                   MGBFS_JOB_PHASE_PATH=str(env_phase)))
    assert proc.returncode == 0, proc.stderr
    part = json.loads((output / "rank-0" / "report.json").read_text())
    raw_states = (output / "rank-0" / "states.json").read_bytes()
    assert hashlib.sha256(raw_states).hexdigest() == part["states_sha256"]
    assert part["graph_digest"] == graph.digest() and part["status"] in ("COMPLETE", "INCOMPLETE")
    # launch.py:399-400: rank geometry must match the admission.
    assert part["rank"] == 0 and part["world"] == 1 and part["device"] == devices[0]
    assert part["plan"] == admission.get("rank_plans", [plan])[0]
    snap = json.loads(raw_states)
    assert len(snap["current"]) <= 1000 and len(snap["previous_small"]) < 1000
    phase = json.loads(env_phase.read_text())
    assert phase["schema"] == 1 and phase["phase"] == "done" and "bfs" in phase["entered"]
    return "rank", part


@pytest.mark.parametrize("name,graph,route", [("lrx9", lrx(9), "rank"), ("directed7", directed_graph(), "rank"),
                                              ("sl2_11", sl2(11), "rank"),
                                              ("directed_matrix", GraphDefinition.matrix(1, 1, [([2], 5), ([3], 7)],
                                                                                         [4]), "direct")])
def test_reference_launcher_protocol_on_one_device(tmp_path, name, graph, route):
    taken, report = reference_launch(graph, tmp_path / "out")
    assert taken == route
    d = graph.to_dict()
    bound = math.prod(range(1, 10)) if name == "lrx9" else 100000
    expected = [len(x) for x in O.c_exact_layers(d, bound)]
    assert report["layer_sizes"] == expected and report["status"] == "COMPLETE"


def test_graph_rank_refuses_worlds_above_one(tmp_path):
    graph = lrx(5)
    path = tmp_path / "graph.json"
    path.write_text(graph.to_json())
    plan = {"world": 2, "capacity": 120, "batch": 120, "backend": "sorted"}
    launch = {"graph_digest": graph.digest(), "devices": [0, 1], "plan": plan, "rank_plans": [plan, plan]}
    (tmp_path / "launch.json").write_text(json.dumps(launch))
    proc = mgbfs("graph-rank", path, tmp_path / "launch.json", "0", tmp_path / "id", tmp_path / "r0", "0", check=False)
    assert error_code(proc) == "GRAPH_RANK_WORLD_UNSUPPORTED"
    launch["graph_digest"] = "0" * 64
    (tmp_path / "launch.json").write_text(json.dumps(launch))
    proc = mgbfs("graph-rank", path, tmp_path / "launch.json", "0", tmp_path / "id", tmp_path / "r1", "0", check=False)
    assert error_code(proc) == "GRAPH_LAUNCH_IDENTITY_OR_GEOMETRY"
