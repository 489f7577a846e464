"""Host-side API contracts on the CPU: graph schema, digest, retention rule,
the CayleyPy dispatch layer (with a stand-in CayleyGraph class, as the
reference's tests use duck types) and the command-line host's validation
paths that run before any device is touched."""
import hashlib
import json
import os
import subprocess
import sys
import types
from pathlib import Path
from types import SimpleNamespace
from unittest.mock import patch

import pytest

from multigpubfs_amd import GraphDefinition, from_cayleypy, lrx
from multigpubfs_amd import cayleypy_backend as api
from multigpubfs_amd._native import CLI_PATH

# ------------------------------------------------------------- schema / digest


def test_validation_error_codes():
    # tests/test_graph_definition.py:20-25 of the reference
    for gens, start in [([[0, 0]], [0, 1]), ([[0, 1]], [0]), ([[0, 1]], [0, 1 << 63])]:
        with pytest.raises(ValueError):
            GraphDefinition.permutation(gens, start)
    with pytest.raises(ValueError, match="GRAPH_MATRIX_GENERATOR"):
        GraphDefinition.matrix(2, 1, [([1], 3)], [1, 2])
    with pytest.raises(ValueError, match="GRAPH_MATRIX_MODULUS"):
        GraphDefinition.matrix(1, 1, [([1], 1)], [1])
    with pytest.raises(ValueError, match="GRAPH_SCHEMA"):
        GraphDefinition({"schema": 2})


def test_digest_is_name_independent_and_order_sensitive():
    g = GraphDefinition.permutation([[1, 0, 2], [0, 2, 1]], [0, 1, 2], name="a")
    h = GraphDefinition.permutation([[1, 0, 2], [0, 2, 1]], [0, 1, 2], name="b")
    assert g.digest() == h.digest()
    assert g.inverse_closed()
    assert g.digest() != GraphDefinition.permutation([[0, 2, 1], [1, 0, 2]], [0, 1, 2]).digest()
    assert GraphDefinition.from_dict(json.loads(g.to_json())).digest() == g.digest()
    core = {"action": g.action, "schema": 2, "start": [0, 1, 2]}
    expected = hashlib.sha256(json.dumps(core, sort_keys=True, separators=(",", ":")).encode()).hexdigest()
    assert g.digest() == expected


def test_inverse_closure_rule():
    assert not GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 0, 1, 2]).inverse_closed()
    assert lrx(7).inverse_closed()
    assert not GraphDefinition.matrix(2, 1, [([1, 1, 0, 1], 257)], [256, 2]).inverse_closed()
    # wrapping can fake a modular inverse (test_graph_definition.py:41-45)
    assert not GraphDefinition.matrix(1, 1, [([(1 << 63) - 1], 3)], [2]).inverse_closed()
    for modulus in (0, 4):
        assert GraphDefinition.matrix(1, 1, [([(1 << 63) - 1], modulus)], [2]).inverse_closed()
    assert GraphDefinition.matrix(1, 1, [([100], 101)], [2]).inverse_closed()
    p = 11
    sl2 = [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p), ([1, 0, p - 1, 1], p)]
    assert GraphDefinition.matrix(2, 2, sl2, [1, 0, 0, 1]).inverse_closed()
    assert not GraphDefinition.matrix(2, 2, sl2[:2], [1, 0, 0, 1]).inverse_closed()


def test_cayleypy_adapter_and_scalar_guards():
    x = SimpleNamespace(generators_type=SimpleNamespace(name="PERMUTATION"), generators_permutations=[[1, 0]],
                        central_state=[0, 0], name="coset", generator_names=["x"])
    g = from_cayleypy(x)
    assert g.start == [0, 0] and g.generator_names == ["x"]
    x = SimpleNamespace(generators_type=SimpleNamespace(name="MATRIX"),
                        generators_matrices=[SimpleNamespace(matrix=[[1, 1], [0, 1]], modulo=257)],
                        central_state=[1, 2], name="rect", generator_names=["t"])
    assert from_cayleypy(x).action["generators"] == [{"matrix": [1, 1, 0, 1], "modulo": 257}]
    from multigpubfs_amd.graph_definition import _flatten
    for bad in (0.5, True, "1", float("inf"), float("nan")):
        with pytest.raises(ValueError):
            _flatten(bad)
    assert _flatten([1.0, -2, 3]) == [1, -2, 3]


# --------------------------------------------------------- CayleyPy dispatch

class FakeCayleyGraph:
    """Stand-in for cayleypy.CayleyGraph: the dispatch layer only needs the
    class attribute `bfs` and a `definition` with CayleyPy's attribute names."""

    def __init__(self, n=3):
        self.definition = SimpleNamespace(generators_type=SimpleNamespace(name="PERMUTATION"),
                                          generators_permutations=lrx(n).action["generators"],
                                          central_state=list(range(n)), name=f"lrx-{n}",
                                          generator_names=["L", "R", "X"])

    def bfs(self, **kwargs):
        return ("torch", kwargs)


@pytest.fixture
def fake_cayleypy(monkeypatch, tmp_path):
    module = types.ModuleType("cayleypy")
    module.CayleyGraph = FakeCayleyGraph
    monkeypatch.setitem(sys.modules, "cayleypy", module)
    original = FakeCayleyGraph.bfs
    yield SimpleNamespace(root=tmp_path, original=original)
    try:
        api.disable_native()
    finally:
        FakeCayleyGraph.bfs = original


def fake_native(calls):
    def native(definition, output, **kwargs):
        calls.append((definition, kwargs))
        output.mkdir(parents=True)
        (output / "states.json").write_text(json.dumps(dict(current=[[2, 1, 0]], previous_small=[[0, 2, 1]])))
        return dict(status="COMPLETE", layer_sizes=[1, 2, 3])
    return native


def test_simple_call_and_compact_samples(fake_cayleypy):
    calls = []
    with patch("multigpubfs_amd.launch.native_runtime", lambda *a: Path("/bin/true")):
        api.enable_native(api.NativeOptions(output_root=fake_cayleypy.root))
    with patch.object(api, "run_graph", side_effect=fake_native(calls)):
        result = FakeCayleyGraph().bfs()
    assert result.bfs_completed and result.layer_sizes == [1, 2, 3]
    assert result.diameter() == 2 and result.num_vertices() == 6
    assert set(result.layer_samples) == {1, 2} and result.layers == {}
    assert calls[0][1]["devices"] is None and calls[0][1]["autotune"] is True
    with pytest.raises(KeyError):
        result.last_layer()


def test_disable_restores_and_one_shot_does_not_patch(fake_cayleypy):
    with patch("multigpubfs_amd.launch.native_runtime", lambda *a: Path("/bin/true")):
        api.enable_native()
        api.enable_native()
    api.disable_native()
    assert FakeCayleyGraph.bfs is fake_cayleypy.original
    calls = []
    with patch.object(api, "run_graph", side_effect=fake_native(calls)):
        api.bfs(FakeCayleyGraph(), output=fake_cayleypy.root / "one")
    assert FakeCayleyGraph.bfs is fake_cayleypy.original and len(calls) == 1


def test_foreign_replacement_is_detected(fake_cayleypy):
    with patch("multigpubfs_amd.launch.native_runtime", lambda *a: Path("/bin/true")):
        api.enable_native()
    FakeCayleyGraph.bfs = lambda self, **kw: "other"
    with pytest.raises(RuntimeError):
        api.disable_native()
    FakeCayleyGraph.bfs = api._HOOK.installed
    api.disable_native()


def test_torch_backend_preserves_kwargs_and_unsupported_raise(fake_cayleypy):
    with patch("multigpubfs_amd.launch.native_runtime", lambda *a: Path("/bin/true")):
        api.enable_native()
    assert FakeCayleyGraph().bfs(backend="torch", max_diameter=2) == ("torch", dict(max_diameter=2))
    for kwargs in (dict(return_all_edges=True), dict(max_diameter=2), dict(max_layer_size_to_store=None),
                   dict(misspelled=True)):
        with patch.object(api, "run_graph") as run:
            with pytest.raises(NotImplementedError):
                FakeCayleyGraph().bfs(**kwargs)
            run.assert_not_called()
    # auto falls back to the original for unsupported arguments
    assert FakeCayleyGraph().bfs(backend="auto", max_diameter=2)[0] == "torch"


def test_single_decoded_start_and_multisource(fake_cayleypy):
    calls = []
    with patch.object(api, "run_graph", side_effect=fake_native(calls)):
        api.bfs(FakeCayleyGraph(), output=fake_cayleypy.root / "start", start_states=[[2, 1, 0]])
    assert calls[0][0].start == [2, 1, 0]
    with pytest.raises(NotImplementedError):
        api.bfs(FakeCayleyGraph(), start_states=[[0, 1, 2], [2, 1, 0]])


def test_incomplete_has_no_diameter(fake_cayleypy):
    def incomplete(definition, output, **kwargs):
        report = fake_native([])(definition, output, **kwargs)
        report["status"] = "INCOMPLETE"
        return report
    with patch.object(api, "run_graph", side_effect=incomplete):
        result = api.bfs(FakeCayleyGraph(), output=fake_cayleypy.root / "inc")
    assert not result.bfs_completed and result.completed_depth == 2
    with pytest.raises(RuntimeError):
        result.diameter()


def test_options_validation():
    # cayleypy_backend.py:24-36 of the reference: same fields and codes.
    for kwargs, code in [(dict(memory_preset="fast"), "INVALID_MEMORY_PRESET"),
                         (dict(segment_rows=0), "INVALID_SEGMENT_ROWS"),
                         (dict(segment_rows=(1 << 32) + 1), "INVALID_SEGMENT_ROWS"),
                         (dict(memory_backend="tape"), "INVALID_MEMORY_BACKEND"),
                         (dict(segment_strategy="fast"), "INVALID_SEGMENT_STRATEGY"),
                         (dict(bitmap_owner_page=3), "INVALID_BITMAP_OWNER_PAGE"),
                         (dict(bitmap_owner_page=1 << 21, memory_backend="permutation_bitmap"),
                          "INVALID_BITMAP_OWNER_PAGE"),
                         (dict(bitmap_owner_page=True), "INVALID_BITMAP_OWNER_PAGE"),
                         (dict(bitmap_owner_page=64), "BITMAP_OWNER_PAGE_REQUIRES_BITMAP")]:
        with pytest.raises(ValueError, match=code):
            api.NativeOptions(**kwargs)
    ok = api.NativeOptions(bitmap_owner_page=64, memory_backend="permutation_bitmap", segment_strategy="warp")
    assert ok.bitmap_owner_page == 64 and api.NativeOptions(bitmap_owner_page=0).bitmap_owner_page == 0


def test_segment_strategy_default_follows_environment(monkeypatch):
    monkeypatch.setenv("MGBFS_SEGMENT_STRATEGY", "batch")
    assert api.NativeOptions().segment_strategy == "batch"
    monkeypatch.setenv("MGBFS_SEGMENT_STRATEGY", "bogus")
    with pytest.raises(ValueError, match="INVALID_SEGMENT_STRATEGY"):
        api.NativeOptions()


def test_options_reach_run_graph(fake_cayleypy):
    calls = []
    options = api.NativeOptions(devices=(0,), max_seconds=7, memory_backend="permutation_bitmap",
                                bitmap_owner_page=1024, segment_strategy="atomic", autotune=False)
    with patch.object(api, "run_graph", side_effect=fake_native(calls)):
        api.bfs(FakeCayleyGraph(), native_options=options, output=fake_cayleypy.root / "opts")
    kwargs = calls[0][1]
    assert kwargs["devices"] == [0] and kwargs["max_seconds"] == 7 and kwargs["autotune"] is False
    assert kwargs["bitmap_owner_page"] == 1024 and kwargs["segment_strategy"] == "atomic"
    assert kwargs["memory_backend"] == "permutation_bitmap"


def test_run_graph_argument_validation(tmp_path):
    from multigpubfs_amd import run_graph
    g = lrx(4)
    for kwargs, code in [(dict(device=0, devices=[0]), "DEVICE_SELECTION_CONFLICT"), (dict(device=-1), "INVALID_DEVICE"),
                         (dict(devices=[0, 0]), "INVALID_DEVICES"), (dict(devices=[0, 1]), "SINGLE_DEVICE_BUILD"),
                         (dict(capacity=0), "INVALID_CAPACITY"), (dict(backend="cuda"), "INVALID_BACKEND")]:
        with pytest.raises(ValueError, match=code):
            run_graph(g, tmp_path / "out", **kwargs)
    m = GraphDefinition.matrix(1, 1, [([2], 5)], [1])
    with pytest.raises(ValueError, match="BITMAP_REQUIRES_PERMUTATION_ACTION"):
        run_graph(m, tmp_path / "out", memory_backend="permutation_bitmap")
    with pytest.raises(ValueError, match="ARCHIVE_REQUIRES_PERMUTATION_ACTION"):
        run_graph(m, tmp_path / "out", archive=True)
    with pytest.raises(ValueError, match="INVALID_ARCHIVE"):
        run_graph(lrx(4), tmp_path / "out", archive="yes")


@pytest.mark.parametrize("kwargs,code", [
    (dict(shards=0), "INVALID_SHARDS"),
    (dict(shards=4097), "INVALID_SHARDS"),
    (dict(segment_rows=0), "INVALID_SEGMENT_ROWS"),
    (dict(segment_rows=(1 << 32) + 1), "INVALID_SEGMENT_ROWS"),
    (dict(bitmap_owner_page=3), "INVALID_BITMAP_OWNER_PAGE"),
    (dict(bitmap_owner_page=64), "BITMAP_OWNER_PAGE_REQUIRES_BITMAP"),
    (dict(segment_strategy="fast"), "INVALID_SEGMENT_STRATEGY"),
    (dict(segment_strategy="warp"), "SEGMENT_STRATEGY_REQUIRES_SEGMENTED_BACKEND"),
    (dict(memory_preset="huge"), "INVALID_MEMORY_PRESET"),
    (dict(memory_preset="max_capacity"), "MEMORY_PRESET_UNSUPPORTED_FOR_LEGACY"),
    (dict(peer_transport="nvlink"), "INVALID_PEER_TRANSPORT"),
    (dict(peer_transport="lsa"), "LSA_REQUIRES_SPECIALIZED_BACKEND"),
    (dict(transport="half"), "INVALID_TRANSPORT"),
    (dict(candidate_order="sorted"), "INVALID_CANDIDATE_ORDER"),
    (dict(backend="shard_ab_hash"), "SPECIALIZED_BACKEND_NOT_PROVIDED"),
    (dict(capacity=0x7FFFFFFF), "INVALID_CAPACITY"),  # legacy arenas use 31-bit rows
])
def test_run_graph_reference_keywords_validate_like_the_reference(tmp_path, kwargs, code):
    # launch.py:151-200: same keyword names, validation order and codes.
    from multigpubfs_amd import run_graph
    with pytest.raises(ValueError, match=code):
        run_graph(lrx(4), tmp_path / "out", **kwargs)


# ------------------------------------------------------ command-line host

def cli(*args):
    return subprocess.run([str(CLI_PATH), *map(str, args)], capture_output=True, text=True)


def last_error(proc):
    return json.loads(proc.stderr.strip().splitlines()[-1])


def test_cli_digest_matches_python_and_retention_rule(tmp_path):
    graphs = [lrx(13), GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 0, 1, 2], name="é"),
              GraphDefinition.matrix(2, 2, [([1, 1, 0, 1], 7), ([1, 6, 0, 1], 7)], [1, 0, 0, 1]),
              GraphDefinition.matrix(1, 1, [([(1 << 63) - 1], 3)], [-5])]
    for i, g in enumerate(graphs):
        path = tmp_path / f"g{i}.json"
        path.write_text(g.to_json())
        proc = cli("graph-digest", path)
        assert proc.returncode == 0, proc.stderr
        doc = json.loads(proc.stdout)
        assert doc["graph_digest"] == g.digest()
        assert doc["inverse_closed"] == g.inverse_closed()


def test_cli_rejects_invalid_graphs_and_existing_output(tmp_path):
    bad = tmp_path / "bad.json"
    for doc, code in [({"schema": 2}, "GRAPH_SCHEMA"),
                      (dict(json.loads(lrx(3).to_json()), start=[0, 1]), "GRAPH_PERMUTATION_SHAPE")]:
        bad.write_text(json.dumps(doc))
        proc = cli("graph", bad, tmp_path / "o1")
        assert proc.returncode != 0 and last_error(proc) == {"status": "ERROR", "error": code}
    good = tmp_path / "good.json"
    good.write_text(lrx(3).to_json())
    (tmp_path / "exists").mkdir()
    proc = cli("graph", good, tmp_path / "exists")
    assert last_error(proc)["error"] == "CLI_GRAPH_OUTPUT_EXISTS"
    assert last_error(cli("graph", good, tmp_path / "o2", "--bogus", "1"))["error"] == "CLI_GRAPH_UNKNOWN_OPTION"
    assert last_error(cli("frobnicate"))["error"] == "CLI_UNKNOWN_COMMAND"


def test_cli_key_info():
    proc = cli("key-info")
    doc = json.loads(proc.stdout)
    assert proc.returncode == 0 and "graph" in doc["commands"] and doc["device_arch"] == "gfx950"


# ------------------------------------------------------ sweep / serving inputs

def test_sweep_instances_are_deterministic_and_shaped():
    from multigpubfs_amd.sweep import make_sweep
    a, b = make_sweep(32), make_sweep(32)
    assert a == b
    for i, g in enumerate(a):
        gd = GraphDefinition(g)
        n, k = gd.state_elements, gd.generator_count
        assert n in (9, 10)
        assert (2 <= k <= 5) if i % 2 else (4 <= k <= 10 and gd.inverse_closed())


def test_graph_descriptor_is_built_once_and_overrides_copy_it():
    from multigpubfs_amd._native import desc_with_inverse_closed, graph_desc
    from multigpubfs_amd.sweep import make_sweep, prepare
    graphs = make_sweep(4)
    prepared = prepare(graphs)
    for g, gd in zip(graphs, prepared):
        d = graph_desc(gd)
        assert graph_desc(gd) is d  # kept on the validated graph
        assert graph_desc(g) is not d  # a plain dict is validated afresh
        n, k = gd.state_elements, gd.generator_count
        assert (d.kind, d.elements, d.rows, d.cols, d.generators) == (0, n, n, 1, k)
        assert d.inverse_closed == int(gd.inverse_closed())
        assert list(d.start[:n]) == gd.start
        assert list(d.permutations[:n * k]) == [v for p in gd.action["generators"] for v in p]
        flipped = desc_with_inverse_closed(d, not gd.inverse_closed())
        assert flipped.inverse_closed == 1 - d.inverse_closed
        assert d.inverse_closed == int(gd.inverse_closed())  # the cached one is untouched
        assert list(flipped.permutations[:n * k]) == list(d.permutations[:n * k])
    assert prepare(prepared)[0] is prepared[0]


@pytest.mark.parametrize("argv, env", [(["--gpus", "2"], {}), ([], {"WORLD_SIZE": "8"})])
def test_bench_refuses_more_than_one_gpu(argv, env):
    # One exact BFS runs on one GPU: a multi-GPU request exits with a message
    # before touching a device.
    bench = Path(__file__).resolve().parent.parent / "bench.py"
    out = subprocess.run([sys.executable, str(bench), *argv], capture_output=True, text=True, timeout=120,
                         env={**os.environ, **env})
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert "single-GPU build" in line["error"]


def test_serving_job_stream_mix_and_rate():
    from multigpubfs_amd.serving import make_jobs
    jobs = make_jobs(2000)
    assert jobs == make_jobs(2000)
    perm = [j for j in jobs if j["graph"]["action"]["kind"] == "permutation"]
    assert 0.65 < len(perm) / len(jobs) < 0.75
    assert 40 * 0.9 < len(jobs) / jobs[-1]["arrival"] < 40 * 1.1
    for j in jobs:
        gd = GraphDefinition(j["graph"])
        assert gd.inverse_closed()
        if gd.action["kind"] == "permutation":
            assert 8 <= gd.state_elements <= 12 and 2 <= gd.generator_count <= 4
        else:
            p = gd.action["generators"][0]["modulo"]
            assert 101 <= p <= 521 and all(p % d for d in range(2, p))


def test_graph_plan_from_inventory_needs_no_device(tmp_path):
    # graph-plan sizes rank plans from a supplied inventory (free bytes less
    # the 2 GiB reserve), so it runs on a host without a GPU.
    gib = 1 << 30
    inventory = [{"device": 0, "free_bytes": 100 * gib, "total_bytes": 288 * gib},
                 {"device": 1, "free_bytes": 50 * gib, "total_bytes": 288 * gib}]
    inv = tmp_path / "inventory.json"
    inv.write_text(json.dumps(inventory))
    p = 13
    sl2 = GraphDefinition.matrix(2, 2, [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p),
                                        ([1, 0, p - 1, 1], p)], [1, 0, 0, 1])
    directed = GraphDefinition.permutation([[1, 2, 3, 4, 5, 6, 0], [1, 0, 3, 2, 5, 4, 6]], list(range(7)))
    # Unipotent upper and lower shears without their inverses: directed.
    shears = GraphDefinition.matrix(2, 2, [([1, 1, 0, 1], p), ([1, 0, 1, 1], p)], [1, 0, 0, 1])
    for g, backend, history in ((lrx(13), "sorted", 3), (directed, "permutation_bitmap", 1), (sl2, "segmented", 3),
                                (shears, "generic", 1)):
        path = tmp_path / f"{backend}.json"
        path.write_text(g.to_json())
        proc = cli("graph-plan", path, inv, 4)
        assert proc.returncode == 0, proc.stderr
        doc = json.loads(proc.stdout)
        assert doc["graph_digest"] == g.digest() and doc["devices"] == [0, 1]
        plans = doc["rank_plans"]
        assert [q["backend"] for q in plans] == [backend, backend]
        assert all(q["history_layers"] == history and q["world"] == 2 and q["shards"] == 4 for q in plans)
        for q, item in zip(plans, inventory):
            assert q["device_bytes"] <= item["free_bytes"] - 2 * gib
        assert doc["plan"] == min(plans, key=lambda q: q["capacity"])
        cuts = doc["owner_cuts"]
        assert cuts[0] == 0 and cuts[-1] == 1 << 32 and cuts == sorted(cuts) and len(cuts) == 3
    # Generic arena plan: the reference formula elements*capacity*state_bytes*2
    # + capacity*8 + next_pow2(2 capacity)*8 + 24 (generic_memory.rs:36-41),
    # plus the generator tables; capacity capped by the bound p^4.
    q = doc["plan"]
    assert q["capacity"] == p ** 4 and q["table_slots"] == 1 << 16 and q["state_bytes"] == 8
    tables = 2 * 4 * 8 + 2 * 4
    assert q["device_bytes"] == 4 * p ** 4 * 8 * 2 + p ** 4 * 8 + (1 << 16) * 8 + 24 + tables
    assert q["generator_bytes"] == 2 * (4 * 8 + 4)
    bad = tmp_path / "bad.json"
    for entry, code in (({"device": 0, "free_bytes": 3 * gib, "total_bytes": 2 * gib}, "GRAPH_GLOBAL_MEMORY_RANGE"),
                        ({"device": 0, "free_bytes": gib, "total_bytes": 2 * gib}, "GENERIC_DISTRIBUTED_HEADROOM"),
                        ({"free_bytes": gib, "total_bytes": 2 * gib}, "GRAPH_GLOBAL_DEVICE")):
        bad.write_text(json.dumps([entry]))
        assert last_error(cli("graph-plan", tmp_path / "generic.json", bad, 1))["error"] == code
    assert last_error(cli("graph-plan", tmp_path / "generic.json", inv, 1, "1000", "2000"))["error"] == \
        "REQUESTED_BATCH_EXCEEDS_ADMISSION"
