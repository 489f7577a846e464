"""Parity of the packed HIP backends with the CPU oracle (through the C ABI).

Layer tables must be bit-exact; for small graphs every layer's state set is
compared (the reference compares full layer state sets on small runs,
tests/segmented_single_gpu.py:40-43) and for large ones the layer table is
compared with the oracle and with the reference's golden tables.
"""
import json
import math
import random
from pathlib import Path

import pytest

from multigpubfs_amd import GraphDefinition, NativeEngine, lrx
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((Path(__file__).parent / "golden" / "lrx_layers.json").read_text())
BACKENDS = ["permutation_bitmap", "sorted"]
NAMES = {"permutation_bitmap": "PERMUTATION_BITMAP_LEHMER", "sorted": "SORTED_THREE_BANK_LDS_UNIQUE"}


def run_all_layers(engine, keep_keys=False):
    keys = [sorted(engine.layer_keys())] if keep_keys else None
    while True:
        status = engine.advance()
        if status == 1:
            break
        assert status == 0, (status, engine.fatal)
        if keep_keys:
            keys.append(sorted(engine.layer_keys()))
    return engine.layer_sizes(), keys


def random_generators(rng, n, k, inverse_closed):
    gens = []
    for _ in range(k):
        g = list(range(n))
        rng.shuffle(g)
        gens.append(g)
    if inverse_closed:
        for g in list(gens):
            h = [0] * n
            for i, v in enumerate(g):
                h[v] = i
            if h not in gens:
                gens.append(h)
    return gens


@pytest.mark.parametrize("backend", BACKENDS)
def test_s8_lrx_layer_sets_bit_exact(backend):
    ref = O.packed_bfs(8, keep_keys=True)
    with NativeEngine(lrx(8), backend=backend) as e:
        assert e.backend_name == NAMES[backend]
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]
    assert sum(sizes) == 40320


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7])
def test_sorted_small_degrees_layer_sets_bit_exact(n):
    # One sub-bucket per layer for most of these: the fine buckets, the
    # start bank's 2^20 fine buckets and the slices of a bank whose
    # partition is much finer than the layer's sub-buckets.
    ref = O.packed_bfs(n, keep_keys=True)
    with NativeEngine(lrx(n), backend="sorted") as e:
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]
    assert sum(sizes) == math.factorial(n)


@pytest.mark.parametrize("backend", BACKENDS)
def test_s10_lrx_matches_reference_golden_table(backend):
    with NativeEngine(lrx(10), backend=backend) as e:
        sizes, _ = run_all_layers(e)
    assert sizes == GOLDEN["layers"]["10"]


@pytest.mark.parametrize("backend", BACKENDS)
def test_s11_lrx_matches_oracle(backend):
    ref = O.packed_bfs(11)["layers"]
    with NativeEngine(lrx(11), backend=backend) as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == ref
    assert info["total_states"] == math.factorial(11)


S13_FLAGS = {0: "default", 8: "no back-edge masks (three banks probed)",
             4 | 1: "tight lane slabs and coarse sub-buckets (both retry paths)"}


@pytest.mark.parametrize("flags", list(S13_FLAGS))
def test_s13_sorted_matches_reference_golden_table(flags):
    # Full LRX on S13 (6,227,020,800 states, 79 layers, peak 369,741,101 at
    # depth 54): the reference's own table (data/reference/lrx13-layers.json,
    # tests/test_lrx_layers.py:44-48).
    print(f"S13 sorted flags={flags} ({S13_FLAGS[flags]})", flush=True)
    golden = GOLDEN["layers"]["13"]
    with NativeEngine(lrx(13), backend="sorted", flags=flags) as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == golden
    assert len(sizes) == 79 and sum(sizes) == math.factorial(13)
    assert max(sizes) == 369_741_101 and sizes.index(max(sizes)) == 54
    assert info["total_states"] == math.factorial(13)
    assert info["history_layers"] == (3 if flags & 8 else 2)
    if flags & 8:
        assert info["candidates"] == 3 * math.factorial(13)
    else:
        # Every state but the start skips at least its discovering edge, and
        # every state but the start is some parent's candidate.
        assert math.factorial(13) - 1 <= info["candidates"] <= 2 * math.factorial(13) + 1
    if flags == 4 | 1:
        assert info["retries"] > 0


def test_s13_bitmap_matches_reference_golden_table():
    print("S13 permutation_bitmap", flush=True)
    with NativeEngine(lrx(13), backend="permutation_bitmap") as e:
        sizes, _ = run_all_layers(e)
    assert sizes == GOLDEN["layers"]["13"]


def lrx_variants(n):
    L, R, X = O.lrx_generators(n)
    return {"LRXX": [L, R, X, X], "LRXL": [L, R, X, L], "XLRX": [X, L, R, X], "LRLRX": [L, R, L, R, X]}


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("name", ["LRXX", "LRXL", "XLRX", "LRLRX"])
def test_repeated_generators_match_oracle(backend, name):
    # A generator given more than once: a discovered state must skip every
    # copy of the generator leading back to its parent, not only the first.
    n = 8
    gens = lrx_variants(n)[name]
    ref = O.packed_bfs(n, gens, keep_keys=True)
    assert ref["layers"] == O.packed_bfs(n)["layers"]
    with NativeEngine(GraphDefinition.permutation(gens, list(range(n))), backend=backend) as e:
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]


@pytest.mark.parametrize("seed", [21, 22, 23])
def test_sorted_random_sets_with_repeats_match_oracle(seed):
    rng = random.Random(seed)
    n = rng.choice([7, 8, 9])
    gens = random_generators(rng, n, rng.randint(1, 3), True)
    gens += [list(rng.choice(gens)) for _ in range(rng.randint(1, 2))]
    rng.shuffle(gens)
    ref = O.packed_bfs(n, gens, keep_keys=True)
    with NativeEngine(GraphDefinition.permutation(gens, list(range(n))), backend="sorted") as e:
        assert e.backend_name == NAMES["sorted"]
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]


def test_s11_sorted_overflow_retry_path_is_exact():
    # Coarse start buckets force sub-bucket overflows and finer retries.
    ref = O.packed_bfs(11)["layers"]
    with NativeEngine(lrx(11), backend="sorted", flags=1) as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == ref
    assert info["retries"] > 0


def cross_layer_edges(n, layers_keys):
    """Edges between consecutive BFS layers of LRX on S_n, counted by brute
    force over the oracle's layer sets (each counted once, from the deeper end)."""
    gens = O.lrx_generators(n)
    sets = [set(k) for k in layers_keys]
    total = 0
    for d in range(1, len(sets)):
        for key in sets[d]:
            s = O.unpack(key, n)
            total += sum(O.pack([s[g[i]] for i in range(n)]) in sets[d - 1] for g in gens)
    return total


@pytest.mark.parametrize("n", [8, 9, 10])
def test_sorted_back_edge_masks_skip_every_back_edge(n):
    # With back-edge masks each state skips all generators leading to the
    # previous layer: candidates = 3 n! - (edges between consecutive layers).
    # For even n, L, R and X are odd permutations, so the graph is bipartite
    # by parity and all 3 n! / 2 edges join consecutive layers.
    ref = O.packed_bfs(n, keep_keys=n < 10)
    cross = 3 * math.factorial(n) // 2 if n == 10 else cross_layer_edges(n, ref["keys"])
    if n == 8:
        assert cross == 3 * math.factorial(n) // 2
    with NativeEngine(lrx(n), backend="sorted") as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == ref["layers"]
    assert info["history_layers"] == 2
    assert info["candidates"] == 3 * math.factorial(n) - cross


@pytest.mark.parametrize("n", [9, 10])
def test_sorted_without_pruning_is_exact(n):
    # Flag 8: no back-edge masks, every child generated, both banks probed.
    ref = O.packed_bfs(n)["layers"]
    with NativeEngine(lrx(n), backend="sorted", flags=8) as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == ref
    assert info["history_layers"] == 3
    assert info["candidates"] == 3 * math.factorial(n)


@pytest.mark.parametrize("flags", [32, 32 | 4, 32 | 1])
def test_s11_sorted_wide_routing_path_is_exact(flags):
    # 64-bit slab addressing (outputs of 2^32 entries or more), alone and with
    # the lane-slab and sub-bucket overflow retries, whose dropped runs that
    # path skips instead of sending to the dump area.
    ref = O.packed_bfs(11)["layers"]
    with NativeEngine(lrx(11), backend="sorted", flags=flags) as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == ref
    if flags != 32:
        assert info["retries"] > 0


def test_s11_sorted_lane_slab_retry_path_is_exact():
    # Owner-lane slabs sized to the mean overflow; the layer is redone with slack.
    ref = O.packed_bfs(11)["layers"]
    with NativeEngine(lrx(11), backend="sorted", flags=4) as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == ref
    assert info["retries"] > 0


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_random_inverse_closed_sets_match_oracle(backend, seed):
    rng = random.Random(seed)
    n = rng.choice([7, 8, 9])
    gens = random_generators(rng, n, rng.randint(1, 3), True)
    start = list(range(n))
    rng.shuffle(start)
    ref = O.packed_bfs(n, gens, start=start, keep_keys=True)
    with NativeEngine(GraphDefinition.permutation(gens, start), backend=backend) as e:
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_bitmap_directed_sets_match_oracle(seed):
    rng = random.Random(seed)
    n = rng.choice([7, 8, 9])
    gens = random_generators(rng, n, rng.randint(2, 5), False)
    ref = O.packed_bfs(n, gens, keep_keys=True)
    with NativeEngine(GraphDefinition.permutation(gens, list(range(n))), backend="permutation_bitmap") as e:
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]


def test_sorted_repeated_labels_match_ordered_set_oracle():
    # Multiset start (repeated labels): outside the Lehmer bitmap domain.
    rng = random.Random(5)
    for start in ([0, 0, 1, 1, 2, 2, 3, 3, 4], [0, 0, 0, 1, 1, 2, 3, 4, 5, 6]):
        n = len(start)
        gens = O.lrx_generators(n) + random_generators(rng, n, 1, True)
        graph = {"action": {"kind": "permutation", "degree": n, "generators": gens}, "start": start}
        ref = O.c_exact_layers(graph, math.factorial(n))
        with NativeEngine(GraphDefinition.permutation(gens, start), backend="sorted") as e:
            sizes, keys = run_all_layers(e, keep_keys=True)
        assert sizes == [len(x) for x in ref]
        assert keys == [sorted(O.pack(s) for s in layer) for layer in ref]


@pytest.mark.parametrize("backend", BACKENDS)
def test_samples_match_last_and_previous_layers(backend):
    ref = O.packed_bfs(8, keep_keys=True)
    with NativeEngine(lrx(8), backend=backend) as e:
        run_all_layers(e)
        cur = e.sample(0, 1000)
        prev = e.sample(1, 1000)
    assert sorted(O.pack(r) for r in cur) == sorted(ref["keys"][-1])
    assert sorted(O.pack(r) for r in prev) == sorted(ref["keys"][-2])


# run() with no layer or time limit lets the permutation-bitmap backend queue
# up to 16 layers per host wait (layer sizes read from device counter marks):
# the tables, the last two layers and resumed runs must match one advance()
# per layer and the oracle.
def test_bitmap_run_to_end_matches_golden_table_and_layer_loop():
    with NativeEngine(lrx(10), backend="permutation_bitmap") as e:
        assert e.run() == 1
        sizes = e.layer_sizes()
        total = e.info()["total_states"]
        last = sorted(e.layer_keys())
    with NativeEngine(lrx(10), backend="permutation_bitmap") as e:
        stepped, _ = run_all_layers(e)
        stepped_last = sorted(e.layer_keys())
    assert sizes == stepped == GOLDEN["layers"]["10"]
    assert total == math.factorial(10)
    assert last == stepped_last


def test_bitmap_run_to_end_samples_match_last_and_previous_layers():
    ref = O.packed_bfs(8, keep_keys=True)
    with NativeEngine(lrx(8), backend="permutation_bitmap") as e:
        e.run()
        cur = e.sample(0, 1000)
        prev = e.sample(1, 1000)
    assert sorted(O.pack(r) for r in cur) == sorted(ref["keys"][-1])
    assert sorted(O.pack(r) for r in prev) == sorted(ref["keys"][-2])


def test_bitmap_run_to_end_after_advances_and_reset():
    ref = O.packed_bfs(9)["layers"]
    with NativeEngine(lrx(9), backend="permutation_bitmap") as e:
        for _ in range(5):
            assert e.advance() == 0
        e.run()
        assert e.layer_sizes() == ref
        e.reset()
        e.run()
        assert e.layer_sizes() == ref


@pytest.mark.parametrize("seed", [21, 22, 23])
def test_bitmap_run_to_end_directed_sets_match_oracle(seed):
    rng = random.Random(seed)
    n = rng.choice([7, 8, 9])
    gens = random_generators(rng, n, rng.randint(2, 5), False)
    ref = O.packed_bfs(n, gens, keep_keys=True)
    with NativeEngine(GraphDefinition.permutation(gens, list(range(n))), backend="permutation_bitmap") as e:
        e.run()
        sizes = e.layer_sizes()
        last = sorted(e.layer_keys())
    assert sizes == ref["layers"]
    assert last == sorted(ref["keys"][-1])


@pytest.mark.parametrize("backend", BACKENDS)
def test_reset_repeats_identical_tables(backend):
    with NativeEngine(lrx(9), backend=backend) as e:
        first, _ = run_all_layers(e)
        e.reset()
        second, _ = run_all_layers(e)
    assert first == second == O.packed_bfs(9)["layers"]


@pytest.mark.parametrize("backend", BACKENDS)
def test_rebind_runs_other_graphs_of_the_same_degree(backend):
    # One engine, several generator sets of degree 9: each run equals a
    # fresh engine's and the oracle's table; device memory is kept.
    rng = random.Random(31)
    graphs = [lrx(9)] + [GraphDefinition.permutation(random_generators(rng, 9, k, True), list(range(9)))
                         for k in (1, 2, 3)]
    with NativeEngine(graphs[0], backend=backend) as e:
        before = e.info()["device_bytes"]
        for g in graphs:
            e.rebind(g)
            sizes, _ = run_all_layers(e)
            gens = g.to_dict()["action"]["generators"]
            assert sizes == O.packed_bfs(9, gens)["layers"]
        assert e.info()["device_bytes"] >= before
        with pytest.raises(RuntimeError):
            e.rebind(lrx(8))  # another degree


@pytest.mark.parametrize("backend", BACKENDS)
def test_capacity_limit_stops_with_resource(backend):
    # A capacity below the widest layer must stop as RESOURCE, never silently.
    with NativeEngine(lrx(8), backend=backend, capacity=100) as e:
        statuses = []
        while True:
            s = e.advance()
            statuses.append(s)
            if s != 0:
                break
        assert statuses[-1] == 2 and e.fatal & 4
        ref = O.packed_bfs(8)["layers"]
        sizes = e.layer_sizes()
        assert sizes == ref[:len(sizes)]
