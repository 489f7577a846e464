"""Segmented open-addressing backend (MGBFS_BACKEND_SEGMENTED) against the
oracle: three rolling banks for inverse-closed graphs, full history for
directed ones, u8 and int64 state rows, forced segment spills, colliding
hashes, and the sorted backend's hand-over when a layer exceeds its limits
(cuda/segmented_history.cu and crates/mgbfs-runtime/src/segmented_history.rs
in the reference; fixtures of tests/generic_visited_gpu.py:40-45)."""
import math
import random

import pytest

from multigpubfs_amd import GraphDefinition, NativeEngine, lrx, run_graph
from oracle import oracle as O

pytestmark = pytest.mark.gpu
COLLIDE, TINY, SORTED_CAP = 2, 128, 256
NAME = "SEGMENTED_OPEN_ADDRESSING_THREE_BANK"

FIXTURES = {
    "lx4": GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 1, 2, 3]),
    "lx4_repeated_labels": GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 0, 1, 2]),
    "matrix_2x1_mod257": GraphDefinition.matrix(2, 1, [([1, 1, 0, 1], 257)], [256, 2]),
    "matrix_wrap_mod0": GraphDefinition.matrix(1, 1, [([2], 0)], [1]),
    "matrix_mixed_moduli": GraphDefinition.matrix(1, 1, [([2], 5), ([3], 7)], [4]),
    "cycle300": GraphDefinition.permutation([list(range(1, 300)) + [0]], list(range(300))),
}


def layers_of(engine):
    rows = [sorted(map(tuple, engine.sample(0, 1 << 30)))]
    while True:
        status = engine.advance()
        if status == 1:
            break
        assert status == 0, (status, engine.fatal)
        rows.append(sorted(map(tuple, engine.sample(0, 1 << 30))))
    return engine.layer_sizes(), rows


def oracle_layers(graph, bound):
    return [sorted(map(tuple, layer)) for layer in O.c_exact_layers(graph.to_dict(), bound)]


@pytest.mark.parametrize("flags", [0, COLLIDE])
@pytest.mark.parametrize("name", list(FIXTURES))
def test_reference_fixtures_layer_sets(name, flags):
    g = FIXTURES[name]
    with NativeEngine(g, backend="segmented", flags=flags) as e:
        assert e.backend_name == NAME
        sizes, rows = layers_of(e)
    want = oracle_layers(g, 1024)
    assert rows == want and sizes == [len(x) for x in want]


def sl2(p):
    return GraphDefinition.matrix(2, 2, [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p),
                                         ([1, 0, p - 1, 1], p)], [1, 0, 0, 1])


@pytest.mark.parametrize("p", [5, 7, 11, 13])
def test_sl2_three_banks_layer_sets(p):
    g = sl2(p)
    with NativeEngine(g) as e:  # AUTO: inverse-closed int64 states -> segmented
        assert e.backend_name == NAME
        sizes, rows = layers_of(e)
        assert e.info()["history_layers"] == 3
    assert sum(sizes) == p * (p * p - 1)
    assert rows == oracle_layers(g, p * (p * p - 1))


@pytest.mark.parametrize("flags", [0, TINY, TINY | COLLIDE])
def test_lrx9_packed_states_and_segment_spills(flags):
    ref = O.packed_bfs(9, keep_keys=True)
    with NativeEngine(lrx(9), backend="segmented", flags=flags) as e:
        keys = [sorted(e.layer_keys())]
        while e.advance() == 0:
            keys.append(sorted(e.layer_keys()))
        assert e.status == 1
        sizes = e.layer_sizes()
        spills = e.info()["retries"]
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]
    if flags & TINY:
        assert spills > 0


def test_degree_16_and_wide_labels():
    # Degree 16 leaves no tag nibble for the packed backends; labels >= 16
    # are outside the packed encoding; labels > 255 need int64 rows.
    n, k = 16, 8
    L = [(i + 1) % k if i < k else i for i in range(n)]
    R = [(i - 1) % k if i < k else i for i in range(n)]
    X = [1, 0] + list(range(2, n))
    for start in (list(range(n)), [3 * i + 1 for i in range(n)], [1000 + 7 * i for i in range(n)]):
        g = GraphDefinition.permutation([L, R, X], start)
        with NativeEngine(g) as e:
            assert e.backend_name == NAME
            sizes, rows = layers_of(e)
        want = oracle_layers(g, math.factorial(k))
        assert sum(sizes) == math.factorial(k) and rows == want


@pytest.mark.parametrize("seed", [41, 42])
@pytest.mark.parametrize("flags", [0, TINY])
def test_directed_full_history(seed, flags):
    rng = random.Random(seed)
    n = 8
    gens = []
    for _ in range(rng.randint(2, 3)):
        g = list(range(n))
        rng.shuffle(g)
        gens.append(g)
    graph = GraphDefinition.permutation(gens, list(range(n)))
    assert not graph.inverse_closed()
    ref = O.packed_bfs(n, gens, keep_keys=True)
    with NativeEngine(graph, backend="segmented", flags=flags) as e:
        keys = [sorted(e.layer_keys())]
        while e.advance() == 0:
            keys.append(sorted(e.layer_keys()))
        assert e.info()["history_layers"] == 0
    assert keys == [sorted(k) for k in ref["keys"]]


@pytest.mark.parametrize("n,flags", [(10, SORTED_CAP), (11, SORTED_CAP), (11, SORTED_CAP | TINY)])
def test_sorted_hands_over_to_segmented_at_its_limits(n, flags):
    # The test flag lowers the sorted backend's limit to 2^18 children per
    # layer: the run starts sorted and finishes on the segmented backend
    # from the sorted backend's last two layers.
    ref = O.packed_bfs(n, keep_keys=(n == 10))
    with NativeEngine(lrx(n), backend="sorted", flags=flags) as e:
        assert e.backend_name == "SORTED_THREE_BANK_LDS_UNIQUE"
        names = set()
        keys = [sorted(e.layer_keys())] if n == 10 else None
        while True:
            s = e.advance()
            names.add(e.backend_name)
            if s != 0:
                break
            if keys is not None:
                keys.append(sorted(e.layer_keys()))
        assert s == 1
        sizes = e.layer_sizes()
        info = e.info()
        prev = e.sample(1, 1000)
    assert NAME in names
    assert sizes == ref["layers"] and info["total_states"] == math.factorial(n)
    assert info["backend"] == 3
    if keys is not None:
        assert keys == [sorted(k) for k in ref["keys"]]
        assert sorted(O.pack(r) for r in prev) == sorted(ref["keys"][-2])


def test_resource_stop_keeps_prefix_and_samples():
    with NativeEngine(sl2(11), backend="segmented", capacity=40) as e:
        statuses = []
        while True:
            s = e.advance()
            statuses.append(s)
            if s != 0:
                break
        assert statuses[-1] == 2
        sizes = e.layer_sizes()
        cur = sorted(map(tuple, e.sample(0, 1000)))
    want = oracle_layers(sl2(11), 1320)
    assert sizes == [len(x) for x in want[:len(sizes)]]
    assert cur == want[len(sizes) - 1]


def test_reset_repeats_and_segmented_memory_backend_through_cli(tmp_path):
    with NativeEngine(sl2(13)) as e:
        first, _ = layers_of(e)
        e.reset()
        second, _ = layers_of(e)
    assert first == second
    report = run_graph(sl2(13), tmp_path / "out", device=0, memory_backend="segmented")
    assert report["layer_sizes"] == first and report["backend"] == NAME
    report = run_graph(lrx(8), tmp_path / "out8", device=0, memory_backend="segmented_hash")
    assert report["layer_sizes"] == O.packed_bfs(8)["layers"] and report["backend"] == NAME
