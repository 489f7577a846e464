"""Generic SoA backend (any width, permutation or int64 matrix actions,
all-visited exact table) against the ordered-set oracle.

Fixtures are the reference's own generic-path fixtures
(tests/generic_visited_gpu.py:43-49; crates/mgbfs-runtime/tests/generic_native_gpu.rs:7-40),
each run with full 64-bit hashes and with every hash colliding (equality
alone decides), plus the RESOURCE stop at a tiny capacity.
"""
import math

import pytest

from multigpubfs_amd import GraphDefinition, NativeEngine, lrx
from oracle import oracle as O

pytestmark = pytest.mark.gpu
COLLIDE = 2

FIXTURES = [
    GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 1, 2, 3]),
    GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 0, 1, 2]),
    GraphDefinition.matrix(2, 1, [([1, 1, 0, 1], 257)], [256, 2]),
    GraphDefinition.matrix(1, 1, [([2], 0)], [1]),
    GraphDefinition.matrix(1, 1, [([2], 5), ([3], 7)], [4]),
    GraphDefinition.permutation([list(range(1, 300)) + [0]], list(range(300))),
]


def run_layers(engine):
    layers = [sorted(map(tuple, engine.sample(0, 1 << 20)))]
    while True:
        status = engine.advance()
        if status != 0:
            return status, layers
        layers.append(sorted(map(tuple, engine.sample(0, 1 << 20))))


@pytest.mark.parametrize("index", range(len(FIXTURES)))
@pytest.mark.parametrize("flags", [0, COLLIDE])
def test_reference_fixtures_exact(index, flags):
    g = FIXTURES[index]
    expected = O.c_exact_layers(g.to_dict(), 1024)
    with NativeEngine(g, backend="generic", capacity=1024, flags=flags) as e:
        assert e.backend_name == "GENERIC_EXACT_ALL_VISITED"
        status, layers = run_layers(e)
    assert status == 1
    assert layers == [sorted(map(tuple, x)) for x in expected]


def test_resource_stop_keeps_completed_prefix():
    g = GraphDefinition.permutation([[1, 2, 3, 0]], [0, 1, 2, 3])
    with NativeEngine(g, backend="generic", capacity=3, flags=COLLIDE) as e:
        status, layers = run_layers(e)
        assert status == 2 and e.fatal & 2
    assert [len(x) for x in layers] == [1, 1, 1]


@pytest.mark.parametrize("p", [5, 7, 11, 13])
def test_sl2_matrix_action_matches_oracle(p):
    gens = [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p), ([1, 0, p - 1, 1], p)]
    g = GraphDefinition.matrix(2, 2, gens, [1, 0, 0, 1])
    expected = O.c_exact_layers(g.to_dict(), p * (p * p - 1))
    with NativeEngine(g, backend="generic", capacity=p * (p * p - 1)) as e:
        status, layers = run_layers(e)
    assert status == 1
    assert layers == [sorted(map(tuple, x)) for x in expected]
    assert sum(map(len, layers)) == p * (p * p - 1)


def test_wide_u8_lrx_s8_matches_packed_oracle():
    ref = O.packed_bfs(8, keep_keys=True)
    with NativeEngine(lrx(8), backend="generic") as e:
        status, layers = run_layers(e)
        assert e.info()["state_bytes"] == 1
    assert status == 1
    assert [sorted(O.pack(r) for r in layer) for layer in layers] == [sorted(k) for k in ref["keys"]]


def test_wide_degree_20_permutations_match_oracle():
    # Degree 20 does not pack into 64 bits; a 3-cycle and a double swap.
    n = 20
    a = list(range(n)); a[0], a[1], a[2] = 1, 2, 0
    b = list(range(n)); b[5], b[6], b[17], b[18] = 6, 5, 18, 17
    g = GraphDefinition.permutation([a, b], list(range(n)))
    expected = O.c_exact_layers(g.to_dict(), 10000)
    with NativeEngine(g, backend="generic", capacity=10000) as e:
        status, layers = run_layers(e)
    assert status == 1 and layers == [sorted(map(tuple, x)) for x in expected]


def test_previous_layer_sample():
    g = lrx(6)
    expected = O.c_exact_layers(g.to_dict(), math.factorial(6))
    with NativeEngine(g, backend="generic") as e:
        run_layers(e)
        prev = sorted(map(tuple, e.sample(1, 1000)))
    assert prev == sorted(map(tuple, expected[-2]))
