"""Parity of the permutation-bitmap backend (HIP) with the CPU oracle.

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


def test_s8_lrx_layer_sets_bit_exact():
    ref = O.packed_bfs(8, keep_keys=True)
    with NativeEngine(lrx(8), backend="permutation_bitmap") as e:
        assert e.backend_name == "PERMUTATION_BITMAP_LEHMER"
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]
    assert sum(sizes) == 40320


def test_s10_lrx_matches_reference_golden_table():
    with NativeEngine(lrx(10), backend="permutation_bitmap") as e:
        sizes, _ = run_all_layers(e)
    assert sizes == GOLDEN["layers"]["10"]


def test_s11_lrx_matches_oracle():
    ref = O.packed_bfs(11)["layers"]
    with NativeEngine(lrx(11), backend="permutation_bitmap") as e:
        sizes, _ = run_all_layers(e)
        info = e.info()
    assert sizes == ref
    assert info["total_states"] == math.factorial(11)
    assert info["candidates"] == 3 * math.factorial(11)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_generator_sets_match_oracle(seed):
    rng = random.Random(seed)
    n = rng.choice([7, 8, 9])
    gens = []
    for _ in range(rng.randint(2, 5)):
        g = list(range(n))
        rng.shuffle(g)
        gens.append(g)
    if seed % 2 == 0:  # inverse-closed half
        inv = []
        for g in gens:
            h = [0] * n
            for i, v in enumerate(g):
                h[v] = i
            inv.append(h)
        gens = gens + inv
    start = list(range(n))
    rng.shuffle(start)
    ref = O.packed_bfs(n, gens, start=start, keep_keys=True)
    with NativeEngine(GraphDefinition.permutation(gens, start), backend="permutation_bitmap") as e:
        sizes, keys = run_all_layers(e, keep_keys=True)
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]


def test_samples_match_last_and_previous_layers():
    ref = O.packed_bfs(8, keep_keys=True)
    with NativeEngine(lrx(8), backend="permutation_bitmap") as e:
        run_all_layers(e)
        cur = e.sample(0, 1000)
        prev = e.sample(1, 1000)
    assert sorted(O.pack(r) for r in cur) == sorted(ref["keys"][-1])
    assert sorted(O.pack(r) for r in prev) == sorted(ref["keys"][-2])


def test_capacity_limit_stops_with_resource():
    # A capacity below the widest layer must stop as RESOURCE, never silently.
    with NativeEngine(lrx(8), backend="permutation_bitmap", capacity=100) as e:
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
