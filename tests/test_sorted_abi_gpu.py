"""The kernel-level sorted-bank layer ABI (mgbfs_sorted_*) driven as a
foreign caller would: hipMalloc'd banks and workspace, one layer call per
BFS layer, retries on the documented error bits.  Layer sets are compared
with the oracle; the call replaces the reference's sorted-run owner
(cuda/segmented_hash_batch.cu:157-215)."""
import ctypes as C
import json
import math
from pathlib import Path

import pytest

from multigpubfs_amd import _native as N
from oracle import oracle as O
from sorted_abi_driver import ERR_BUCKET, ERR_FUTURE, SortedAbiRun

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((Path(__file__).parent / "golden" / "lrx_layers.json").read_text())


def test_s9_lrx_layer_sets_bit_exact():
    ref = O.packed_bfs(9, keep_keys=True)
    run = SortedAbiRun(O.lrx_generators(9), list(range(9)))
    try:
        sizes, keys = run.run(keep_keys=True)
    finally:
        run.close()
    assert sizes == ref["layers"]
    assert keys == [sorted(k) for k in ref["keys"]]


def test_s10_lrx_matches_reference_golden_table_with_forced_retries():
    # Coarse sub-buckets (error bit 64) and 64-row banks (error bit 4) on
    # every large layer: each is redone as the header prescribes.
    run = SortedAbiRun(O.lrx_generators(10), list(range(10)), coarse=True, initial_rows=64)
    try:
        sizes, _ = run.run()
    finally:
        run.close()
    assert sizes == GOLDEN["layers"]["10"]
    assert run.retries[ERR_BUCKET] > 0 and run.retries[ERR_FUTURE] > 0
    # With back-edge masks every state but the start skips its back edges;
    # for even n LRX is bipartite by parity, so 3 n!/2 edges join layers.
    assert run.candidates == 3 * math.factorial(10) - 3 * math.factorial(10) // 2


def test_unmasked_three_bank_path_on_degree_15():
    # Degree 15 leaves 4 tag bits: three generators get no back-edge masks,
    # so every child is generated and both retained banks are probed.  LRX
    # acts on positions 0..7 only (8! states), with a shuffled start.
    n, k = 15, 8
    L = [(i + 1) % k if i < k else i for i in range(n)]
    R = [(i - 1) % k if i < k else i for i in range(n)]
    X = [1, 0] + list(range(2, n))
    start = [3, 0, 7, 5, 1, 6, 2, 4, 8, 9, 10, 11, 12, 13, 14]
    graph = {"action": {"kind": "permutation", "degree": n, "generators": [L, R, X]}, "start": start}
    ref = O.c_exact_layers(graph, math.factorial(k))
    run = SortedAbiRun([L, R, X], start)
    try:
        assert run.plan(1).history_banks == 3 and run.plan(1).masked == 0
        sizes, keys = run.run(keep_keys=True)
    finally:
        run.close()
    assert sizes == [len(x) for x in ref] and sum(sizes) == math.factorial(k)
    assert keys == [sorted(O.pack(s) for s in layer) for layer in ref]
    assert run.candidates == 3 * math.factorial(k)


def test_plan_rejects_what_the_layer_cannot_run():
    lib = N.bind_sorted(N.load())
    info = N.SortedPlanInfo()
    lrx = O.lrx_generators(8)
    table = (C.c_uint32 * 24)(*[v for g in lrx for v in g])
    assert lib.mgbfs_sorted_plan(8, 3, table, 0, 1, 0, 0, C.byref(info)) == 0
    assert info.history_banks == 2 and info.masked == 0
    assert lib.mgbfs_sorted_plan(8, 3, table, 1, 1000, 0, 0, C.byref(info)) == 0
    assert info.masked == 1 and info.candidates_bound == 2000
    directed = (C.c_uint32 * 16)(*(lrx[0] + lrx[2]))  # L and X: L's inverse is missing
    assert lib.mgbfs_sorted_plan(8, 2, directed, 0, 1, 0, 0, C.byref(info)) != 0
    assert lib.mgbfs_sorted_plan(16, 3, table, 0, 1, 0, 0, C.byref(info)) != 0
    assert lib.mgbfs_sorted_plan(8, 3, table, 1, 1 << 32, 0, 0, C.byref(info)) != 0
