"""Pins the CPU oracle (oracle/) before it is trusted as the GPU checker.

Every check below is something the restatement could get wrong in a
plausible way (a transposed operand, a wrong shift direction, an off-by-one
in the rank digits, a dropped wrap) and that an independent source fixes:
the reference's golden layer tables, the reference's own fixtures, or
mathematics (closed forms, exhaustive enumeration, brute force).
"""
import itertools
import json
import math
from pathlib import Path

import pytest

from oracle import oracle as O

GOLDEN = json.loads((Path(__file__).parent / "golden" / "lrx_layers.json").read_text())


def tuple_lrx_layers(n):
    """Independent tuple BFS (the reference test's own oracle shape,
    tests/test_cpu_small_bfs.py:12-24)."""
    start = tuple(range(n))
    seen = {start}
    cur = {start}
    out = []
    while cur:
        out.append(sorted(cur))
        nxt = set()
        for w in cur:
            for c in (w[1:] + w[:1], w[-1:] + w[:-1], (w[1], w[0]) + w[2:]):
                if c not in seen:
                    seen.add(c)
                    nxt.add(c)
        cur = nxt
    return out


# ------------------------------------------------------------ golden tables

def test_golden_tables_are_the_reference_files():
    s13 = GOLDEN["layers"]["13"]
    assert sum(s13) == math.factorial(13) and len(s13) == 79 and max(s13) == 369741101
    s10 = GOLDEN["layers"]["10"]
    assert sum(s10) == math.factorial(10) and len(s10) == 46


@pytest.mark.parametrize("generic", [False, True])
def test_packed_bfs_matches_reference_s10_table(generic):
    gens = O.lrx_generators(10) if generic else None
    assert O.packed_bfs(10, gens)["layers"] == GOLDEN["layers"]["10"]


def test_planted_wrong_generator_is_caught_by_the_golden_table():
    # Layer counts only see the graph up to isomorphism: X on (1,2) is the
    # conjugate of X on (0,1) by the rotation and gives the same table, so
    # the planted mistakes are ones that change the graph.  The packed
    # transitions themselves are pinned against tuple rotations below.
    def with_gen(index, perm):
        gens = O.lrx_generators(10)
        gens[index] = perm
        return O.packed_bfs(10, gens)["layers"]
    assert with_gen(2, [2, 1, 0] + list(range(3, 10))) != GOLDEN["layers"]["10"]   # X on (0,2)
    assert with_gen(1, O.lrx_generators(10)[0]) != GOLDEN["layers"]["10"]         # R dropped for L
    assert with_gen(0, [(i + 2) % 10 for i in range(10)]) != GOLDEN["layers"]["10"]  # L shifts by 2


@pytest.mark.slow
def test_packed_bfs_matches_reference_s13_table():
    assert O.packed_bfs(13)["layers"] == GOLDEN["layers"]["13"]


def test_lrx_closed_forms_small_n():
    # LRX generates S_n (an n-cycle and a transposition): |V| = n!.
    for n in range(2, 10):
        layers = O.packed_bfs(n)["layers"]
        assert sum(layers) == math.factorial(n)


def test_packed_layers_equal_tuple_bfs_sets():
    for n in (4, 6, 8):
        ref = tuple_lrx_layers(n)
        got = O.packed_bfs(n, keep_keys=True)["keys"]
        assert [sorted(tuple(O.unpack(k, n)) for k in layer) for layer in got] == ref


def test_packed_transitions_are_rotations_and_swap():
    for n in range(2, 17):
        state = list(range(n))[::-1]
        word = O.pack(state)
        l, r, x = O.lrx_transitions(word, n)
        assert O.unpack(l, n) == state[1:] + state[:1]
        assert O.unpack(r, n) == state[-1:] + state[:-1]
        assert O.unpack(x, n) == [state[1], state[0]] + state[2:]
        gens = O.lrx_generators(n)
        assert [O.c_packed_apply(word, g) for g in gens] == [l, r, x]


# ------------------------------------------------ Coxeter / Mahonian closed forms

def mahonian(n):
    counts = [1]
    for size in range(2, n + 1):
        nxt = [0] * (len(counts) + size - 1)
        for inv, c in enumerate(counts):
            for add in range(size):
                nxt[inv + add] += c
        counts = nxt
    return counts


def coxeter(n):
    gens = []
    for i in range(n - 1):
        g = list(range(n))
        g[i], g[i + 1] = g[i + 1], g[i]
        gens.append(g)
    return gens


def test_coxeter_layers_are_mahonian_and_reference_kats():
    for n in (5, 7, 8, 9):
        layers = O.packed_bfs(n, coxeter(n))["layers"]
        assert layers == mahonian(n)
    s8 = O.packed_bfs(8, coxeter(8))["layers"]
    assert len(s8) - 1 == 28 and max(s8) == 3836            # rust/src/cayley.rs:366-377
    s9 = O.packed_bfs(9, coxeter(9))["layers"]
    assert len(s9) == 37 and max(s9) == 29228 and s9.index(29228) == 18  # REF-017:46-47


# ------------------------------------------------------------------- ranks

def test_rank_exhaustive_lexicographic_order():
    for n in (1, 4, 8):
        for index, perm in enumerate(itertools.permutations(range(n))):
            assert O.c_rank(perm) == index
            if n <= 4 or index % 97 == 0:
                assert O.c_unrank(index, n) == list(perm)
                assert O.lehmer_rank(perm) == index


def test_rank_kat_round_trips():
    # crates/mgbfs-runtime/tests/permutation_bitmap_contracts.rs:36-41
    for n in (13, 20):
        size = math.factorial(n)
        for r in (0, 1, 4294967313, size // 2, size - 1):
            if r < size:
                assert O.c_rank(O.c_unrank(r, n)) == r
    assert math.factorial(20) == 2432902008176640000
    with pytest.raises(ValueError):
        O.c_unrank(math.factorial(8), 8)


def test_multiset_rank_exhaustive():
    # permutation_bitmap_contracts.rs:14-27 starts, lexicographic enumeration
    for start in ([0], [0, 0, 1, 2, 2, 3], [-7, -7, 0, 99, 99], [4, 4, 4, 4], list(range(6))):
        distinct = sorted(set(itertools.permutations(start)))
        for index, state in enumerate(distinct):
            assert O.c_multiset_rank(state) == index
        size = math.factorial(len(start))
        for v in set(start):
            size //= math.factorial(start.count(v))
        assert len(distinct) == size


# -------------------------------------------------- action fixtures (reference)

def perm_graph(gens, start):
    return {"action": {"kind": "permutation", "degree": len(start), "generators": gens}, "start": start}


def mat_graph(rows, cols, gens, start):
    return {"action": {"kind": "matrix", "rows": rows, "cols": cols,
                       "generators": [{"matrix": m, "modulo": mod} for m, mod in gens]}, "start": start}


def test_action_kats_from_reference_tests():
    # tests/test_graph_definition.py:6-19
    g = perm_graph([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 0, 1, 2])
    for succ in (O.successor, lambda gr, s, k: O.c_successor(gr, s, k)):
        assert succ(g, [0, 0, 1, 2], 0) == [0, 1, 2, 0]
    assert sum(map(len, O.exact_layers(g, 100))) == 12
    m = mat_graph(2, 1, [([1, 1, 0, 1], 257)], [256, 2])
    assert O.successor(m, [256, 2], 0) == [1, 2] == O.c_successor(m, [256, 2], 0)
    w = mat_graph(1, 1, [([2], 7)], [(1 << 63) - 1])
    assert O.successor(w, w["start"], 0) == [5] == O.c_successor(w, w["start"], 0)
    z = mat_graph(1, 1, [([2], 0)], [(1 << 63) - 1])
    assert O.successor(z, z["start"], 0) == [-2] == O.c_successor(z, z["start"], 0)
    mm = mat_graph(1, 1, [([2], 5), ([3], 7)], [4])
    assert [O.successor(mm, [4], k) for k in (0, 1)] == [[3], [5]]
    assert [O.c_successor(mm, [4], k) for k in (0, 1)] == [[3], [5]]
    # a fake inverse under wrapping (test_graph_definition.py:41-45)
    f = mat_graph(1, 1, [([(1 << 63) - 1], 3)], [2])
    assert O.c_successor(f, [2], 0) == [1]


def test_python_and_c_exact_layers_agree_on_reference_fixtures():
    fixtures = [
        perm_graph([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 1, 2, 3]),
        perm_graph([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 0, 1, 2]),
        mat_graph(2, 1, [([1, 1, 0, 1], 257)], [256, 2]),
        mat_graph(1, 1, [([2], 0)], [1]),
        mat_graph(1, 1, [([2], 5), ([3], 7)], [4]),
        perm_graph([list(range(1, 30)) + [0]], list(range(30))),
    ]
    for g in fixtures:
        assert O.c_exact_layers(g, 4096) == O.exact_layers(g, 4096)


def test_exact_layers_capacity_is_reported():
    g = perm_graph([[1, 2, 0]], [0, 1, 2])
    with pytest.raises(ValueError):
        O.exact_layers(g, 2)
    with pytest.raises(ValueError):
        O.c_exact_layers(g, 2)


def sl2(p):
    return mat_graph(2, 2, [([1, 1, 0, 1], p), ([1, 0, 1, 1], p), ([1, p - 1, 0, 1], p), ([1, 0, p - 1, 1], p)],
                     [1, 0, 0, 1])


def test_sl2_order_closed_form():
    for p in (2, 3, 5, 7, 11, 13):
        layers = O.c_exact_layers(sl2(p), p * (p * p - 1))
        assert sum(map(len, layers)) == p * (p * p - 1)
    assert O.c_exact_layers(sl2(5), 120) == O.exact_layers(sl2(5), 120)


def test_ordered_set_bfs_equals_packed_bfs_on_random_generators():
    import random
    rng = random.Random(7)
    for trial in range(6):
        n = rng.choice([5, 6, 7])
        gens = []
        for _ in range(rng.randint(2, 3)):
            g = list(range(n))
            rng.shuffle(g)
            gens.append(g)
        packed = O.packed_bfs(n, gens, keep_keys=True)
        sets = O.c_exact_layers(perm_graph(gens, list(range(n))), math.factorial(n))
        assert [len(x) for x in sets] == packed["layers"]
        assert [sorted(O.pack(s) for s in layer) for layer in sets] == [sorted(k) for k in packed["keys"]]
