"""CPU oracle for exact Cayley-graph BFS.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this module, as the checker.  The product (multigpubfs_amd) never
imports it and has no CPU fallback.

Two restatements of the reference's CPU semantics:

* ``exact_layers`` — pure Python, ordered sets, for tiny graphs.  Restates
  multigpubfs/graph_definition.py:70-83 (successor, _multiply with wrap_i64
  and Python ``%``) and :109-117 (exact_layers).
* ``liboracle.so`` (mgbfs_oracle.c, built by oracle/Makefile) — the same
  ordered-set BFS in C (``c_exact_layers``) plus a multi-threaded packed-u64
  BFS with an exact Lehmer-rank bitmap (``packed_bfs``), restating
  scripts/cpu_small_bfs.py:21-80 and gpu/cayley_bfs.cu:58-124.

What pins each function is listed in the header of mgbfs_oracle.c and tested
in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
I64_MIN = -(1 << 63)
I64_MAX = (1 << 63) - 1


def wrap_i64(x: int) -> int:
    return ((x + (1 << 63)) & ((1 << 64) - 1)) - (1 << 63)


# ----------------------------------------------------------- pure Python

def successor(graph: dict, state: list[int], generator: int) -> list[int]:
    """graph_definition.py:70-83."""
    action = graph["action"]
    g = action["generators"][generator]
    if action["kind"] == "permutation":
        return [state[i] for i in g]
    n, m = action["rows"], action["cols"]
    mx, mod = g["matrix"], g["modulo"]
    out = []
    for i in range(n):
        for j in range(m):
            total = wrap_i64(sum(mx[i * n + k] * state[k * m + j] for k in range(n)))
            out.append(total % mod if mod else total)
    return out


def exact_layers(graph: dict, maximum_states: int) -> list[list[list[int]]]:
    """graph_definition.py:109-117 — each layer sorted lexicographically."""
    if type(maximum_states) is not int or maximum_states < 1:
        raise ValueError("ORACLE_CAPACITY")
    action = graph["action"]
    count = len(action["generators"])
    start = tuple(graph["start"])
    seen = {start}
    frontier = {start}
    layers = []
    while frontier:
        layers.append([list(x) for x in sorted(frontier)])
        future = {tuple(successor(graph, list(x), g)) for x in frontier for g in range(count)} - seen
        if len(seen) + len(future) > maximum_states:
            raise ValueError("ORACLE_CAPACITY")
        seen.update(future)
        frontier = future
    return layers


def lrx_transitions(word: int, n: int, bits: int = 4) -> tuple[int, int, int]:
    """Packed L, R, X (scripts/cpu_small_bfs.py:21-27)."""
    symbol = (1 << bits) - 1
    mask = (1 << (n * bits)) - 1
    difference = (word ^ (word >> bits)) & symbol
    return ((word >> bits) | ((word & symbol) << ((n - 1) * bits)),
            ((word << bits) & mask) | (word >> ((n - 1) * bits)),
            word ^ difference ^ (difference << bits))


def lrx_generators(n: int) -> list[list[int]]:
    """L (child[i]=parent[i+1]), R (its inverse), X (swap 0,1)
    — multigpubfs/specialized.py:24, crates/mgbfs-core/src/matrix.rs:188-193."""
    left = [(i + 1) % n for i in range(n)]
    right = [(i - 1) % n for i in range(n)]
    swap = [1, 0] + list(range(2, n))
    return [left, right, swap]


def pack(state, bits: int = 4) -> int:
    return sum(int(v) << (bits * i) for i, v in enumerate(state))


def unpack(word: int, n: int, bits: int = 4) -> list[int]:
    return [(word >> (bits * i)) & ((1 << bits) - 1) for i in range(n)]


def lehmer_rank(perm) -> int:
    """rust/src/cayley.rs:504-516 — lexicographic, position 0 most significant."""
    n = len(perm)
    r = 0
    for i in range(n):
        r = r * (n - i) + sum(1 for v in perm[i + 1:] if v < perm[i])
    return r


# ------------------------------------------------------------- C library

_LIB = None


def build() -> Path:
    """Compile liboracle.so from oracle/mgbfs_oracle.c (gcc)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return HERE / "liboracle.so"


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists() or path.stat().st_mtime < (HERE / "mgbfs_oracle.c").stat().st_mtime:
            build()
        L = C.CDLL(str(path))
        u32, u64, p = C.c_uint32, C.c_uint64, C.c_void_p
        L.orc_exact_layers.argtypes = [u32, u32, u32, u32, u32, p, p, p, p, u64, p, p, u32, p]
        L.orc_exact_layers.restype = C.c_int
        L.orc_packed_bfs.argtypes = [u32, u32, p, u64, u32, p, u32, p, p, u64, u64, p]
        L.orc_packed_bfs.restype = C.c_int
        L.orc_rank.argtypes = [p, u32]
        L.orc_rank.restype = u64
        L.orc_unrank.argtypes = [u64, u32, p]
        L.orc_unrank.restype = C.c_int
        L.orc_multiset_rank.argtypes = [p, u32, p]
        L.orc_multiset_rank.restype = C.c_int
        L.orc_packed_apply.argtypes = [u64, p, u32]
        L.orc_packed_apply.restype = u64
        L.orc_successor.argtypes = [u32, u32, u32, u32, p, p, u32, p, p]
        L.orc_successor.restype = C.c_int
        _LIB = L
    return _LIB


def _arr(ty, values):
    values = list(values)
    return (ty * max(1, len(values)))(*values)


def c_exact_layers(graph: dict, maximum_states: int) -> list[list[list[int]]]:
    """C ordered-set BFS; same contract as exact_layers()."""
    a = graph["action"]
    w = len(graph["start"])
    G = len(a["generators"])
    if a["kind"] == "permutation":
        kind, rows, cols = 0, w, 1
        perms = _arr(C.c_uint32, [v for g in a["generators"] for v in g])
        mats = _arr(C.c_int64, [])
        mods = _arr(C.c_uint32, [])
    else:
        kind, rows, cols = 1, a["rows"], a["cols"]
        perms = _arr(C.c_uint32, [])
        mats = _arr(C.c_int64, [v for g in a["generators"] for v in g["matrix"]])
        mods = _arr(C.c_uint32, [g["modulo"] for g in a["generators"]])
    start = _arr(C.c_int64, graph["start"])
    states = (C.c_int64 * (w * maximum_states))()
    max_layers = 1 << 16
    sizes = (C.c_uint64 * max_layers)()
    nlayers = C.c_uint32()
    rc = lib().orc_exact_layers(kind, w, rows, cols, G, perms, mats, mods, start, maximum_states,
                                states, sizes, max_layers, C.byref(nlayers))
    if rc == 2:
        raise ValueError("ORACLE_CAPACITY")
    if rc:
        raise RuntimeError(f"orc_exact_layers status {rc}")
    out, at = [], 0
    flat = list(states)
    for d in range(nlayers.value):
        k = sizes[d]
        out.append([flat[(at + i) * w:(at + i + 1) * w] for i in range(k)])
        at += k
    return out


def packed_bfs(n: int, generators=None, start=None, threads: int | None = None,
               keep_keys: bool = False, stop_after_edges: int = 0):
    """Packed-u64 Lehmer-bitmap BFS over permutations of 0..n-1.

    generators=None selects LRX packed words.  Returns dict(layers, keys, edges)
    where keys is a list of per-layer sorted key lists when keep_keys.
    """
    if start is None:
        start = list(range(n))
    if threads is None:
        threads = os.cpu_count() or 1
    gens = 0 if generators is None else len(generators)
    perms = _arr(C.c_uint32, [v for g in (generators or []) for v in g])
    max_layers = 4096
    sizes = (C.c_uint64 * max_layers)()
    nlayers = C.c_uint32()
    edges = C.c_uint64()
    cap = math.factorial(n) if keep_keys else 0
    keys = (C.c_uint64 * cap)() if keep_keys else None
    rc = lib().orc_packed_bfs(n, gens, perms, pack(start), threads, sizes, max_layers,
                              C.byref(nlayers), keys, cap, stop_after_edges, C.byref(edges))
    if rc:
        raise RuntimeError(f"orc_packed_bfs status {rc}")
    layers = [sizes[i] for i in range(nlayers.value)]
    result = dict(layers=layers, edges=edges.value)
    if keep_keys:
        flat = list(keys[:sum(layers)])
        out, at = [], 0
        for k in layers:
            out.append(flat[at:at + k])
            at += k
        result["keys"] = out
    return result


def c_rank(perm) -> int:
    return lib().orc_rank(_arr(C.c_int64, perm), len(perm))


def c_unrank(rank: int, n: int) -> list[int]:
    out = (C.c_int64 * n)()
    if lib().orc_unrank(rank, n, out):
        raise ValueError("ORACLE_RANK_RANGE")
    return list(out)


def c_multiset_rank(state) -> int:
    r = C.c_uint64()
    if lib().orc_multiset_rank(_arr(C.c_int64, state), len(state), C.byref(r)):
        raise ValueError("ORACLE_MULTISET")
    return r.value


def c_packed_apply(key: int, perm) -> int:
    return lib().orc_packed_apply(key, _arr(C.c_uint32, perm), len(perm))


def c_successor(graph: dict, state, generator: int) -> list[int]:
    a = graph["action"]
    w = len(state)
    out = (C.c_int64 * w)()
    g = a["generators"][generator]
    if a["kind"] == "permutation":
        rc = lib().orc_successor(0, w, w, 1, _arr(C.c_uint32, g), None, 0, _arr(C.c_int64, state), out)
    else:
        rc = lib().orc_successor(1, w, a["rows"], a["cols"], None, _arr(C.c_int64, g["matrix"]),
                                 g["modulo"], _arr(C.c_int64, state), out)
    if rc:
        raise ValueError("ORACLE_SUCCESSOR")
    return list(out)
