"""mgbfs_sorted_layer_noalloc (static per-GPU scratch) against
mgbfs_sorted_layer (caller workspace) through the C ABI: identical next
banks, offsets, counters and error words on two calls in a row, the second
on a smaller layer; a whole S10 BFS through the no-alloc entry point against
the reference's golden table; and the size limit."""
import ctypes as C
import json
from pathlib import Path

import pytest

from hip_memory import DeviceArray, synchronize
from multigpubfs_amd import _native as N
from oracle import oracle as O
from sorted_abi_driver import SortedAbiRun

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((Path(__file__).parent / "golden" / "lrx_layers.json").read_text())
SCRATCH_BYTES = 64 << 20


def noalloc_layer(run):
    lib = N.bind_sorted(N.load())
    lib.mgbfs_sorted_layer_noalloc.argtypes = [C.POINTER(N.SortedLayerArgs), C.c_void_p]
    lib.mgbfs_sorted_layer_noalloc.restype = C.c_int

    def call(args, info):
        assert info.workspace_bytes <= SCRATCH_BYTES
        return lib.mgbfs_sorted_layer_noalloc(C.byref(args), None)
    return lib, call


def advance_to(run, depth):
    while run.depth < depth:
        assert run.advance()


def one_call(run, lib, noalloc, capacity):
    """One layer from the driver's current bank into fresh outputs, through
    either entry point; returns (sorted next values, offsets, counters, error)."""
    info = run.plan(run.counts[run.cur])
    nxt, offs = DeviceArray(capacity), DeviceArray(N.SORTED_OFFSETS, "I")
    counters, error = DeviceArray(2).zero(), DeviceArray(1, "I").zero()
    args = N.SortedLayerArgs()
    args.degree, args.generators, args.permutations = run.degree, len(run.gens), run.table
    args.depth, args.buckets = run.depth, run.buckets[run.cur]
    args.bank, args.offsets, args.count = run.banks[run.cur].ptr, run.offsets[run.cur].ptr, run.counts[run.cur]
    if run.prev >= 0:
        args.previous, args.previous_offsets = run.banks[run.prev].ptr, run.offsets[run.prev].ptr
        args.previous_buckets = run.buckets[run.prev]
    args.next, args.next_offsets, args.next_capacity = nxt.ptr, offs.ptr, capacity
    args.counters, args.error = counters.ptr, error.ptr
    if noalloc:
        rc = lib.mgbfs_sorted_layer_noalloc(C.byref(args), None)
    else:
        ws = DeviceArray(info.workspace_bytes, "B")
        rc = lib.mgbfs_sorted_layer(C.byref(args), ws.ptr, ws.nbytes, None)
    assert rc == 0, rc
    synchronize()
    offsets = offs.read(info.next_buckets + 1)
    out = (sorted(nxt.read(offsets[-1])), offsets, counters.read(2), error.read(1))
    for a in (nxt, offs, counters, error):
        a.free()
    return out


def test_noalloc_matches_workspace_entry_on_two_calls_large_then_small():
    lib, _ = noalloc_layer(None)
    sizes = GOLDEN["layers"]["10"]
    peak = sizes.index(max(sizes))
    big = SortedAbiRun(O.lrx_generators(10), list(range(10)))
    small = SortedAbiRun(O.lrx_generators(10), list(range(10)))
    try:
        advance_to(big, peak)
        advance_to(small, 6)
        assert big.plan(big.counts[big.cur]).workspace_bytes > 4 * small.plan(small.counts[small.cur]).workspace_bytes
        results = []
        for run in (big, small):
            cap = run.plan(run.counts[run.cur]).candidates_bound
            want = one_call(run, lib, False, cap)
            got = one_call(run, lib, True, cap)
            assert want[3] == [0] and want[1][-1] == sizes[run.depth + 1]
            assert got == want
            results.append(got)
    finally:
        big.close()
        small.close()
    assert len(results[1][0]) < len(results[0][0])


def test_noalloc_runs_whole_s10_bfs_exactly():
    run = SortedAbiRun(O.lrx_generators(10), list(range(10)))
    _, run.layer_call = noalloc_layer(run)
    try:
        sizes, _ = run.run()
    finally:
        run.close()
    assert sizes == GOLDEN["layers"]["10"]


def test_noalloc_rejects_layers_beyond_the_scratch():
    lib, _ = noalloc_layer(None)
    count = 3_049_721  # peak layer of LRX on S11: its plan needs about 141 MiB
    gens = O.lrx_generators(11)
    table = (C.c_uint32 * 33)(*[v for g in gens for v in g])
    info = N.SortedPlanInfo()
    assert lib.mgbfs_sorted_plan(11, 3, table, 5, count, 0, 0, C.byref(info)) == 0
    assert info.workspace_bytes > SCRATCH_BYTES
    # Real, zeroed buffers of the declared sizes: the call must return before
    # enqueueing anything, and would stay in bounds even if it did not.
    bank, offs = DeviceArray(count).zero(), DeviceArray(N.SORTED_OFFSETS, "I").zero()
    nxt, noffs = DeviceArray(info.candidates_bound).zero(), DeviceArray(N.SORTED_OFFSETS, "I").zero()
    counters, error = DeviceArray(2).zero(), DeviceArray(1, "I").zero()
    args = N.SortedLayerArgs()
    args.degree, args.generators, args.permutations, args.depth = 11, 3, table, 5
    args.buckets, args.bank, args.offsets, args.count = 1024, bank.ptr, offs.ptr, count
    args.next, args.next_offsets, args.next_capacity = nxt.ptr, noffs.ptr, info.candidates_bound
    args.counters, args.error = counters.ptr, error.ptr
    try:
        assert lib.mgbfs_sorted_layer_noalloc(C.byref(args), None) == 2  # hipErrorOutOfMemory
        synchronize()
        assert error.read(1) == [0] and counters.read(2) == [0, 0]
    finally:
        for a in (bank, offs, nxt, noffs, counters, error):
            a.free()
