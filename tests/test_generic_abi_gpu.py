"""The generic kernel-level ABI (mgbfs_generic_seed / expand / gather,
cuda/generic_visited.h:3-11) driven as the reference's own GPU test drives
its CUDA build (tests/generic_visited_gpu.py:7-47): caller-owned hipMalloc'd
SoA arena, frontier and future row lists, slot table, counters and error
word; one gather and one expand per layer; each layer's state set against the
oracle's exact_layers, with full 64-bit and with all-colliding (0-bit)
hashes, and a capacity stop that must raise the visited-capacity bit."""
import ctypes as C

import pytest

from hip_memory import DeviceArray, synchronize
from multigpubfs_amd import GraphDefinition
from multigpubfs_amd import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SEED = 1234567

# The reference test's fixtures (tests/generic_visited_gpu.py:40-45).
FIXTURES = {
    "lx4": GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 1, 2, 3]),
    "lx4_repeated_labels": GraphDefinition.permutation([[1, 2, 3, 0], [1, 0, 2, 3]], [0, 0, 1, 2]),
    "matrix_2x1_mod257": GraphDefinition.matrix(2, 1, [([1, 1, 0, 1], 257)], [256, 2]),
    "matrix_wrap_mod0": GraphDefinition.matrix(1, 1, [([2], 0)], [1]),
    "matrix_mixed_moduli": GraphDefinition.matrix(1, 1, [([2], 5), ([3], 7)], [4]),
    "cycle300": GraphDefinition.permutation([list(range(1, 300)) + [0]], list(range(300))),
}


def bind():
    lib = N.load()
    u, p, h = C.c_uint32, C.c_void_p, C.c_uint64
    lib.mgbfs_generic_expand_i64_wide.argtypes = [u, u, u, u, u, p, u, u, p, p, p, p, h, p, u, p, p, u, p, h, u, p, p]
    lib.mgbfs_generic_seed_i64_wide.argtypes = [u, p, u, u, p, h, h, u, p, p]
    lib.mgbfs_generic_gather_i64.argtypes = [u, p, u, p, u, p, u, p]
    for f in ("expand_i64_wide", "seed_i64_wide", "gather_i64"):
        getattr(lib, "mgbfs_generic_" + f).restype = C.c_int
    return lib


def run(graph, bits, capacity=1024):
    lib = bind()
    d = graph.to_dict()
    a = d["action"]
    width = len(d["start"])
    kind = 0 if a["kind"] == "permutation" else 1
    n, m = (a["degree"], 1) if kind == 0 else (a["rows"], a["cols"])
    G = len(a["generators"])
    state = [0] * (capacity * width)
    for e, x in enumerate(d["start"]):
        state[e * capacity] = x
    slots = 1
    while slots < capacity * 4:
        slots *= 2
    visited = DeviceArray(capacity * width, "q").write(state)
    parents = DeviceArray(capacity * width, "q").zero()
    front, future = DeviceArray(capacity, "I").zero(), DeviceArray(capacity, "I").zero()
    table = DeviceArray(slots).write([(1 << 64) - 1] * slots)
    vc, fc, error = DeviceArray(1, "I").write([1]), DeviceArray(1, "I").zero(), DeviceArray(1, "I").zero()
    perms = DeviceArray(max(1, G * n), "I").write([v for g in a["generators"] for v in g] if kind == 0 else [0])
    mats = DeviceArray(max(1, G * n * n), "q").write(
        [v for g in a["generators"] for v in g["matrix"]] if kind else [0])
    mods = DeviceArray(G, "I").write([g["modulo"] for g in a["generators"]] if kind else [0] * G)
    bufs = [visited, parents, front, future, table, vc, fc, error, perms, mats, mods]
    try:
        assert lib.mgbfs_generic_seed_i64_wide(width, visited.ptr, capacity, 1, table.ptr, slots, SEED, bits,
                                               error.ptr, None) == 0
        count, layers, fatal = 1, [], 0
        for _ in range(2048):
            if not count:
                break
            ids = front.read(count)
            raw = visited.read()
            layers.append(sorted([raw[e * capacity + i] for e in range(width)] for i in ids))
            assert lib.mgbfs_generic_gather_i64(width, visited.ptr, capacity, front.ptr, count, parents.ptr,
                                                capacity, None) == 0
            fc.zero()
            assert lib.mgbfs_generic_expand_i64_wide(kind, width, n, m, G, parents.ptr, count, capacity, perms.ptr,
                                                     mats.ptr, mods.ptr, table.ptr, slots, visited.ptr, capacity,
                                                     vc.ptr, future.ptr, capacity, fc.ptr, SEED, bits, error.ptr,
                                                     None) == 0
            synchronize()
            fatal = error.read(1)[0]
            if fatal:
                break
            count = fc.read(1)[0]
            front, future = future, front
        else:
            raise AssertionError("depth bound")
        return layers, fatal
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("bits", [64, 0])
@pytest.mark.parametrize("name", list(FIXTURES))
def test_seed_expand_gather_loop_matches_exact_layers(name, bits):
    graph = FIXTURES[name]
    layers, fatal = run(graph, bits)
    assert fatal == 0
    assert layers == O.c_exact_layers(graph.to_dict(), 1024)


def test_capacity_three_raises_visited_capacity_bit():
    layers, fatal = run(GraphDefinition.permutation([[1, 2, 3, 0]], [0, 1, 2, 3]), 0, capacity=3)
    assert fatal & 2
    assert layers == [[[0, 1, 2, 3]], [[1, 2, 3, 0]], [[2, 3, 0, 1]]]


def test_hardware_info_reports_the_device():
    lib = N.load()
    lib.mgbfs_generic_hardware_info.argtypes = [C.c_int, C.POINTER(C.c_uint64)]
    lib.mgbfs_generic_hardware_info.restype = C.c_int
    values = (C.c_uint64 * 8)()
    assert lib.mgbfs_generic_hardware_info(0, values) == 0
    cus, major, minor, l2, wave, per_cu, per_block, lds = list(values)
    info = N.device_info(0)
    assert cus == info["compute_units"] and wave == info["wave_size"] == 64
    assert (major, minor) == (9, 5)  # gfx950
    assert l2 > 0 and per_block >= 1024 and per_cu >= per_block and lds >= 64 * 1024
    assert lib.mgbfs_generic_hardware_info(0, None) != 0
