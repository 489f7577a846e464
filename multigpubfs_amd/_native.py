"""ctypes binding of libmgbfs_hip.so (include/mgbfs_hip.h).

The native library is mandatory: there is no CPU fallback.  Loading fails
loudly (NativeUnavailable) when the in-tree build is missing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libmgbfs_hip.so"
CLI_PATH = PKG / "bin" / "mgbfs"

MGBFS_LAYER, MGBFS_COMPLETE, MGBFS_RESOURCE = 0, 1, 2
BACKENDS = {"auto": 0, "permutation_bitmap": 1, "sorted": 2, "segmented": 3, "generic": 4}
BACKEND_NAMES = {v: k for k, v in BACKENDS.items()}


class NativeUnavailable(RuntimeError):
    """The HIP library is not built or cannot be loaded."""


class GraphDesc(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("elements", C.c_uint32), ("rows", C.c_uint32), ("cols", C.c_uint32),
                ("generators", C.c_uint32), ("inverse_closed", C.c_uint32),
                ("permutations", C.POINTER(C.c_uint32)), ("matrices", C.POINTER(C.c_int64)),
                ("moduli", C.POINTER(C.c_uint32)), ("start", C.POINTER(C.c_int64)),
                ("expected_states", C.c_uint64)]


class EngineOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("backend", C.c_uint32), ("capacity", C.c_uint64),
                ("vram_budget", C.c_uint64), ("flags", C.c_uint32), ("batch", C.c_uint32)]


class EngineInfo(C.Structure):
    _fields_ = [("backend", C.c_uint32), ("history_layers", C.c_uint32), ("state_bytes", C.c_uint32),
                ("depth", C.c_uint32), ("current_count", C.c_uint64), ("total_states", C.c_uint64),
                ("device_bytes", C.c_uint64), ("peak_layer", C.c_uint64), ("candidates", C.c_uint64),
                ("unique_candidates", C.c_uint64), ("kernel_seconds", C.c_double),
                ("kernel_launches", C.c_uint64), ("algorithmic_bytes", C.c_uint64), ("retries", C.c_uint64)]


_LIB = None


def load():
    """Load the in-tree libmgbfs_hip.so (raises NativeUnavailable)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise NativeUnavailable(f"{LIB_PATH} is not built; run `python -m multigpubfs_amd.build`")
    try:
        lib = C.CDLL(str(LIB_PATH))
    except OSError as exc:
        raise NativeUnavailable(str(exc)) from exc
    p, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32
    lib.mgbfs_engine_create.argtypes = [C.POINTER(GraphDesc), C.POINTER(EngineOptions), C.POINTER(p),
                                        C.c_char_p, C.c_size_t]
    lib.mgbfs_engine_create.restype = C.c_int
    lib.mgbfs_engine_advance.argtypes = [p, C.POINTER(u64), C.POINTER(u32)]
    lib.mgbfs_engine_advance.restype = C.c_int
    lib.mgbfs_engine_run.argtypes = [p, u32, C.c_double]
    lib.mgbfs_engine_run.restype = C.c_int
    lib.mgbfs_engine_query.argtypes = [p, C.POINTER(EngineInfo)]
    lib.mgbfs_engine_query.restype = C.c_int
    lib.mgbfs_engine_phases.argtypes = [p, C.POINTER(C.c_double), C.POINTER(u64), u32]
    lib.mgbfs_engine_phases.restype = C.c_int
    lib.mgbfs_engine_reset.argtypes = [p]
    lib.mgbfs_engine_reset.restype = C.c_int
    lib.mgbfs_engine_rebind.argtypes = [p, C.POINTER(GraphDesc)]
    lib.mgbfs_engine_rebind.restype = C.c_int
    lib.mgbfs_engine_layer_sizes.argtypes = [p, C.POINTER(u64), u32, C.POINTER(u32)]
    lib.mgbfs_engine_layer_sizes.restype = C.c_int
    lib.mgbfs_engine_sample.argtypes = [p, u32, u64, C.POINTER(C.c_int64), C.POINTER(u64)]
    lib.mgbfs_engine_sample.restype = C.c_int
    lib.mgbfs_engine_layer_keys.argtypes = [p, C.POINTER(u64), u64, C.POINTER(u64)]
    lib.mgbfs_engine_layer_keys.restype = C.c_int
    lib.mgbfs_engine_backend_name.argtypes = [p]
    lib.mgbfs_engine_backend_name.restype = C.c_char_p
    lib.mgbfs_engine_destroy.argtypes = [p]
    lib.mgbfs_engine_destroy.restype = None
    lib.mgbfs_hip_abi_version.argtypes = []
    lib.mgbfs_hip_abi_version.restype = u32
    lib.mgbfs_hip_device_info.argtypes = [i32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32),
                                          C.POINTER(u64), C.POINTER(u64)]
    lib.mgbfs_hip_device_info.restype = C.c_int
    _LIB = lib
    return lib


def _array(ty, values):
    values = list(values)
    return (ty * max(1, len(values)))(*values)


class NativeEngine:
    """One exact BFS engine instance on one GPU (mgbfs_engine_* wrapper)."""

    def __init__(self, graph: dict, *, device: int = 0, backend: str | int = "auto", capacity: int = 0,
                 vram_budget: int = 0, inverse_closed: bool | None = None, flags: int = 0):
        lib = load()
        desc = graph_desc(graph)
        self.elements = desc.elements
        if inverse_closed is not None:
            desc = desc_with_inverse_closed(desc, inverse_closed)
        self._keep = [desc]
        opts = EngineOptions()
        opts.device = device
        opts.backend = BACKENDS[backend] if isinstance(backend, str) else int(backend)
        opts.capacity = capacity
        opts.vram_budget = vram_budget
        opts.flags = flags
        handle = C.c_void_p()
        err = C.create_string_buffer(256)
        rc = lib.mgbfs_engine_create(C.byref(desc), C.byref(opts), C.byref(handle), err, 256)
        if rc:
            raise RuntimeError(err.value.decode() or f"ENGINE_CREATE_{rc}")
        self._lib, self._h = lib, handle
        self.fatal = 0
        self.status = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mgbfs_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def advance(self) -> int:
        nc, fatal = C.c_uint64(), C.c_uint32()
        rc = self._lib.mgbfs_engine_advance(self._h, C.byref(nc), C.byref(fatal))
        if rc < 0:
            raise RuntimeError(f"ENGINE_ADVANCE_{rc}: {self._lib.mgbfs_engine_backend_name(None).decode()}")
        self.fatal = fatal.value
        self.status = rc
        return rc

    def run(self, max_layers: int = 0, max_seconds: float = 0.0) -> int:
        rc = self._lib.mgbfs_engine_run(self._h, max_layers, max_seconds)
        if rc < 0:
            raise RuntimeError(f"ENGINE_RUN_{rc}: {self._lib.mgbfs_engine_backend_name(None).decode()}")
        self.status = rc
        return rc

    def rebind(self, graph):
        """Switch to another graph of the same shape, keeping the engine's
        device allocations (mgbfs_engine_rebind), and reset."""
        desc = graph_desc(graph)
        rc = self._lib.mgbfs_engine_rebind(self._h, C.byref(desc))
        if rc:
            raise RuntimeError(f"ENGINE_REBIND_{rc}: {self._lib.mgbfs_engine_backend_name(None).decode()}")
        self._keep = [desc]
        self.status = None
        self.fatal = 0

    def reset(self):
        rc = self._lib.mgbfs_engine_reset(self._h)
        if rc:
            raise RuntimeError(f"ENGINE_RESET_{rc}")
        self.status = None
        self.fatal = 0

    def layer_sizes(self) -> list[int]:
        n = C.c_uint32()
        self._lib.mgbfs_engine_layer_sizes(self._h, None, 0, C.byref(n))
        buf = (C.c_uint64 * max(1, n.value))()
        self._lib.mgbfs_engine_layer_sizes(self._h, buf, n.value, C.byref(n))
        return list(buf[:n.value])

    def info(self) -> dict:
        info = EngineInfo()
        self._lib.mgbfs_engine_query(self._h, C.byref(info))
        out = {name: getattr(info, name) for name, _ in EngineInfo._fields_}
        out["backend_name"] = self.backend_name
        return out

    def phases(self, count: int = 4) -> list[tuple[float, int]]:
        secs = (C.c_double * count)()
        byts = (C.c_uint64 * count)()
        self._lib.mgbfs_engine_phases(self._h, secs, byts, count)
        return [(secs[i], byts[i]) for i in range(count)]

    @property
    def backend_name(self) -> str:
        return self._lib.mgbfs_engine_backend_name(self._h).decode()

    def layer_keys(self) -> list[int]:
        n = C.c_uint64()
        cur = self.info()["current_count"]
        buf = (C.c_uint64 * max(1, cur))()
        rc = self._lib.mgbfs_engine_layer_keys(self._h, buf, cur, C.byref(n))
        if rc:
            raise RuntimeError(f"ENGINE_LAYER_KEYS_{rc}")
        return list(buf[:n.value])

    def sample(self, which: int = 0, limit: int = 1000) -> list[list[int]]:
        sizes = self.layer_sizes()
        available = sizes[-1] if which == 0 else (sizes[-2] if len(sizes) > 1 else 0)
        limit = min(limit, available)
        rows = (C.c_int64 * max(1, limit * self.elements))()
        n = C.c_uint64()
        rc = self._lib.mgbfs_engine_sample(self._h, which, limit, rows, C.byref(n))
        if rc:
            raise RuntimeError(f"ENGINE_SAMPLE_{rc}")
        w = self.elements
        flat = rows[:n.value * w]
        return [list(flat[i * w:(i + 1) * w]) for i in range(n.value)]


def device_info(device: int = 0) -> dict:
    lib = load()
    cu, ws, lds = C.c_uint32(), C.c_uint32(), C.c_uint32()
    tot, fr = C.c_uint64(), C.c_uint64()
    rc = lib.mgbfs_hip_device_info(device, C.byref(cu), C.byref(ws), C.byref(lds), C.byref(tot), C.byref(fr))
    if rc:
        raise RuntimeError(f"HIP_DEVICE_INFO_{rc}")
    return dict(compute_units=cu.value, wave_size=ws.value, lds_bytes=lds.value, total_bytes=tot.value,
                free_bytes=fr.value)


HISTORY = {"auto": 0, "full": 1, "rolling": 2}


class BatchOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_slots", C.c_uint32), ("vram_budget", C.c_uint64),
                ("layers_per_graph", C.c_uint32), ("grid", C.c_uint32), ("threads", C.c_uint32),
                ("hash_load_pct", C.c_uint32), ("history", C.c_uint32), ("layer_pct", C.c_uint32),
                ("bitmap_bias_permille", C.c_uint32)]


class SortedPlanInfo(C.Structure):
    _fields_ = [("workspace_bytes", C.c_uint64), ("candidates_bound", C.c_uint64), ("sub_buckets", C.c_uint32),
                ("next_buckets", C.c_uint32), ("masked", C.c_uint32), ("history_banks", C.c_uint32)]


class SortedLayerArgs(C.Structure):
    _fields_ = [("degree", C.c_uint32), ("generators", C.c_uint32), ("permutations", C.POINTER(C.c_uint32)),
                ("depth", C.c_uint32), ("sub_buckets", C.c_uint32), ("slack", C.c_uint32), ("buckets", C.c_uint32),
                ("bank", C.c_void_p), ("offsets", C.c_void_p), ("count", C.c_uint64),
                ("previous", C.c_void_p), ("previous_offsets", C.c_void_p), ("previous_buckets", C.c_uint32),
                ("reserved", C.c_uint32), ("next", C.c_void_p), ("next_offsets", C.c_void_p),
                ("next_capacity", C.c_uint64), ("counters", C.c_void_p), ("error", C.c_void_p)]


SORTED_OFFSETS = (1 << 20) + 1
SORTED_SEED_BUCKETS = 1 << 20


def bind_sorted(lib):
    """Signatures of the kernel-level sorted-bank layer ABI (mgbfs_sorted_*)."""
    if getattr(lib, "_sorted_bound", False):
        return lib
    p, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
    lib.mgbfs_sorted_plan.argtypes = [u32, u32, C.POINTER(u32), u32, u64, u32, u32, C.POINTER(SortedPlanInfo)]
    lib.mgbfs_sorted_seed.argtypes = [u32, u64, p, p, p]
    lib.mgbfs_sorted_layer.argtypes = [C.POINTER(SortedLayerArgs), p, u64, p]
    lib.mgbfs_sorted_unpack.argtypes = [u32, C.POINTER(u64), u64, C.POINTER(u64)]
    for name in ("plan", "seed", "layer", "unpack"):
        getattr(lib, "mgbfs_sorted_" + name).restype = C.c_int
    lib._sorted_bound = True
    return lib


MGBFS_BATCH_FULL = 3
VISITED = {"auto": 0, "bitmap": 1, "hash": 2}


def _bind_batch(lib):
    if getattr(lib, "_batch_bound", False):
        return lib
    p, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32
    lib.mgbfs_batch_create.argtypes = [C.POINTER(BatchOptions), C.POINTER(p), C.c_char_p, C.c_size_t]
    lib.mgbfs_batch_plan.argtypes = [p, C.POINTER(GraphDesc), u32, u32, C.POINTER(u64), C.POINTER(u32)]
    lib.mgbfs_batch_admit.argtypes = [p, C.POINTER(GraphDesc), u32, u32, C.POINTER(i32)]
    lib.mgbfs_batch_step.argtypes = [p, u32, C.POINTER(u32)]
    lib.mgbfs_batch_run.argtypes = [p]
    lib.mgbfs_batch_poll.argtypes = [p, C.POINTER(i32), u32, C.POINTER(u32)]
    lib.mgbfs_batch_result.argtypes = [p, i32, C.POINTER(u64), u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    lib.mgbfs_batch_sample.argtypes = [p, i32, u64, C.POINTER(C.c_int64), C.POINTER(u64)]
    lib.mgbfs_batch_release.argtypes = [p, i32]
    lib.mgbfs_batch_usage.argtypes = [p, C.POINTER(u64), C.POINTER(u64), C.POINTER(u32)]
    lib.mgbfs_batch_instance_info.argtypes = [p, i32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u64),
                                              C.POINTER(u32)]
    lib.mgbfs_batch_destroy.argtypes = [p]
    lib.mgbfs_batch_destroy.restype = None
    for name in ("create", "plan", "admit", "step", "run", "poll", "result", "sample", "release", "usage",
                 "instance_info"):
        getattr(lib, "mgbfs_batch_" + name).restype = C.c_int
    lib._batch_bound = True
    return lib


def desc_with_inverse_closed(desc, inverse_closed):
    """A copy of `desc` with its inverse-closure flag overridden; the graph's
    own (cached) descriptor stays as validated.  The copy shares the arrays."""
    out = GraphDesc.from_buffer_copy(desc)
    out._keep = desc._keep
    out.inverse_closed = int(inverse_closed)
    return out


def graph_desc(graph):
    """mgbfs_graph_desc for a GraphDefinition (arrays kept alive on the
    descriptor).  A GraphDefinition's contents never change after validation,
    so its descriptor is built once and kept on it; the library copies what
    it needs from a descriptor during each call."""
    from .graph_definition import GraphDefinition
    gd = graph if isinstance(graph, GraphDefinition) else GraphDefinition(graph)
    cached = gd.__dict__.get("_desc")
    if cached is not None:
        return cached
    data = gd._data  # read only
    a = data["action"]
    desc = GraphDesc()
    desc.elements = len(data["start"])
    desc.generators = len(a["generators"])
    desc.inverse_closed = int(gd.inverse_closed())
    desc.expected_states = data["expected_max_unique_states"] or 0
    keep = [_array(C.c_int64, data["start"])]
    desc.start = keep[0]
    if a["kind"] == "permutation":
        desc.kind, desc.rows, desc.cols = 0, desc.elements, 1
        keep.append(_array(C.c_uint32, [v for g in a["generators"] for v in g]))
        desc.permutations = keep[-1]
    else:
        desc.kind, desc.rows, desc.cols = 1, a["rows"], a["cols"]
        keep.append(_array(C.c_int64, [v for g in a["generators"] for v in g["matrix"]]))
        desc.matrices = keep[-1]
        keep.append(_array(C.c_uint32, [g["modulo"] for g in a["generators"]]))
        desc.moduli = keep[-1]
    desc._keep = keep
    gd._desc = desc
    return desc


class BatchEngine:
    """Many BFS instances on one GPU, advanced together (mgbfs_batch_*)."""

    def __init__(self, *, device=0, max_slots=256, vram_budget=0, layers_per_graph=0, grid=0, threads=0,
                 hash_load_pct=0, history="auto", layer_pct=0, bitmap_bias_permille=0):
        """history of inverse-closed hash instances: "auto" (full history
        when it fits the free budget, else three rolling tables), "full" or
        "rolling"; zeros take the engine's defaults (mgbfs_batch_options in
        include/mgbfs_hip.h)."""
        self._lib = _bind_batch(load())
        if history not in HISTORY:
            raise ValueError("BATCH_OPTION_HISTORY")
        opts = BatchOptions(device, max_slots, vram_budget, layers_per_graph, grid, threads, hash_load_pct,
                            HISTORY[history], layer_pct, bitmap_bias_permille)
        h = C.c_void_p()
        err = C.create_string_buffer(256)
        if self._lib.mgbfs_batch_create(C.byref(opts), C.byref(h), err, 256):
            raise RuntimeError(err.value.decode() or "BATCH_CREATE")
        self._h = h
        self.max_slots = max_slots

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mgbfs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError(f"{what}: {self._lib.mgbfs_engine_backend_name(None).decode()}")
        return rc

    def plan(self, graph, visited="auto", max_layers=0):
        nbytes, chosen = C.c_uint64(), C.c_uint32()
        self._check(self._lib.mgbfs_batch_plan(self._h, C.byref(graph_desc(graph)), VISITED[visited], max_layers,
                                               C.byref(nbytes), C.byref(chosen)), "BATCH_PLAN")
        return nbytes.value, {1: "bitmap", 2: "hash"}[chosen.value]

    def admit(self, graph, visited="auto", max_layers=0):
        """Slot index, or None when no slot or VRAM budget is free."""
        slot = C.c_int32()
        rc = self._check(self._lib.mgbfs_batch_admit(self._h, C.byref(graph_desc(graph)), VISITED[visited],
                                                     max_layers, C.byref(slot)), "BATCH_ADMIT")
        return None if rc == MGBFS_BATCH_FULL else slot.value

    def step(self, layers=4):
        active = C.c_uint32()
        self._check(self._lib.mgbfs_batch_step(self._h, layers, C.byref(active)), "BATCH_STEP")
        return active.value

    def run(self):
        self._check(self._lib.mgbfs_batch_run(self._h), "BATCH_RUN")

    def poll(self):
        buf = (C.c_int32 * self.max_slots)()
        n = C.c_uint32()
        self._check(self._lib.mgbfs_batch_poll(self._h, buf, self.max_slots, C.byref(n)), "BATCH_POLL")
        return list(buf[:n.value])

    def result(self, slot):
        n, status, error = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self._lib.mgbfs_batch_result(self._h, slot, None, 0, C.byref(n), C.byref(status),
                                                 C.byref(error)), "BATCH_RESULT")
        buf = (C.c_uint64 * max(1, n.value))()
        self._check(self._lib.mgbfs_batch_result(self._h, slot, buf, n.value, C.byref(n), C.byref(status),
                                                 C.byref(error)), "BATCH_RESULT")
        return list(buf[:n.value]), status.value, error.value

    def sample(self, slot, width, limit=1000):
        sizes, _, _ = self.result(slot)
        limit = min(limit, sizes[-1] if sizes else 0)
        rows = (C.c_int64 * max(1, limit * width))()
        n = C.c_uint64()
        self._check(self._lib.mgbfs_batch_sample(self._h, slot, limit, rows, C.byref(n)), "BATCH_SAMPLE")
        flat = rows[:n.value * width]
        return [list(flat[i * width:(i + 1) * width]) for i in range(n.value)]

    def release(self, slot):
        self._check(self._lib.mgbfs_batch_release(self._h, slot), "BATCH_RELEASE")

    def instance_info(self, slot):
        """{visited: bitmap|hash, history_layers: 3 (rolling) or 0 (full),
        layer_capacity, restarts} of an admitted slot."""
        v, h, r = C.c_uint32(), C.c_uint32(), C.c_uint32()
        cap = C.c_uint64()
        self._check(self._lib.mgbfs_batch_instance_info(self._h, slot, C.byref(v), C.byref(h), C.byref(cap),
                                                        C.byref(r)), "BATCH_INSTANCE_INFO")
        return {"visited": "bitmap" if v.value == 1 else "hash", "history_layers": h.value,
                "layer_capacity": cap.value, "restarts": r.value}

    def usage(self):
        used, budget, slots = C.c_uint64(), C.c_uint64(), C.c_uint32()
        self._lib.mgbfs_batch_usage(self._h, C.byref(used), C.byref(budget), C.byref(slots))
        return dict(used=used.value, budget=budget.value, slots=slots.value)
