"""A foreign caller's layer loop over the kernel-level sorted-bank ABI
(include/mgbfs_hip.h, mgbfs_sorted_*): hipMalloc'd banks, offsets,
counters, error word and workspace, one mgbfs_sorted_layer call per layer,
the error-bit retries the header describes, and rotation of three banks.
The layer call itself is a parameter, so other entry points of the same
contract can be driven by the same loop."""
import ctypes as C

from hip_memory import DeviceArray, synchronize
from multigpubfs_amd import _native as N

ERR_FUTURE, ERR_BUCKET, ERR_LANE = 4, 64, 256


def pack(state):
    return sum(int(v) << (4 * i) for i, v in enumerate(state))


class SortedAbiRun:
    def __init__(self, generators, start, *, coarse=False, initial_rows=0, layer=None):
        self.lib = N.bind_sorted(N.load())
        self.degree = len(start)
        self.gens = [list(g) for g in generators]
        self.table = (C.c_uint32 * (len(self.gens) * self.degree))(*[v for g in self.gens for v in g])
        self.coarse, self.initial_rows = coarse, initial_rows
        self.layer_call = layer or self.workspace_layer
        self.banks = [DeviceArray(max(1, initial_rows or 1024)) for _ in range(3)]
        self.offsets = [DeviceArray(N.SORTED_OFFSETS, "I") for _ in range(3)]
        self.buckets = [0, 0, 0]
        self.counts = [0, 0, 0]
        self.counters = DeviceArray(2)
        self.error = DeviceArray(1, "I")
        self.workspace = None
        self.retries = {ERR_FUTURE: 0, ERR_BUCKET: 0, ERR_LANE: 0}
        self.candidates = 0
        rc = self.lib.mgbfs_sorted_seed(self.degree, pack(start), self.banks[0].ptr, self.offsets[0].ptr, None)
        assert rc == 0, rc
        synchronize()
        self.buckets[0], self.counts[0] = N.SORTED_SEED_BUCKETS, 1
        self.cur, self.prev, self.depth = 0, -1, 0
        self.layers = [1]

    def plan(self, count, sub_buckets=0, slack=0):
        info = N.SortedPlanInfo()
        rc = self.lib.mgbfs_sorted_plan(self.degree, len(self.gens), self.table, self.depth, count, sub_buckets,
                                        slack, C.byref(info))
        assert rc == 0, rc
        return info

    def workspace_layer(self, args, info):
        if self.workspace is None or self.workspace.nbytes < info.workspace_bytes:
            if self.workspace is not None:
                self.workspace.free()
            self.workspace = DeviceArray(info.workspace_bytes, "B")
        return self.lib.mgbfs_sorted_layer(C.byref(args), self.workspace.ptr, self.workspace.nbytes, None)

    def keys(self, bank=None):
        bank = self.cur if bank is None else bank
        values = self.banks[bank].read(self.counts[bank])
        arr = (C.c_uint64 * max(1, len(values)))(*values)
        out = (C.c_uint64 * max(1, len(values)))()
        assert self.lib.mgbfs_sorted_unpack(self.degree, arr, len(values), out) == 0
        return sorted(out[:len(values)])

    def advance(self):
        """One layer; returns its size (0 when the BFS is complete)."""
        nxt = next(b for b in range(3) if b not in (self.cur, self.prev))
        F = self.counts[self.cur]
        sub, slack = 0, 0
        if self.coarse:
            sub = max(1, self.plan(F).sub_buckets // 8)
        while True:
            info = self.plan(F, sub, slack)
            if not self.initial_rows and self.banks[nxt].count < info.candidates_bound:
                self.banks[nxt].free()
                self.banks[nxt] = DeviceArray(info.candidates_bound)
            args = N.SortedLayerArgs()
            args.degree, args.generators, args.permutations = self.degree, len(self.gens), self.table
            args.depth, args.sub_buckets, args.slack = self.depth, sub, slack
            args.bank, args.offsets = self.banks[self.cur].ptr, self.offsets[self.cur].ptr
            args.buckets, args.count = self.buckets[self.cur], F
            if self.prev >= 0:
                args.previous, args.previous_offsets = self.banks[self.prev].ptr, self.offsets[self.prev].ptr
                args.previous_buckets = self.buckets[self.prev]
            args.next, args.next_offsets = self.banks[nxt].ptr, self.offsets[nxt].ptr
            args.next_capacity = self.banks[nxt].count
            args.counters, args.error = self.counters.zero().ptr, self.error.zero().ptr
            rc = self.layer_call(args, info)
            assert rc == 0, rc
            synchronize()
            err = self.error.read(1)[0]
            total = self.offsets[nxt].read(1, offset=info.next_buckets)[0]
            if err & ERR_BUCKET:
                self.retries[ERR_BUCKET] += 1
                sub = 2 * info.sub_buckets
                continue
            if err & ERR_LANE:
                self.retries[ERR_LANE] += 1
                slack = 2 * max(1, slack)
                continue
            if err & ERR_FUTURE:
                self.retries[ERR_FUTURE] += 1
                self.banks[nxt].free()
                self.banks[nxt] = DeviceArray(total)
                continue
            assert err == 0, err
            break
        self.candidates += self.counters.read(2)[1]
        if total == 0:
            return 0
        self.counts[nxt], self.buckets[nxt] = total, info.next_buckets
        self.prev, self.cur = self.cur, nxt
        self.depth += 1
        self.layers.append(total)
        return total

    def run(self, keep_keys=False):
        keys = [self.keys()] if keep_keys else None
        while self.advance():
            if keep_keys:
                keys.append(self.keys())
        return self.layers, keys

    def close(self):
        for a in self.banks + self.offsets + [self.counters, self.error, self.workspace]:
            if a is not None:
                a.free()
