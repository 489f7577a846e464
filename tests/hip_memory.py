"""Caller-owned device memory for tests of the kernel-level C ABI.

The kernel-level entry points take device pointers that the caller
allocated (the reference's convention, cuda/generic_visited.h:13-21); these
helpers call the HIP runtime directly through ctypes, as a foreign caller of
the library would.
"""
import ctypes as C
import struct

_HIP = None
H2D, D2H = 1, 2


def hip():
    global _HIP
    if _HIP is None:
        lib = C.CDLL("libamdhip64.so")
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipSetDevice.argtypes = [C.c_int]
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize", "hipSetDevice"):
            getattr(lib, name).restype = C.c_int
        _HIP = lib
    return _HIP


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} returned {rc}")


class DeviceArray:
    """hipMalloc'd buffer of `count` elements of a struct format char."""

    def __init__(self, count, fmt="Q"):
        self.fmt, self.count = fmt, count
        self.itemsize = struct.calcsize(fmt)
        self.nbytes = max(8, count * self.itemsize)
        self.ptr = C.c_void_p()
        check(hip().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")

    def zero(self):
        check(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        return self

    def read(self, count=None, offset=0):
        count = self.count - offset if count is None else count
        buf = C.create_string_buffer(max(1, count * self.itemsize))
        if count:
            check(hip().hipMemcpy(buf, self.ptr.value + offset * self.itemsize, count * self.itemsize, D2H),
                  "hipMemcpy D2H")
        return list(struct.unpack(f"<{count}{self.fmt}", buf.raw[:count * self.itemsize]))

    def write(self, values):
        data = struct.pack(f"<{len(values)}{self.fmt}", *values)
        check(hip().hipMemcpy(self.ptr, data, len(data), H2D), "hipMemcpy H2D")
        return self

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
