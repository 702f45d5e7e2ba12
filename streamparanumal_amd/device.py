"""Device, memory, streams, events and graphs over libspx (include/spx.h).

Mirrors the reference's platform_t / deviceMemory<T> surface
(include/platform.hpp:60-176, include/memory.hpp:285-460) for Python callers.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, libspx

_DTYPES = {np.dtype(np.float32): _lib.FLOAT, np.dtype(np.float64): _lib.DOUBLE,
           np.dtype(np.int32): _lib.INT32, np.dtype(np.int64): _lib.INT64}


def ogs_type(dtype):
    """ogs::Type code of a numpy dtype (float, double, int32, int64)."""
    dt = np.dtype(dtype)
    if dt not in _DTYPES:
        raise TypeError(f"ogs supports float32, float64, int32, int64; got {dt}")
    return _DTYPES[dt]


def device_count():
    c = C.c_int(0)
    check(libspx().spx_device_count(C.byref(c)), "spx_device_count")
    return c.value


def open_device(device_id=0):
    """Open the HIP device (occa::device::setup, platformDeviceConfig.cpp:159)."""
    check(libspx().spx_device_open(device_id), "spx_device_open")


def finish():
    """Wait for all enqueued work (occa::device::finish)."""
    check(libspx().spx_device_finish(), "spx_device_finish")


def props():
    p = _lib.DeviceProps()
    check(libspx().spx_device_props(C.byref(p)), "spx_device_props")
    return {"name": p.name.decode(), "arch": p.arch.decode(), "compute_units": p.compute_units,
            "clock_khz": p.clock_khz, "memory_clock_khz": p.memory_clock_khz,
            "memory_bus_width": p.memory_bus_width, "global_mem_bytes": p.global_mem_bytes,
            "l2_bytes": p.l2_bytes, "lds_bytes_per_block": p.lds_bytes_per_block, "warp_size": p.warp_size}


def check_kernel_errors():
    check(libspx().spx_check_kernel_errors(), "spx_check_kernel_errors")


class DeviceArray:
    """A device allocation of ``n`` elements of ``dtype`` (deviceMemory<T>)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self._ptr = C.c_void_p()
        self.nbytes = self.n * self.dtype.itemsize
        check(libspx().spx_malloc(C.byref(self._ptr), self.nbytes), "spx_malloc")
        self._owner = True

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.copy_from(a)
        return d

    @property
    def ptr(self):
        return self._ptr.value

    def offset(self, k):
        """Raw address of element k (occa::memory::operator+)."""
        if not 0 <= k <= self.n:
            raise IndexError(f"offset {k} outside [0, {self.n}]")
        return (self._ptr.value or 0) + k * self.dtype.itemsize

    def copy_from(self, a, asynchronous=False):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError(f"size mismatch: {a.size} vs {self.n}")
        if self.n:
            check(libspx().spx_memcpy(self.ptr, a.ctypes.data, self.nbytes, 1 if asynchronous else 0),
                  "spx_memcpy")

    def to_numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            check(libspx().spx_memcpy(out.ctypes.data, self.ptr, self.nbytes, 0), "spx_memcpy")
        return out

    def read(self, start, count):
        """Elements [start, start + count) copied to a new host array."""
        if not (0 <= start and 0 <= count and start + count <= self.n):
            raise IndexError(f"range [{start}, {start + count}) outside [0, {self.n}]")
        out = np.empty(count, dtype=self.dtype)
        if count:
            check(libspx().spx_memcpy(out.ctypes.data, self.offset(start), count * self.dtype.itemsize, 0),
                  "spx_memcpy")
        return out

    def fill_zero(self):
        check(libspx().spx_fill_bytes(self.ptr, self.nbytes, 0), "spx_fill_bytes")

    def free(self):
        if self._owner and self._ptr.value:
            libspx().spx_free(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:  # interpreter shutdown
            pass


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        check(libspx().spx_event_create(C.byref(self.h)), "spx_event_create")

    def record(self):
        check(libspx().spx_event_record(self.h), "spx_event_record")

    def elapsed_ms(self, stop):
        ms = C.c_float()
        check(libspx().spx_event_elapsed_ms(self.h, stop.h, C.byref(ms)), "spx_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if self.h.value:
                libspx().spx_event_destroy(self.h)
        except Exception:
            pass


class Graph:
    """A captured, replayable sequence of launches on the current stream."""

    def __init__(self, h):
        self.h = h

    @classmethod
    def capture(cls, fn):
        check(libspx().spx_graph_begin(), "spx_graph_begin")
        try:
            fn()
        finally:
            h = C.c_void_p()
            check(libspx().spx_graph_end(C.byref(h)), "spx_graph_end")
        return cls(h)

    def launch(self):
        check(libspx().spx_graph_launch(self.h), "spx_graph_launch")

    def __del__(self):
        try:
            if self.h.value:
                libspx().spx_graph_destroy(self.h)
        except Exception:
            pass


def set_tuning(kernel, block_threads=0, xcd_remap=-1, unroll=0, grid_blocks=0, nontemporal=-1, variant=-1):
    """Launch parameters of one kernel family (0 / -1 keep the library default)."""
    t = _lib.Tuning(block_threads, xcd_remap, unroll, grid_blocks, nontemporal, variant)
    check(libspx().spx_set_tuning(kernel.encode(), C.byref(t)), "spx_set_tuning")


def fill_uniform(arr, seed, skip=0):
    """U[-1,1) from the reference LCG (random.cpp:45-163), element i = draw skip+i."""
    check(libspx().spx_fill_uniform_f64(arr.n, arr.ptr, seed, skip), "spx_fill_uniform_f64")


def fill_loguniform(arr, seed, skip=0):
    """+-10^(-6u), |x| log-uniform in [1e-6, 1], sign from the next draw's top bit: element i uses
    draws 2(skip+i) and 2(skip+i)+1 of the reference LCG."""
    check(libspx().spx_fill_loguniform_f64(arr.n, arr.ptr, seed, skip), "spx_fill_loguniform_f64")


def convert(src, dst, n=None):
    """Convert between FP64 and a storage type on the device: float64 -> float32 / uint16 (bf16,
    nearest even), or float32 / uint16 -> float64 (exact)."""
    n = src.n if n is None else n
    if n > min(src.n, dst.n):
        raise ValueError("convert: n exceeds an array")
    pair = (src.dtype, dst.dtype)
    fns = {(np.dtype(np.float64), np.dtype(np.float32)): "spx_convert_f64_to_f32",
           (np.dtype(np.float64), np.dtype(np.uint16)): "spx_convert_f64_to_bf16",
           (np.dtype(np.float32), np.dtype(np.float64)): "spx_convert_f32_to_f64",
           (np.dtype(np.uint16), np.dtype(np.float64)): "spx_convert_bf16_to_f64"}
    if pair not in fns:
        raise TypeError(f"convert: no conversion {pair[0]} -> {pair[1]}")
    check(getattr(libspx(), fns[pair])(n, src.ptr, dst.ptr), fns[pair])
