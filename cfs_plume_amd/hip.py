"""Minimal ctypes binding of the HIP runtime: device buffers, streams, events.

It binds the same libamdhip64 instance that libcfs_plume.so is linked
against, so handles and device pointers are shared with the C ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3

_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        _lib.lib()  # resolves libamdhip64.so.7 first, as the C ABI sees it
        h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        sig = {
            "hipMalloc": (i, [ctypes.POINTER(vp), sz]),
            "hipFree": (i, [vp]),
            "hipMemcpy": (i, [vp, vp, sz, i]),
            "hipMemcpyAsync": (i, [vp, vp, sz, i, vp]),
            "hipMemset": (i, [vp, i, sz]),
            "hipDeviceSynchronize": (i, []),
            "hipGetDeviceCount": (i, [ctypes.POINTER(i)]),
            "hipSetDevice": (i, [i]),
            "hipDeviceGetAttribute": (i, [ctypes.POINTER(i), i, i]),
            "hipGetErrorString": (ctypes.c_char_p, [i]),
            "hipStreamCreate": (i, [ctypes.POINTER(vp)]),
            "hipStreamDestroy": (i, [vp]),
            "hipStreamSynchronize": (i, [vp]),
            "hipEventCreate": (i, [ctypes.POINTER(vp)]),
            "hipEventDestroy": (i, [vp]),
            "hipEventRecord": (i, [vp, vp]),
            "hipEventSynchronize": (i, [vp]),
            "hipEventElapsedTime": (i, [ctypes.POINTER(ctypes.c_float), vp, vp]),
            "hipStreamBeginCapture": (i, [vp, i]),
            "hipStreamEndCapture": (i, [vp, ctypes.POINTER(vp)]),
            "hipGraphInstantiate": (i, [ctypes.POINTER(vp), vp, vp, ctypes.c_char_p, sz]),
            "hipGraphLaunch": (i, [vp, vp]),
            "hipGraphExecDestroy": (i, [vp]),
            "hipGraphDestroy": (i, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _hip = h
    return _hip


class HipError(RuntimeError):
    pass


def check(err: int, what: str):
    if err != 0:
        msg = hip().hipGetErrorString(err).decode()
        raise HipError(f"{what} failed: {msg} ({err})")


def device_count() -> int:
    n = ctypes.c_int(0)
    err = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if err == 0 else 0


def gpu_available() -> bool:
    try:
        return device_count() > 0
    except OSError:
        return False


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def cu_count(device: int = 0) -> int:
    v = ctypes.c_int(0)
    check(hip().hipDeviceGetAttribute(ctypes.byref(v), 63, device), "hipDeviceGetAttribute")  # MultiprocessorCount
    return v.value


class DeviceArray:
    """A device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, int) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 4)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def view(cls, ptr: int, shape, dtype, owner=None) -> "DeviceArray":
        """A non-owning view of device memory held by `owner` (kept alive)."""
        d = cls.__new__(cls)
        d.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, int) else (int(shape),)
        d.dtype = np.dtype(dtype)
        d.nbytes = int(np.prod(d.shape)) * d.dtype.itemsize
        d.ptr = ptr
        d._view = True
        d._owner = owner
        return d

    @classmethod
    def zeros(cls, shape, dtype) -> "DeviceArray":
        d = cls(shape, dtype)
        d.fill_zero()
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, hipMemcpyDeviceToHost), "hipMemcpy D2H")
        return out

    def fill_zero(self):
        check(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if getattr(self, "_view", False):  # the owner frees the memory
            self.ptr = None
            return
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle:
                hip().hipStreamDestroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        check(hip().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(hip().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                hip().hipEventDestroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Graph:
    """A captured HIP graph of work issued on `stream` inside `fn()`."""

    def __init__(self, stream: Stream, fn):
        check(hip().hipStreamBeginCapture(stream.handle, 0), "hipStreamBeginCapture")  # global mode
        try:
            fn()
        finally:
            g = ctypes.c_void_p()
            err = hip().hipStreamEndCapture(stream.handle, ctypes.byref(g))
        check(err, "hipStreamEndCapture")
        self.graph = g.value
        ex = ctypes.c_void_p()
        check(hip().hipGraphInstantiate(ctypes.byref(ex), self.graph, None, None, 0), "hipGraphInstantiate")
        self.exec = ex.value
        self.stream = stream

    def launch(self):
        check(hip().hipGraphLaunch(self.exec, self.stream.handle), "hipGraphLaunch")

    def __del__(self):
        try:
            if self.exec:
                hip().hipGraphExecDestroy(self.exec)
            if self.graph:
                hip().hipGraphDestroy(self.graph)
        except Exception:
            pass
