"""Minimal HIP runtime access (device buffers, copies, sync) through ctypes,
so that benchmarks can hand device pointers to the C ABI without torch."""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """An int32 buffer in device memory."""

    def __init__(self, n: int):
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.n, 1) * 4), "hipMalloc")

    @classmethod
    def from_numpy(cls, x) -> "DeviceArray":
        x = np.ascontiguousarray(np.asarray(x, np.int32))
        d = cls(x.size)
        if x.size:
            _ok(hip().hipMemcpy(d.ptr, x.ctypes.data, x.size * 4, 1), "hipMemcpy H2D")
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(max(self.n, 1), np.int32)
        if self.n:
            _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.n * 4, 2), "hipMemcpy D2H")
        return out[: self.n]

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A HIP stream (default priority)."""

    def __init__(self):
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(self.ptr)), "hipStreamCreate")

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def close(self) -> None:
        if self.ptr:
            hip().hipStreamDestroy(self.ptr)
            self.ptr = ctypes.c_void_p()
