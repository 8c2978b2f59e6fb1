"""Minimal HIP runtime access (device buffers, copies, events) via ctypes.

Used by bench.py to keep queries, proxies and results resident in HBM and to
time kernels with HIP events recorded on the library's own stream.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        _hip = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
        vp = C.c_void_p
        _hip.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        _hip.hipMemcpyAsync.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
        _hip.hipEventCreate.argtypes = [C.POINTER(vp)]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
        _hip.hipHostMalloc.argtypes = [C.POINTER(vp), C.c_size_t, C.c_uint]
        _hip.hipHostFree.argtypes = [vp]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
        _hip.hipHostGetDevicePointer.argtypes = [C.POINTER(vp), vp, C.c_uint]
    return _hip


def device_name(device: int = 0) -> str:
    """The HIP device name (hipDeviceGetName), without starting any process."""
    buf = C.create_string_buffer(256)
    h = hip()
    h.hipDeviceGetName.argtypes = [C.c_char_p, C.c_int, C.c_int]
    if h.hipDeviceGetName(buf, 256, device) != 0:
        return "gpu"
    return buf.value.decode(errors="replace").strip() or "gpu"


def _ck(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A device allocation with the shape/dtype of a host array."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if not isinstance(shape, int) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ck(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ck(hip().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def at(self, elem: int) -> C.c_void_p:
        """Device address of element `elem` (flat index)."""
        return C.c_void_p(self.ptr.value + int(elem) * self.dtype.itemsize)

    def copy_from_host_async(self, a: np.ndarray, stream, elem: int = 0) -> None:
        """Queue a copy of host array a to element `elem` on `stream`."""
        a = np.ascontiguousarray(a, self.dtype)
        if (int(elem) + a.size) * self.dtype.itemsize > self.nbytes:
            raise ValueError("copy_from_host_async: out of range")
        _ck(hip().hipMemcpyAsync(self.at(elem), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D, stream),
            "hipMemcpyAsync H2D")

    def copy_from_device(self, src: "DeviceArray", nbytes: int) -> None:
        """Synchronous device-to-device copy of nbytes from src."""
        if nbytes > min(self.nbytes, src.nbytes):
            raise ValueError("copy_from_device: out of range")
        _ck(hip().hipMemcpy(self.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _ck(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory (hipHostMalloc) viewed as a numpy array: copies
    between it and the device run at the link's rate without staging."""

    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        shape = (shape,) if isinstance(shape, int) else tuple(shape)
        nbytes = max(int(np.prod(shape)) * self.dtype.itemsize, 1)
        p = C.c_void_p()
        _ck(hip().hipHostMalloc(C.byref(p), nbytes, 0), "hipHostMalloc")
        self.ptr = p
        buf = (C.c_char * nbytes).from_address(p.value)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(shape))).reshape(shape)

    def device_ptr(self) -> C.c_void_p:
        """The address kernels use to read or write this memory directly over
        the link (hipHostGetDevicePointer)."""
        dp = C.c_void_p()
        _ck(hip().hipHostGetDevicePointer(C.byref(dp), self.ptr, 0), "hipHostGetDevicePointer")
        return dp

    def free(self):
        if self.ptr:
            self.a = None
            hip().hipHostFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_async(dst, src, nbytes: int, kind: int, stream) -> None:
    """hipMemcpyAsync of nbytes between raw addresses on `stream`."""
    _ck(hip().hipMemcpyAsync(dst, src, int(nbytes), kind, stream), "hipMemcpyAsync")


class Event:
    def __init__(self):
        e = C.c_void_p()
        _ck(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.e = e

    def record(self, stream):
        _ck(hip().hipEventRecord(self.e, stream), "hipEventRecord")

    def sync(self):
        _ck(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def ms_since(self, start: "Event") -> float:
        v = C.c_float()
        _ck(hip().hipEventElapsedTime(C.byref(v), start.e, self.e), "hipEventElapsedTime")
        return float(v.value)


def stream_sync(stream):
    _ck(hip().hipStreamSynchronize(stream), "hipStreamSynchronize")
