"""Minimal ctypes binding of the HIP runtime (libamdhip64): device buffers, 2-D copies, streams, events.

The denoiser's C ABI takes plain device pointers and byte pitches; this module is the plumbing that lets
Python tests, the benchmark and the trainer allocate, fill and read such buffers without torch.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


class HipError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        candidates = ["libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"]
        err = None
        for c in candidates:
            try:
                _hip = ctypes.CDLL(c)
                break
            except OSError as e:  # pragma: no cover - depends on the image
                err = e
        if _hip is None:
            raise HipError(f"cannot load the HIP runtime: {err}")
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _hip.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipGraphInstantiate.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_char_p, ctypes.c_size_t]
        _hip.hipGraphLaunch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipGraphExecDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipGraphDestroy.argtypes = [ctypes.c_void_p]
    return _hip


def check(code: int, what: str) -> None:
    if code != 0:
        raise HipError(f"{what}: {lib().hipGetErrorString(code).decode()} ({code})")


def device_count() -> int:
    n = ctypes.c_int(0)
    r = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if r == 0 else 0


def synchronize() -> None:
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd block of device memory."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    def free(self) -> None:
        if self.ptr:
            lib().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.free()
        except Exception:
            pass


class Surface:
    """A row-major 2-D device surface of `channels` x `dtype` texels with a byte pitch.

    Either owns its memory (allocated here) or views memory owned elsewhere (ptr given).
    """

    def __init__(self, width: int, height: int, dtype=np.float16, channels: int = 4, ptr: int | None = None,
                 pitch: int | None = None, owner=None):
        self.width, self.height = int(width), int(height)
        self.dtype = np.dtype(dtype)
        self.channels = channels
        self.texel_bytes = self.dtype.itemsize * channels
        if ptr is None:
            self.pitch = pitch or ((self.width * self.texel_bytes + 255) // 256 * 256)
            self._buf = DeviceBuffer(self.pitch * self.height)
            self.ptr = self._buf.ptr
        else:
            self.pitch = int(pitch)
            self._buf = owner
            self.ptr = int(ptr)

    def upload(self, host: np.ndarray) -> "Surface":
        a = np.ascontiguousarray(host)
        assert a.nbytes == self.width * self.height * self.texel_bytes, (a.shape, a.dtype, self.width, self.height)
        check(lib().hipMemcpy2D(ctypes.c_void_p(self.ptr), self.pitch, a.ctypes.data_as(ctypes.c_void_p),
                                self.width * self.texel_bytes, self.width * self.texel_bytes, self.height,
                                hipMemcpyHostToDevice), "hipMemcpy2D H2D")
        return self

    def download(self) -> np.ndarray:
        out = np.empty((self.height, self.width, self.channels), self.dtype)
        check(lib().hipMemcpy2D(out.ctypes.data_as(ctypes.c_void_p), self.width * self.texel_bytes,
                                ctypes.c_void_p(self.ptr), self.pitch, self.width * self.texel_bytes, self.height,
                                hipMemcpyDeviceToHost), "hipMemcpy2D D2H")
        return out

    def copy_from(self, other: "Surface") -> None:
        assert (other.width, other.height, other.texel_bytes) == (self.width, self.height, self.texel_bytes)
        check(lib().hipMemcpy2D(ctypes.c_void_p(self.ptr), self.pitch, ctypes.c_void_p(other.ptr), other.pitch,
                                self.width * self.texel_bytes, self.height, hipMemcpyDeviceToDevice), "hipMemcpy2D D2D")


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(p)), "hipStreamCreate")
        self.handle = p.value

    def synchronize(self) -> None:
        check(lib().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle:
            lib().hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(p)), "hipEventCreate")
        self.handle = p.value

    def record(self, stream: int | None = None) -> None:
        check(lib().hipEventRecord(ctypes.c_void_p(self.handle), ctypes.c_void_p(stream or 0)), "hipEventRecord")

    def synchronize(self) -> None:
        check(lib().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle), ctypes.c_void_p(end.handle)),
              "hipEventElapsedTime")
        return ms.value


class Graph:
    """The launches that `record(stream_handle)` enqueues on `stream`, captured once (hipStreamBeginCapture,
    relaxed mode) and replayed by `launch()` as one hipGraphLaunch: the device runs them back to back, with
    no gap for the host to issue each one.  Only kernel launches belong in it (no memsets or copies)."""

    def __init__(self, stream: Stream, record):
        L = lib()
        s = ctypes.c_void_p(stream.handle)
        self.stream = stream
        self.graph, self.exec = ctypes.c_void_p(), ctypes.c_void_p()
        check(L.hipStreamBeginCapture(s, 2), "hipStreamBeginCapture")
        try:
            record(stream.handle)
        finally:
            check(L.hipStreamEndCapture(s, ctypes.byref(self.graph)), "hipStreamEndCapture")
        check(L.hipGraphInstantiate(ctypes.byref(self.exec), self.graph, None, None, 0), "hipGraphInstantiate")

    def launch(self) -> None:
        check(lib().hipGraphLaunch(self.exec, ctypes.c_void_p(self.stream.handle)), "hipGraphLaunch")

    def destroy(self) -> None:
        if self.exec:
            lib().hipGraphExecDestroy(self.exec)
            self.exec = ctypes.c_void_p()
        if self.graph:
            lib().hipGraphDestroy(self.graph)
            self.graph = ctypes.c_void_p()


def gpu_available() -> bool:
    if os.environ.get("HUSH_FORCE_NO_GPU"):
        return False
    try:
        return device_count() > 0
    except HipError:
        return False
