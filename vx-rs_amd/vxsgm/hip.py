"""Minimal HIP runtime access (device buffers, streams, events) over the same
libamdhip64 that libvxsgm.so links, for callers that hand device pointers to
the stage-level, batched and stream-ordered entry points.

Using one runtime matters: a second HIP runtime loaded into the same process
(for example a framework's bundled copy) does not share devices, streams or
allocations with this library.
"""
import ctypes

import numpy as np

from . import LIB_PATH, lib as _vxlib  # noqa: F401  (loads libvxsgm first)

# The loader has already mapped libamdhip64.so.7 for libvxsgm.so; this handle
# resolves to that copy.
_hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)

_vp = ctypes.c_void_p
_H2D, _D2H, _D2D = 1, 2, 3


def _sig(name, res, *args):
    f = getattr(_hip, name)
    f.restype = res
    f.argtypes = list(args)


_sig("hipMalloc", ctypes.c_int, ctypes.POINTER(_vp), ctypes.c_size_t)
_sig("hipFree", ctypes.c_int, _vp)
_sig("hipMemcpy", ctypes.c_int, _vp, _vp, ctypes.c_size_t, ctypes.c_int)
_sig("hipMemset", ctypes.c_int, _vp, ctypes.c_int, ctypes.c_size_t)
_sig("hipDeviceSynchronize", ctypes.c_int)
_sig("hipStreamCreate", ctypes.c_int, ctypes.POINTER(_vp))
_sig("hipStreamDestroy", ctypes.c_int, _vp)
_sig("hipStreamSynchronize", ctypes.c_int, _vp)
_sig("hipEventCreate", ctypes.c_int, ctypes.POINTER(_vp))
_sig("hipEventDestroy", ctypes.c_int, _vp)
_sig("hipEventRecord", ctypes.c_int, _vp, _vp)
_sig("hipEventSynchronize", ctypes.c_int, _vp)
_sig("hipEventElapsedTime", ctypes.c_int, ctypes.POINTER(ctypes.c_float), _vp, _vp)
_sig("hipGetErrorString", ctypes.c_char_p, ctypes.c_int)
_sig("hipSetDevice", ctypes.c_int, ctypes.c_int)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {_hip.hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """A hipMalloc allocation, freed with the object."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        _ok(_hip.hipMalloc(ctypes.byref(p), max(1, self.nbytes)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _ok(_hip.hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ok(_hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(n, dtype)
        _ok(_hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def fill(self, byte):
        _ok(_hip.hipMemset(self.ptr, byte, self.nbytes), "hipMemset")

    def close(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = _vp()
        _ok(_hip.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(_hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle:
            _hip.hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = _vp()
        _ok(_hip.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        _ok(_hip.hipEventRecord(self.handle, stream.handle if stream else None),
            "hipEventRecord")

    def synchronize(self):
        _ok(_hip.hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _ok(_hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle),
            "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                _hip.hipEventDestroy(self.handle)
        except Exception:
            pass


def synchronize():
    _ok(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


def set_device(i):
    _ok(_hip.hipSetDevice(i), "hipSetDevice")
