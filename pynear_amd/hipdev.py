"""Minimal device-memory, stream and event helpers over the HIP runtime that libpynear_amd.so uses.

The device-resident entry points (``*_device``) take device pointers and a stream. These helpers
allocate and move data with the very HIP runtime instance the library is linked against (the
``libamdhip64.so.7`` it loaded), so buffers, streams and events are valid for it; a second HIP
runtime in the process (for instance one bundled with another package) is never involved.
"""

import ctypes

import numpy as np

from . import _native

_H = None
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _H
    if _H is None:
        _native.lib()  # loads libpynear_amd.so and, with it, its HIP runtime
        H = ctypes.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [ctypes.POINTER(_vp), _sz]
        H.hipFree.argtypes = [_vp]
        H.hipMemcpy.argtypes = [_vp, _vp, _sz, ctypes.c_int]
        H.hipMemcpyAsync.argtypes = [_vp, _vp, _sz, ctypes.c_int, _vp]
        H.hipMemsetAsync.argtypes = [_vp, ctypes.c_int, _sz, _vp]
        H.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint]
        H.hipStreamDestroy.argtypes = [_vp]
        H.hipStreamSynchronize.argtypes = [_vp]
        H.hipDeviceSynchronize.argtypes = []
        H.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
        H.hipEventDestroy.argtypes = [_vp]
        H.hipEventRecord.argtypes = [_vp, _vp]
        H.hipEventSynchronize.argtypes = [_vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
        H.hipGetErrorString.restype = ctypes.c_char_p
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        _H = H
    return _H


def _ok(e):
    if e != 0:
        raise RuntimeError(f"HIP error {e}: {hip().hipGetErrorString(e).decode()}")


def synchronize():
    _ok(hip().hipDeviceSynchronize())


class DeviceArray:
    """A device allocation with numpy-style shape/dtype bookkeeping."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        _ok(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            _ok(hip().hipMemcpy(_vp(d.ptr), _vp(a.ctypes.data), a.nbytes, H2D))
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _ok(hip().hipMemcpy(_vp(out.ctypes.data), _vp(self.ptr), self.nbytes, D2H))
        return out

    def free(self):
        if self.ptr is not None and _H is not None:
            _H.hipFree(_vp(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()


class Stream:
    """A non-blocking HIP stream at default priority."""

    def __init__(self):
        s = _vp()
        _ok(hip().hipStreamCreateWithFlags(ctypes.byref(s), 1))
        self.handle = s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(_vp(self.handle)))

    def __del__(self):
        if getattr(self, "handle", None) is not None and _H is not None:
            _H.hipStreamDestroy(_vp(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        e = _vp()
        _ok(hip().hipEventCreate(ctypes.byref(e)))
        self.handle = e.value

    def record(self, stream=None):
        _ok(hip().hipEventRecord(_vp(self.handle), _vp(stream.handle if stream else None)))

    def synchronize(self):
        _ok(hip().hipEventSynchronize(_vp(self.handle)))

    def elapsed_ms(self, end):
        t = ctypes.c_float(0)
        _ok(hip().hipEventElapsedTime(ctypes.byref(t), _vp(self.handle), _vp(end.handle)))
        return float(t.value)

    def __del__(self):
        if getattr(self, "handle", None) is not None and _H is not None:
            _H.hipEventDestroy(_vp(self.handle))
            self.handle = None
