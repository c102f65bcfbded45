"""Device memory, streams and events through the library's own C ABI."""
import ctypes

import numpy as np

from ._lib import check, lib


def device_count():
    c = ctypes.c_int(0)
    st = lib.cfft_device_count(ctypes.byref(c))
    return c.value if st == 0 else 0


def synchronize():
    check(lib.cfft_device_synchronize())


def mem_info():
    f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib.cfft_device_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class Stream:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib.cfft_stream_create(ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(lib.cfft_stream_synchronize(self.handle))

    def close(self):
        if self.handle:
            lib.cfft_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib.cfft_event_create(ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        check(lib.cfft_event_record(self.handle, stream.handle if stream else None))

    def synchronize(self):
        check(lib.cfft_event_synchronize(self.handle))

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0)
        check(lib.cfft_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib.cfft_event_destroy(self.handle)
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation with a dtype and shape, for passing to the C ABI."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib.cfft_device_alloc(self.nbytes, ctypes.byref(p)))
        self.dptr = p.value or 0

    @classmethod
    def from_host(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.shape, arr.dtype)
        b.upload(arr, stream)
        return b

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        check(lib.cfft_memcpy_htod(self.dptr, arr.ctypes.data, self.nbytes,
                                   stream.handle if stream else None))
        if stream is None:
            check(lib.cfft_stream_synchronize(None))

    def download(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib.cfft_memcpy_dtoh(out.ctypes.data, self.dptr, self.nbytes,
                                   stream.handle if stream else None))
        return out

    def zero(self, stream=None):
        check(lib.cfft_memset(self.dptr, 0, self.nbytes, stream.handle if stream else None))

    def free(self):
        if self.dptr:
            lib.cfft_device_free(self.dptr)
            self.dptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
