"""Mirror of `concrete_fft::ordered` (src/ordered.rs) over the device C ABI.

    plan = Plan(1024, Method.UserProvided(FftAlgo.Dif4))
    plan.fwd(buf)   # buf: numpy complex128 (host, synchronous) or DeviceBuffer
    plan.inv(buf)   # unnormalised, like the reference
"""
import ctypes
import enum

import numpy as np

from ._lib import CfftPanic, check, lib, ptr
from .device import DeviceBuffer


class FftAlgo(enum.IntEnum):
    """src/ordered.rs:26-45 (same discriminants)."""
    Dif2 = 0
    Dit2 = 1
    Dif4 = 2
    Dit4 = 3
    Dif8 = 4
    Dit8 = 5
    Dif16 = 6
    Dit16 = 7


class Method:
    """src/ordered.rs:48-59: UserProvided(algo) or Measure(seconds)."""

    USER_PROVIDED = 0
    MEASURE = 1

    def __init__(self, kind, algo=FftAlgo.Dif4, seconds=0.0):
        self.kind = kind
        self.algo = FftAlgo(algo)
        self.seconds = float(seconds)

    @classmethod
    def UserProvided(cls, algo):
        return cls(cls.USER_PROVIDED, algo)

    @classmethod
    def Measure(cls, seconds):
        return cls(cls.MEASURE, FftAlgo.Dif4, seconds)


class _PlanBase:
    handle = None

    def fft_size(self):
        n = ctypes.c_size_t(0)
        check(lib.cfft_plan_fft_size(self.handle, ctypes.byref(n)))
        return n.value

    def fft_scratch(self):
        """(bytes, alignment) of the scratch the reference asks its caller for."""
        b, a = ctypes.c_size_t(0), ctypes.c_size_t(0)
        check(lib.cfft_plan_fft_scratch(self.handle, ctypes.byref(b), ctypes.byref(a)))
        return b.value, a.value

    def _algo(self):
        a, b = ctypes.c_int(0), ctypes.c_size_t(0)
        check(lib.cfft_plan_algo(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return FftAlgo(a.value), b.value

    def _run(self, buf, fwd, stream=None, batch=None):
        n = self.fft_size()
        if isinstance(buf, DeviceBuffer):
            total = int(np.prod(buf.shape))
            if buf.dtype != np.complex128:
                raise TypeError("device buffer must be complex128")
            if batch is None:
                if total % n:
                    raise CfftPanic(6, "assertion `left == right` failed\n  left: %d\n right: %d"
                                    % (n, total))
                batch = total // n
            f = lib.cfft_fwd if fwd else lib.cfft_inv
            check(f(self.handle, buf.dptr, batch, n, stream.handle if stream else None))
            return buf
        if not isinstance(buf, np.ndarray) or buf.dtype != np.complex128 or not buf.flags.c_contiguous:
            raise TypeError("host buffer must be a C-contiguous numpy complex128 array")
        if buf.size % n or (buf.ndim == 1 and buf.size != n):
            raise CfftPanic(6, "assertion `left == right` failed\n  left: %d\n right: %d"
                            % (n, buf.shape[-1]))
        f = lib.cfft_fwd_host if fwd else lib.cfft_inv_host
        check(f(self.handle, buf.ctypes.data, buf.size // n))
        return buf

    def fwd(self, buf, stream=None, batch=None):
        """Forward transform in place (src/ordered.rs:342 / src/unordered.rs:826)."""
        return self._run(buf, True, stream, batch)

    def inv(self, buf, stream=None, batch=None):
        """Inverse transform in place, unnormalised (src/ordered.rs:368 / src/unordered.rs:927)."""
        return self._run(buf, False, stream, batch)

    def close(self):
        if self.handle:
            lib.cfft_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan(_PlanBase):
    """ordered::Plan (src/ordered.rs:186-373): standard order in and out,
    n <= 1024 (with large=True, n <= 65536)."""

    def __init__(self, n, method, large=False):
        """large=True admits 2^11 <= n <= 2^16 (cfft_ordered_plan_create_large):
        the reference's ordered recursion beyond its n < 2^11 assertion."""
        h = ctypes.c_void_p()
        create = lib.cfft_ordered_plan_create_large if large else lib.cfft_ordered_plan_create
        check(create(n, method.kind, int(method.algo), method.seconds, ctypes.byref(h)))
        self.handle = h.value

    def algo(self):
        return self._algo()[0]
