"""Mirror of `concrete_fft::unordered` (src/unordered.rs) over the device C ABI.

    plan = Plan(2048, Method.UserProvided(FftAlgo.Dif4, 32))
    plan.fwd(z)                       # z in place, permuted Fourier order
    plan.inv(z)                       # back to standard order, unnormalised
    plan.fwd_monomial(degree, z)
    data = plan.serialize_fourier_buffer(z)      # bincode byte stream
    plan2.deserialize_fourier_buffer(data, z)
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .device import DeviceBuffer
from .ordered import FftAlgo, _PlanBase


class Method:
    """src/unordered.rs:526-537: UserProvided{base_algo, base_n} or Measure(seconds)."""

    USER_PROVIDED = 0
    MEASURE = 1

    def __init__(self, kind, base_algo=FftAlgo.Dif4, base_n=0, seconds=0.0):
        self.kind = kind
        self.base_algo = FftAlgo(base_algo)
        self.base_n = int(base_n)
        self.seconds = float(seconds)

    @classmethod
    def UserProvided(cls, base_algo, base_n):
        return cls(cls.USER_PROVIDED, base_algo, base_n)

    @classmethod
    def Measure(cls, seconds):
        return cls(cls.MEASURE, FftAlgo.Dif4, 0, seconds)


class Plan(_PlanBase):
    """unordered::Plan (src/unordered.rs:496-1036)."""

    def __init__(self, n, method):
        h = ctypes.c_void_p()
        check(lib.cfft_unordered_plan_create(n, method.kind, int(method.base_algo), method.base_n,
                                             method.seconds, ctypes.byref(h)))
        self.handle = h.value

    def algo(self):
        """(base_algo, base_n), src/unordered.rs:783."""
        return self._algo()

    def permutation(self):
        """idx[i] = position of frequency i (bit_rev_twice)."""
        n = self.fft_size()
        idx = np.zeros(n, dtype=np.uintp)
        check(lib.cfft_plan_permutation(self.handle, idx.ctypes.data))
        return idx.astype(np.int64)

    def fwd_monomial(self, degree, buf, stream=None):
        """src/unordered.rs:844-900."""
        if isinstance(buf, DeviceBuffer):
            check(lib.cfft_fwd_monomial(self.handle, degree, buf.dptr,
                                        stream.handle if stream else None))
            return buf
        n = self.fft_size()
        if buf.dtype != np.complex128 or buf.size != n:
            raise ValueError("buffer must be complex128 of length %d" % n)
        check(lib.cfft_fwd_monomial_host(self.handle, degree, buf.ctypes.data))
        return buf

    def serialize_fourier_buffer(self, buf):
        """bincode(options()) stream of the buffer in standard order (src/unordered.rs:951)."""
        buf = np.ascontiguousarray(buf, dtype=np.complex128)
        ln = ctypes.c_size_t(0)
        check(lib.cfft_serialize_fourier_buffer(self.handle, buf.ctypes.data, None, 0,
                                                ctypes.byref(ln)))
        out = np.zeros(ln.value, dtype=np.uint8)
        check(lib.cfft_serialize_fourier_buffer(self.handle, buf.ctypes.data, out.ctypes.data,
                                                out.size, ctypes.byref(ln)))
        return out.tobytes()

    def deserialize_fourier_buffer(self, data, buf):
        """src/unordered.rs:988-1036: scatter a standard-order stream into permuted order."""
        raw = np.frombuffer(data, dtype=np.uint8)
        check(lib.cfft_deserialize_fourier_buffer(self.handle, raw.ctypes.data, raw.size,
                                                  buf.ctypes.data))
        return buf

    def fourier_to_standard(self, src, dst, batch=1, stream=None):
        check(lib.cfft_fourier_to_standard(self.handle, src.dptr, dst.dptr, batch,
                                           stream.handle if stream else None))

    def fourier_from_standard(self, src, dst, batch=1, stream=None):
        check(lib.cfft_fourier_from_standard(self.handle, src.dptr, dst.dptr, batch,
                                             stream.handle if stream else None))
