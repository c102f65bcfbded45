"""Mirror of `concrete_fft::fft128` (src/fft128/mod.rs) over the device C ABI."""
import ctypes

import numpy as np

from ._lib import check, lib
from .device import DeviceBuffer


class Plan:
    """fft128::Plan (src/fft128/mod.rs:1830-1961): negacyclic double-double FFT,
    n >= 32; fwd leaves bit-reversed order, inv expects it; unnormalised."""

    def __init__(self, n):
        h = ctypes.c_void_p()
        check(lib.cfft128_plan_create(n, ctypes.byref(h)))
        self.handle = h.value

    def fft_size(self):
        n = ctypes.c_size_t(0)
        check(lib.cfft_plan_fft_size(self.handle, ctypes.byref(n)))
        return n.value

    def _run(self, fwd, re0, re1, im0, im1, stream=None, batch=None):
        n = self.fft_size()
        bufs = (re0, re1, im0, im1)
        if all(isinstance(x, DeviceBuffer) for x in bufs):
            if batch is None:
                batch = int(np.prod(re0.shape)) // n
            f = lib.cfft128_fwd if fwd else lib.cfft128_inv
            check(f(self.handle, re0.dptr, re1.dptr, im0.dptr, im1.dptr, batch,
                    stream.handle if stream else None))
            return
        for x in bufs:
            if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
                raise TypeError("buffers must be C-contiguous float64 arrays")
            if x.size != re0.size or x.size % n:
                raise ValueError("buffer length mismatch")
        f = lib.cfft128_fwd_host if fwd else lib.cfft128_inv_host
        check(f(self.handle, re0.ctypes.data, re1.ctypes.data, im0.ctypes.data, im1.ctypes.data,
                re0.size // n))

    def fwd(self, re0, re1, im0, im1, stream=None, batch=None):
        self._run(True, re0, re1, im0, im1, stream, batch)

    def inv(self, re0, re1, im0, im1, stream=None, batch=None):
        self._run(False, re0, re1, im0, im1, stream, batch)

    def close(self):
        if self.handle:
            lib.cfft_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_negacyclic_twiddles(n):
    """Host twiddle tables (re0, re1, im0, im1) of init_negacyclic_twiddles."""
    t = [np.zeros(n, dtype=np.float64) for _ in range(4)]
    check(lib.cfft128_init_negacyclic_twiddles(n, *[x.ctypes.data for x in t]))
    return t


def negacyclic_fwd_fft(re0, re1, im0, im1, tw, batch=1, mul_variant=0, stream=None):
    """#[doc(hidden)] free function with caller twiddles (device buffers)."""
    n = int(np.prod(tw[0].shape))
    check(lib.cfft128_negacyclic_fwd_fft(n, re0.dptr, re1.dptr, im0.dptr, im1.dptr,
                                         tw[0].dptr, tw[1].dptr, tw[2].dptr, tw[3].dptr, batch,
                                         mul_variant, stream.handle if stream else None))


def negacyclic_inv_fft(re0, re1, im0, im1, tw, batch=1, mul_variant=0, stream=None):
    n = int(np.prod(tw[0].shape))
    check(lib.cfft128_negacyclic_inv_fft(n, re0.dptr, re1.dptr, im0.dptr, im1.dptr,
                                         tw[0].dptr, tw[1].dptr, tw[2].dptr, tw[3].dptr, batch,
                                         mul_variant, stream.handle if stream else None))
