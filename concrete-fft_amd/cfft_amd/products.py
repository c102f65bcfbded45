"""Fused negacyclic products on an unordered plan (caller-side in the
reference: TFHE-rs builds them on `unordered::Plan`)."""
import numpy as np

from ._lib import check, lib
from .device import DeviceBuffer


def _h(stream):
    return stream.handle if stream else None


def negacyclic_polymul_u64(plan, a, b, c, batch, stream=None):
    """C = A * B mod (X^N + 1) on the 2^-64 torus; N = 2 * plan.fft_size().
    a: u64 torus coefficients, b: i64 integer coefficients, c: u64 output;
    all DeviceBuffers of batch * N words."""
    check(lib.cfft_negacyclic_polymul_u64(plan.handle, a.dptr, b.dptr, c.dptr, batch, _h(stream)))


def negacyclic_polymul_u64_host(plan, a, b):
    """Host convenience: a (batch, N) uint64, b (batch, N) int64 -> (batch, N) uint64."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.int64)
    batch = a.shape[0] if a.ndim == 2 else 1
    da, db = DeviceBuffer.from_host(a), DeviceBuffer.from_host(b)
    dc = DeviceBuffer(a.shape, np.uint64)
    negacyclic_polymul_u64(plan, da, db, dc, batch)
    return dc.download()


def exact_polymul_u64(plan, a, b, c, batch, margin=None, stream=None):
    """Exact C = A * B mod (X^N + 1, 2^64) via 16-bit limbs."""
    check(lib.cfft_exact_polymul_u64(plan.handle, a.dptr, b.dptr, c.dptr,
                                     margin.dptr if margin is not None else None, batch,
                                     _h(stream)))


def exact_polymul_u64_host(plan, a, b, with_margin=False):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    batch = a.shape[0] if a.ndim == 2 else 1
    da, db = DeviceBuffer.from_host(a), DeviceBuffer.from_host(b)
    dc = DeviceBuffer(a.shape, np.uint64)
    dm = DeviceBuffer((batch,), np.float64) if with_margin else None
    exact_polymul_u64(plan, da, db, dc, batch, dm)
    out = dc.download()
    if with_margin:
        return out, dm.download()
    return out


def exact128_polymul_u64(plan128, a, b, c, batch, margin=None, stream=None):
    """Exact C = A * B mod (X^N + 1, 2^64) over the double-double transform of
    an fft128 plan of size N/2 (32-bit limbs)."""
    check(lib.cfft128_exact_polymul_u64(plan128.handle, a.dptr, b.dptr, c.dptr,
                                        margin.dptr if margin is not None else None, batch,
                                        _h(stream)))


def exact128_polymul_u64_host(plan128, a, b, with_margin=False):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    batch = a.shape[0] if a.ndim == 2 else 1
    da, db = DeviceBuffer.from_host(a), DeviceBuffer.from_host(b)
    dc = DeviceBuffer(a.shape, np.uint64)
    dm = DeviceBuffer((batch,), np.float64) if with_margin else None
    exact128_polymul_u64(plan128, da, db, dc, batch, dm)
    out = dc.download()
    if with_margin:
        return out, dm.download()
    return out


def ggsw_to_fourier(plan, key, k1, levels, key_fourier, stream=None):
    check(lib.cfft_ggsw_to_fourier(plan.handle, key.dptr, k1, levels, key_fourier.dptr,
                                   _h(stream)))


def external_product(plan, key_fourier, k1, levels, base_log, glwe, out, batch, stream=None):
    check(lib.cfft_external_product(plan.handle, key_fourier.dptr, k1, levels, base_log,
                                    glwe.dptr, out.dptr, batch, _h(stream)))


def external_product_noalloc(plan, key_fourier, k1, levels, base_log, glwe, out, batch,
                             stream=None):
    """external_product over the library's fixed per-device scratch: no device
    memory is allocated (calls overlapping on one device race)."""
    check(lib.cfft_external_product_noalloc(plan.handle, key_fourier.dptr, k1, levels, base_log,
                                            glwe.dptr, out.dptr, batch, _h(stream)))
