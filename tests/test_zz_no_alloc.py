"""cfft_external_product_noalloc through the C ABI: the same bits as
cfft_external_product on two calls in a row (the second smaller), batches
that need several passes over the fixed 32 MiB scratch, no device memory
taken by the calls, and a GLWE too large for the scratch refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M = 2**64
K1, LEVELS, BASE_LOG = 2, 3, 8


def _setup(N, algo, base_n, seed):
    from cfft_amd import products
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(seed)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo(algo), base_n))
    key = rng.integers(0, M, (K1 * LEVELS, K1, N), dtype=np.uint64)
    dk = DeviceBuffer.from_host(key)
    kf = DeviceBuffer((K1 * LEVELS, K1, N // 2), np.complex128)
    products.ggsw_to_fourier(plan, dk, K1, LEVELS, kf)
    return rng, plan, kf


def _both(plan, kf, glwe, stream):
    """(noalloc result, cfft_external_product result) for the same GLWEs."""
    from cfft_amd import check, lib
    from cfft_amd.device import DeviceBuffer
    batch = glwe.shape[0]
    dg = DeviceBuffer.from_host(glwe)
    o1 = DeviceBuffer(glwe.shape, np.uint64)
    o2 = DeviceBuffer(glwe.shape, np.uint64)
    check(lib.cfft_external_product_noalloc(plan.handle, kf.dptr, K1, LEVELS, BASE_LOG, dg.dptr,
                                            o1.dptr, batch, stream.handle))
    stream.synchronize()
    r1 = o1.download()
    check(lib.cfft_external_product(plan.handle, kf.dptr, K1, LEVELS, BASE_LOG, dg.dptr, o2.dptr,
                                    batch, stream.handle))
    stream.synchronize()
    return r1, o2.download()


@pytest.mark.parametrize("N,algo,base_n,first,second", [
    (4096, 6, 1024, 400, 37),   # 170 GLWEs per pass: 3 passes, then 1
    (2048, 6, 1024, 700, 5),    # 341 per pass: 3 passes, then 1
    (1024, 6, 512, 9, 3),       # wave-local n = 512 kernels
    (2048, 2, 64, 50, 7),       # generic split kernels (Dif4 plan)
])
def test_noalloc_equals_allocating_entry_point(cfft, N, algo, base_n, first, second):
    from cfft_amd import lib
    from cfft_amd.device import Stream
    assert lib.cfft_noalloc_scratch_bytes() == 32 << 20
    rng, plan, kf = _setup(N, algo, base_n, N + first)
    s = Stream()
    for batch in (first, second):
        glwe = rng.integers(0, M, (batch, K1, N), dtype=np.uint64)
        got, want = _both(plan, kf, glwe, s)
        assert np.array_equal(got, want), (N, batch, int(np.sum(got != want)))


def test_noalloc_takes_no_device_memory(cfft):
    from cfft_amd import check, lib
    from cfft_amd.device import DeviceBuffer, Stream, mem_info
    rng, plan, kf = _setup(4096, 6, 1024, 5)
    s = Stream()
    glwe = rng.integers(0, M, (600, K1, 4096), dtype=np.uint64)
    dg = DeviceBuffer.from_host(glwe)
    do = DeviceBuffer(glwe.shape, np.uint64)
    # first call on this device resolves the scratch's address
    check(lib.cfft_external_product_noalloc(plan.handle, kf.dptr, K1, LEVELS, BASE_LOG, dg.dptr,
                                            do.dptr, 1, s.handle))
    s.synchronize()
    free0, _ = mem_info()
    for _ in range(3):
        check(lib.cfft_external_product_noalloc(plan.handle, kf.dptr, K1, LEVELS, BASE_LOG,
                                                dg.dptr, do.dptr, 600, s.handle))
    s.synchronize()
    free1, _ = mem_info()
    # a scratch-cache entry would take at least 32 MiB
    assert free1 >= free0 - (4 << 20), (free0, free1)


def test_noalloc_refuses_oversized_glwe(cfft):
    from cfft_amd import CfftPanic, check, lib
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.unordered import FftAlgo, Method, Plan
    plan = Plan(8192, Method.UserProvided(FftAlgo.Dif16, 1024))
    k1, levels = 64, 5  # 64 * 5 * 8192 * 16 bytes = 40 MiB of spectra per GLWE
    d = DeviceBuffer((16,), np.uint64)
    with pytest.raises(CfftPanic, match="scratch"):
        check(lib.cfft_external_product_noalloc(plan.handle, d.dptr, k1, levels, 8, d.dptr,
                                                d.dptr, 1, None))
