"""The host-buffer entry points (cfft_fwd_host, cfft_inv_host,
cfft_fwd_monomial_host, cfft128_*_host) against the device-buffer ones, bit
for bit.  Calls of at most 64 KiB run their kernels on the pinned staging
buffer itself, larger ones copy through a device buffer, and batches above
64 MiB run in chunks; the cases sit on both sides of each limit."""
import numpy as np
import pytest

from cfft_amd import DeviceBuffer, fft128, ordered, unordered
from cfft_amd.ordered import FftAlgo


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _both(plan, x, fwd):
    h = x.copy()
    (plan.fwd if fwd else plan.inv)(h)
    d = DeviceBuffer.from_host(x)
    (plan.fwd if fwd else plan.inv)(d)
    return h, d.download()


@pytest.mark.gpu
@pytest.mark.parametrize("n,base_n,batch", [
    (1024, 1024, 1),     # 16 KiB: on the pinned buffer
    (1024, 1024, 4),     # 64 KiB: the largest call on the pinned buffer
    (1024, 1024, 5),     # through a device buffer
    (4096, 512, 1),      # two-level plan on the pinned buffer
    (65536, 1024, 1),    # 1 MiB through a device buffer
    (1024, 1024, 4100),  # 64 MiB + 64 KiB: two chunks
])
def test_unordered_host_equals_device(cfft, n, base_n, batch):
    print("host path: unordered n=%d base=%d batch=%d" % (n, base_n, batch), flush=True)
    plan = unordered.Plan(n, unordered.Method.UserProvided(FftAlgo.Dif16, base_n))
    rng = np.random.default_rng(n + batch)
    x = _rand(rng, batch, n) if batch > 1 else _rand(rng, n)
    for fwd in (True, False):
        h, d = _both(plan, x, fwd)
        assert np.array_equal(h.view(np.uint64), d.view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("n,batch", [(256, 1), (1024, 8), (16384, 1)])
def test_ordered_host_equals_device(cfft, n, batch):
    print("host path: ordered n=%d batch=%d" % (n, batch), flush=True)
    plan = ordered.Plan(n, ordered.Method.UserProvided(FftAlgo.Dit4), large=True)
    rng = np.random.default_rng(n)
    x = _rand(rng, batch, n) if batch > 1 else _rand(rng, n)
    for fwd in (True, False):
        h, d = _both(plan, x, fwd)
        assert np.array_equal(h.view(np.uint64), d.view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1024, 8192])
def test_fwd_monomial_host_equals_device(cfft, n):
    print("host path: fwd_monomial n=%d" % n, flush=True)
    plan = unordered.Plan(n, unordered.Method.UserProvided(FftAlgo.Dif4, min(n, 1024)))
    for degree in (0, 1, n // 3, n - 1):
        m = np.full(n, np.nan, dtype=np.complex128)  # every point is written
        plan.fwd_monomial(degree, m)
        d = DeviceBuffer.from_host(np.zeros(n, dtype=np.complex128))
        plan.fwd_monomial(degree, d)
        assert np.array_equal(m.view(np.uint64), d.download().view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("n,batch", [(1024, 1), (1024, 3), (4096, 2)])
def test_fft128_host_equals_device(cfft, n, batch):
    print("host path: fft128 n=%d batch=%d" % (n, batch), flush=True)
    plan = fft128.Plan(n)
    rng = np.random.default_rng(n * batch)
    shape = (batch, n) if batch > 1 else (n,)
    hi = [rng.standard_normal(shape) for _ in range(2)]
    lo = [h * 2.0 ** -60 * rng.standard_normal(shape) for h in hi]
    arrs = [hi[0], lo[0], hi[1], lo[1]]
    for fwd in (True, False):
        h = [a.copy() for a in arrs]
        (plan.fwd if fwd else plan.inv)(*h)
        d = [DeviceBuffer.from_host(a.copy()) for a in arrs]
        (plan.fwd if fwd else plan.inv)(*d)
        for a, b in zip(h, d):
            assert np.array_equal(a.view(np.uint64), b.download().view(np.uint64))
