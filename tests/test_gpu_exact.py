"""Exact u64 negacyclic product (16-bit limbs, f64 transforms, integer
rounding and carries): bit-exact against the u64 schoolbook oracle, for any
plan, with the reported rounding margin below 1/2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M = 2**64


@pytest.mark.parametrize("N,algo,base_n", [(4, 6, 2), (64, 6, 32), (256, 2, 128),
                                           (2048, 6, 1024), (2048, 3, 32), (4096, 6, 1024),
                                           (4096, 0, 512)])
def test_exact_bit_exact_random(cfft, oracle, N, algo, base_n):
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(N + algo)
    batch = 6
    a = rng.integers(0, M, (batch, N), dtype=np.uint64)
    b = rng.integers(0, M, (batch, N), dtype=np.uint64)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo(algo), base_n))
    c, margin = products.exact_polymul_u64_host(plan, a, b, with_margin=True)
    exact = oracle.schoolbook_u64(a, b)
    assert np.array_equal(c, exact), (N, int(np.sum(c != exact)))
    assert margin.max() < 0.5 and margin.min() >= 0.0
    print("N=%d max rounding margin %.3g" % (N, margin.max()))


@pytest.mark.parametrize("N", [2048, 4096])
@pytest.mark.parametrize("batch", [1, 3, 5])
def test_exact_weight_spectra_odd_batches(cfft, oracle, N, batch):
    # Dif16 plans with base 1024 form the four weight spectra inside the
    # forward kernel (two products per workgroup at N = 2048, so an odd
    # batch leaves half of the last workgroup without a product)
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(1000 * batch + N)
    a = rng.integers(0, M, (batch, N), dtype=np.uint64)
    b = rng.integers(0, M, (batch, N), dtype=np.uint64)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    c, margin = products.exact_polymul_u64_host(plan, a, b, with_margin=True)
    assert np.array_equal(c, oracle.schoolbook_u64(a, b))
    assert margin.max() < 0.5 and margin.min() >= 0.0


def test_exact_edge_values(cfft, oracle):
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 2048
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    rows = []
    ones = np.full(N, M - 1, dtype=np.uint64)        # all 0xffff...ffff
    zeros = np.zeros(N, dtype=np.uint64)
    top = np.full(N, 1 << 63, dtype=np.uint64)
    alt = np.where(np.arange(N) % 2 == 0, np.uint64(0x8000800080008000), np.uint64(0x7fff7fff7fff7fff))
    mono = zeros.copy()
    mono[N - 1] = 1
    for x, y in [(ones, ones), (zeros, ones), (top, top), (alt, ones), (alt, alt), (mono, alt),
                 (ones, mono)]:
        rows.append((x, y))
    a = np.stack([r[0] for r in rows])
    b = np.stack([r[1] for r in rows])
    c = products.exact_polymul_u64_host(plan, a, b)
    assert np.array_equal(c, oracle.schoolbook_u64(a, b))


def test_exact_independent_of_plan(cfft):
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 1024
    rng = np.random.default_rng(5)
    a = rng.integers(0, M, (3, N), dtype=np.uint64)
    b = rng.integers(0, M, (3, N), dtype=np.uint64)
    outs = [products.exact_polymul_u64_host(Plan(N // 2, Method.UserProvided(FftAlgo(al), bn)), a, b)
            for al, bn in [(6, 512), (1, 32), (4, 128)]]
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("N", [64, 256, 2048, 4096, 8192])
def test_exact128_bit_exact(cfft, oracle, N):
    # the double-double path: 32-bit limbs over an fft128 plan of size N/2
    from cfft_amd import fft128, products
    rng = np.random.default_rng(N + 128)
    batch = 5
    a = rng.integers(0, M, (batch, N), dtype=np.uint64)
    b = rng.integers(0, M, (batch, N), dtype=np.uint64)
    a[0, :3] = [M - 1, 1 << 63, (1 << 32) - 1]
    b[0, :3] = [M - 1, (1 << 31), (1 << 63) + (1 << 31)]
    plan = fft128.Plan(N // 2)
    c, margin = products.exact128_polymul_u64_host(plan, a, b, with_margin=True)
    assert np.array_equal(c, oracle.schoolbook_u64(a, b))
    assert margin.max() < 0.5 and margin.min() >= 0.0
    print("N=%d fft128 path max rounding margin %.3g" % (N, margin.max()))


def test_exact_paths_agree_on_edge_values(cfft):
    from cfft_amd import fft128, products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 4096
    a = np.stack([np.full(N, M - 1, np.uint64), np.full(N, 1 << 63, np.uint64),
                  np.full(N, 0x80000000, np.uint64)])
    b = np.stack([np.full(N, M - 1, np.uint64), np.full(N, 1 << 63, np.uint64),
                  np.full(N, 0x7fffffff80000000, np.uint64)])
    c64 = products.exact_polymul_u64_host(Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024)),
                                          a, b)
    c128 = products.exact128_polymul_u64_host(fft128.Plan(N // 2), a, b)
    assert np.array_equal(c64, c128)


def test_exact_calls_on_two_streams_share_scratch(cfft, oracle):
    # consecutive calls of different sizes alternate between two streams with
    # no synchronisation between them; the products' scratch buffers are
    # reused across calls and streams, and every result must still be exact
    from cfft_amd import fft128, products
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 2048
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    plan128 = fft128.Plan(N // 2)
    streams = [Stream(), Stream()]
    rng = np.random.default_rng(77)
    calls = []
    for i, batch in enumerate([5, 2, 7, 1, 6, 3, 4, 2]):
        s = streams[i % 2]
        a = rng.integers(0, M, (batch, N), dtype=np.uint64)
        b = rng.integers(0, M, (batch, N), dtype=np.uint64)
        da, db = DeviceBuffer.from_host(a, s), DeviceBuffer.from_host(b, s)
        dc = DeviceBuffer((batch, N), np.uint64)
        if i % 3 == 2:
            products.exact128_polymul_u64(plan128, da, db, dc, batch, stream=s)
        else:
            products.exact_polymul_u64(plan, da, db, dc, batch, stream=s)
        calls.append((a, b, da, db, dc))
    for s in streams:
        s.synchronize()
    for a, b, _, _, dc in calls:
        assert np.array_equal(dc.download(), oracle.schoolbook_u64(a, b))


@pytest.mark.parametrize("N,base_n", [(4096, 1024), (2048, 1024), (1024, 512)])
def test_exact_worst_case_limbs(cfft, oracle, N, base_n):
    # every coefficient 0x8000800080008000: all four balanced limbs are
    # -2^15, so |D_3| reaches 4 N 2^30 and the rounding margin is tightest
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    v = 0x8000800080008000
    a = np.full((2, N), v, np.uint64)
    b = np.full((2, N), v, np.uint64)
    b[1, ::3] = M - v  # limbs of opposite sign on every third coefficient
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, base_n))
    c, margin = products.exact_polymul_u64_host(plan, a, b, with_margin=True)
    assert np.array_equal(c, oracle.schoolbook_u64(a, b))
    # exact while < 1/2; measured 0.0098 at N = 4096, so wide headroom remains
    assert 0.0 <= margin.max() < 2.0 ** -5, margin
