"""External product (GGSW (x) GLWE on a resident Fourier-domain key) on the
GPU: key transform and product bit-exact against the oracle for the same
plan; the trivial GGSW of 1 recomposes the rounded GLWE at the serving sizes;
argument errors."""
import numpy as np
import pytest

from test_oracle_products import gadget_key

pytestmark = pytest.mark.gpu

M = 2**64


def _plan(N, algo, base_n):
    from cfft_amd.unordered import FftAlgo, Method, Plan
    return Plan(N // 2, Method.UserProvided(FftAlgo(algo), base_n))


def _key_fourier(plan, key, k1, levels):
    from cfft_amd import products
    from cfft_amd.device import DeviceBuffer
    dk = DeviceBuffer.from_host(key)
    kf = DeviceBuffer(key.shape[:-1] + (key.shape[-1] // 2,), np.complex128)
    products.ggsw_to_fourier(plan, dk, k1, levels, kf)
    return kf


def _extprod(plan, kf, k1, levels, base_log, glwe):
    from cfft_amd import products
    from cfft_amd.device import DeviceBuffer
    dg = DeviceBuffer.from_host(glwe)
    do = DeviceBuffer(glwe.shape, np.uint64)
    products.external_product(plan, kf, k1, levels, base_log, dg, do, glwe.shape[0])
    return do.download()


@pytest.mark.parametrize("N,algo,base_n", [(2048, 6, 1024), (1024, 6, 512), (4096, 6, 1024),
                                           (64, 2, 32), (2048, 1, 128)])
def test_ggsw_to_fourier_bit_exact(cfft, oracle, N, algo, base_n):
    k1, levels = 2, 3
    rng = np.random.default_rng(N + algo)
    key = rng.integers(0, M, (k1 * levels, k1, N), dtype=np.uint64)
    plan = _plan(N, algo, base_n)
    got = _key_fourier(plan, key, k1, levels).download()
    want = oracle.torus_to_fourier(key, algo, base_n)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("N,k1,levels,base_log,algo,base_n,batch", [
    (1024, 2, 3, 8, 6, 512, 5),
    (2048, 2, 3, 8, 6, 1024, 4),
    (4096, 2, 3, 8, 6, 1024, 3),
    (2048, 1, 2, 10, 2, 64, 3),
    (512, 3, 4, 6, 7, 256, 7),
    (64, 2, 1, 16, 6, 32, 9),
    (4, 2, 3, 8, 0, 2, 3),
    (4096, 2, 3, 8, 2, 1024, 40),  # generic split kernels (Dif4 plan), 240 rows
])
def test_external_product_bit_exact(cfft, oracle, N, k1, levels, base_log, algo, base_n, batch):
    rng = np.random.default_rng(N * 7 + k1 + levels)
    key = rng.integers(0, M, (k1 * levels, k1, N), dtype=np.uint64)
    glwe = rng.integers(0, M, (batch, k1, N), dtype=np.uint64)
    glwe[0, 0, :4] = [0, M - 1, 1 << 63, (1 << 63) - 1]
    plan = _plan(N, algo, base_n)
    kf = _key_fourier(plan, key, k1, levels)
    got = _extprod(plan, kf, k1, levels, base_log, glwe)
    kfh = oracle.torus_to_fourier(key, algo, base_n)
    for b in range(batch):
        want = oracle.external_product(N, k1, levels, base_log, kfh, algo, base_n, glwe[b])
        assert np.array_equal(got[b], want), (b, int(np.sum(got[b] != want)))


@pytest.mark.parametrize("N", [1024, 2048, 4096])
def test_trivial_ggsw_recomposes_at_serving_sizes(cfft, N):
    k1, levels, base_log = 2, 3, 8
    rng = np.random.default_rng(N)
    glwe = rng.integers(0, M, (64, k1, N), dtype=np.uint64)
    plan = _plan(N, 6, min(N // 2, 1024))
    kf = _key_fourier(plan, gadget_key(N, k1, levels, base_log), k1, levels)
    out = _extprod(plan, kf, k1, levels, base_log, glwe)
    rep = levels * base_log
    rounded = ((glwe + np.uint64(1 << (63 - rep))) >> np.uint64(64 - rep)) << np.uint64(64 - rep)
    err = np.abs((out - rounded).view(np.int64)).max()
    assert err < 2**16, err


def test_external_product_errors(cfft):
    from cfft_amd import CfftPanic, products
    from cfft_amd.device import DeviceBuffer
    plan = _plan(4096, 6, 1024)
    kf = DeviceBuffer((16,), np.complex128)
    g = DeviceBuffer((16,), np.uint64)
    o = DeviceBuffer((16,), np.uint64)
    with pytest.raises(CfftPanic, match="base_log"):
        products.external_product(plan, kf, 2, 3, 0, g, o, 1)
    with pytest.raises(CfftPanic, match="base_log"):
        products.external_product(plan, kf, 2, 9, 8, g, o, 1)
    with pytest.raises(CfftPanic, match="k1"):
        products.external_product(plan, kf, 65, 1, 8, g, o, 1)
    with pytest.raises(CfftPanic, match="k1"):
        products.external_product(plan, kf, 0, 3, 8, g, o, 1)
    from cfft_amd.ordered import FftAlgo, Method, Plan
    with pytest.raises(CfftPanic, match="unordered"):
        products.external_product(Plan(1024, Method.UserProvided(FftAlgo.Dif4)), kf, 2, 3, 8, g,
                                  o, 1)


@pytest.mark.parametrize("N", [1024, 2048, 4096])
def test_batch_splits_agree(cfft, oracle, N):
    # one call over 200 GLWEs equals two calls over 100 each (different
    # launch shapes of the split row / column kernels) and the oracle
    k1, levels, base_log = 2, 3, 8
    rng = np.random.default_rng(N + 99)
    key = rng.integers(0, M, (k1 * levels, k1, N), dtype=np.uint64)
    glwe = rng.integers(0, M, (200, k1, N), dtype=np.uint64)
    plan = _plan(N, 6, min(N // 2, 1024))
    kf = _key_fourier(plan, key, k1, levels)
    whole = _extprod(plan, kf, k1, levels, base_log, glwe)
    parts = np.concatenate([_extprod(plan, kf, k1, levels, base_log, glwe[i:i + 100])
                            for i in (0, 100)])
    assert np.array_equal(whole, parts)
    kfh = oracle.torus_to_fourier(key, 6, min(N // 2, 1024))
    for b in (0, 199):
        assert np.array_equal(whole[b], oracle.external_product(N, k1, levels, base_log, kfh, 6,
                                                                min(N // 2, 1024), glwe[b]))


def test_large_batch_runs_in_passes(cfft, oracle):
    # k1 * levels = 16 rows per GLWE: 16384 rows make one pass of 1024 GLWEs,
    # so 1100 GLWEs run in two passes over the same scratch
    N, k1, levels, base_log = 1024, 2, 8, 4
    rng = np.random.default_rng(321)
    key = rng.integers(0, M, (k1 * levels, k1, N), dtype=np.uint64)
    glwe = rng.integers(0, M, (1100, k1, N), dtype=np.uint64)
    plan = _plan(N, 6, 512)
    kf = _key_fourier(plan, key, k1, levels)
    whole = _extprod(plan, kf, k1, levels, base_log, glwe)
    parts = np.concatenate([_extprod(plan, kf, k1, levels, base_log, glwe[i:j])
                            for i, j in ((0, 500), (500, 1100))])
    assert np.array_equal(whole, parts)
    kfh = oracle.torus_to_fourier(key, 6, 512)
    for b in (0, 1023, 1024, 1099):
        assert np.array_equal(whole[b], oracle.external_product(N, k1, levels, base_log, kfh, 6,
                                                                512, glwe[b])), b
