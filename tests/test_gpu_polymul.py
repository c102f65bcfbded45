"""GPU parity of the fused negacyclic product: bit-identical to the oracle's
reference-based CPU product, and its torus error against the exact u64
schoolbook equals the CPU path's (it is the same computation)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_inputs(seed, batch, N):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2**64, size=(batch, N), dtype=np.uint64)
    b = rng.integers(-128, 128, size=(batch, N), dtype=np.int64)
    return a, b


def torus_err(c, exact):
    return np.abs((c - exact).view(np.int64)).max()


@pytest.mark.parametrize("N,algo,base_n", [(2048, 6, 1024), (2048, 2, 32), (2048, 7, 512),
                                           (1024, 4, 512), (4096, 3, 1024), (256, 0, 128),
                                           (64, 6, 32), (16384, 6, 1024), (2048, 4, 1024),
                                           (4096, 6, 1024), (8192, 6, 1024), (512, 6, 256),
                                           (2048, 7, 1024), (2048, 5, 1024), (2048, 6, 512),
                                           (2048, 0, 1024), (2048, 1, 1024), (2048, 2, 1024),
                                           (2048, 3, 1024), (2048, 0, 512), (2048, 1, 512),
                                           (2048, 2, 512), (2048, 3, 512), (2048, 4, 512),
                                           (2048, 5, 512), (1024, 0, 512), (1024, 1, 512),
                                           (1024, 2, 512), (1024, 3, 512), (1024, 5, 512),
                                           (1024, 7, 512), (4096, 0, 1024), (4096, 5, 1024),
                                           (4096, 7, 1024), (4096, 2, 512), (4096, 4, 512),
                                           (4096, 6, 512), (4096, 7, 512), (8192, 5, 1024),
                                           (8192, 2, 512), (8192, 7, 512), (16384, 5, 1024),
                                           (16384, 0, 512), (16384, 6, 512), (512, 0, 256),
                                           (512, 3, 256), (512, 5, 256), (512, 7, 256)])
def test_polymul_bit_exact_vs_oracle(cfft, oracle, N, algo, base_n):
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    batch = 33
    a, b = rand_inputs(N + algo, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo(algo), base_n))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    ref = oracle.polymul_f64(a, b, algo, base_n)
    assert np.array_equal(c, ref), (N, algo, base_n, int(np.sum(c != ref)))


def test_polymul_torus_error_equals_cpu(cfft, oracle):
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N, batch = 2048, 16
    a, b = rand_inputs(2, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    exact = oracle.schoolbook_u64(a, b.view(np.uint64))
    cpu = oracle.polymul_f64(a, b, 6, 1024)
    e_gpu, e_cpu = torus_err(c, exact), torus_err(cpu, exact)
    assert e_gpu <= e_cpu
    # f64 FFT product of a 64-bit torus polynomial: error well below 2^-24 of the torus
    assert e_gpu < 2**40


def test_polymul_exact_on_coarse_grid(cfft, oracle):
    # torus inputs on the 2^-20 grid (A = a * 2^44) and digits in [-4, 4): the
    # exact product lies on the same grid, and the f64 error (|C| <= 2^11 torus
    # units, relative error ~2^-50) is far below half a grid step, so rounding
    # the f64 product to the grid recovers the schoolbook result exactly
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N, batch = 1024, 8
    rng = np.random.default_rng(11)
    a = (rng.integers(-2**19, 2**19, size=(batch, N), dtype=np.int64) << 44).view(np.uint64)
    b = rng.integers(-4, 4, size=(batch, N), dtype=np.int64)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif8, 512))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    exact = oracle.schoolbook_u64(a, b.view(np.uint64))
    rounded = ((c + np.uint64(1 << 43)) >> np.uint64(44)) << np.uint64(44)
    assert np.array_equal(rounded, exact)
    assert torus_err(c, exact) < 2**40


def test_polymul_monomial_shift(cfft, oracle):
    # B = X^k: the product is the negacyclic shift of A; A lies on the 2^-31
    # grid and the f64 error (~2^13 units of 2^-64) is far below half a step
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 2048
    rng = np.random.default_rng(12)
    a = (rng.integers(-2**30, 2**30, size=(1, N), dtype=np.int64) << 33).view(np.uint64)
    for k in (0, 1, 5, N - 1):
        b = np.zeros((1, N), dtype=np.int64)
        b[0, k] = 1
        plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
        c = products.negacyclic_polymul_u64_host(plan, a, b)
        shifted = np.concatenate([(-a[0, N - k:].view(np.int64)).view(np.uint64), a[0, :N - k]])
        rounded = ((c[0] + np.uint64(1 << 32)) >> np.uint64(33)) << np.uint64(33)
        assert np.array_equal(rounded, shifted), k


def test_polymul_small_outputs_bit_exact(cfft, oracle):
    # products below 2^-11 in magnitude take the rounding branch of the torus
    # conversion (their ulp is finer than 2^-64); an all-zero row gives exact 0
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N, batch = 2048, 8
    rng = np.random.default_rng(21)
    a = rng.integers(-2**40, 2**40, size=(batch, N), dtype=np.int64).view(np.uint64)
    b = rng.integers(-3, 4, size=(batch, N), dtype=np.int64)
    a[0] = 0
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    ref = oracle.polymul_f64(a, b, 6, 1024)
    assert np.array_equal(c, ref), int(np.sum(c != ref))
    assert not c[0].any()


@pytest.mark.parametrize("batch", [1, 2, 15, 16, 17, 1001, 4097, 8195])
def test_half_wave_products_n1024(cfft, oracle, batch):
    # N = 1024 on a Dif16 plan with base 512: one product per 32-lane
    # half-wave, 16 per workgroup, at most one workgroup per CU (4096
    # products per round on 256 CUs); odd batches leave the last half idle
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 1024
    a, b = rand_inputs(batch + 5, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 512))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    assert np.array_equal(c, oracle.polymul_f64(a, b, 6, 512, threads=16))


@pytest.mark.parametrize("batch", [1, 3, 33, 8193, 16390])
def test_quarter_wave_products_n512(cfft, oracle, batch):
    # N = 512 on a Dif16 plan with base 256: one product per 16-lane
    # quarter-wave, 32 per workgroup, at most one workgroup per CU (8192
    # products per round on 256 CUs); the last wave may be partly idle
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 512
    a, b = rand_inputs(batch + 9, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 256))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    assert np.array_equal(c, oracle.polymul_f64(a, b, 6, 256, threads=16))


@pytest.mark.parametrize("batch", [1, 7, 8, 9, 2049])
def test_persistent_grid_boundaries(cfft, oracle, batch):
    # N = 2048 Dif16 / 1024 runs the persistent kernel: 8 products per
    # workgroup, at most one workgroup per CU, products strided over waves
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 2048
    a, b = rand_inputs(batch, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    assert np.array_equal(c, oracle.polymul_f64(a, b, 6, 1024))


@pytest.mark.parametrize("batch", [1, 8, 9, 2047, 2048, 2049, 8192, 8193])
def test_persistent_product_batches(cfft, oracle, batch):
    """N = 2048 on the Dif16 / 1024 plan (the persistent kernel: 8 waves per
    workgroup, one workgroup per CU, products strided over waves, the next
    product's B loaded during the inverse): bit-exact against the oracle at
    batch sizes around its wave and round boundaries, waves without a
    product included."""
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 2048
    a, b = rand_inputs(1000 + batch, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    ref = oracle.polymul_f64(a, b, 6, 1024, threads=16)
    assert np.array_equal(c, ref), (batch, int(np.sum(c != ref)))


@pytest.mark.parametrize("batch", [1, 7, 9, 1031, 2049, 4100])
def test_base512_product_batches(cfft, oracle, batch):
    """N = 2048 on the Dif16 / 512 plan (the persistent kernel: one product
    per wave at a time, at most one workgroup per CU, the depth pass's
    partner values swapped between the half-waves, the next product's B
    loaded during the inverse): bit-exact against the oracle with partial
    workgroups and for batches past a full grid of 256 workgroups."""
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 2048
    a, b = rand_inputs(500 + batch, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 512))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    ref = oracle.polymul_f64(a, b, 6, 512, threads=16)
    assert np.array_equal(c, ref), (batch, int(np.sum(c != ref)))


@pytest.mark.parametrize("batch", [1, 9, 2049, 4100])
def test_dif8_product_batches(cfft, oracle, batch):
    """N = 2048 on the Dif8 / 1024 plan (one product per 128-thread
    workgroup, the register-resident Dif8 schedule): bit-exact against the
    oracle from one product to a few per CU."""
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    N = 2048
    a, b = rand_inputs(700 + batch, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif8, 1024))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    ref = oracle.polymul_f64(a, b, 4, 1024, threads=16)
    assert np.array_equal(c, ref), (batch, int(np.sum(c != ref)))


@pytest.mark.parametrize("N,batch", [(4096, 1), (4096, 1030), (8192, 3), (8192, 515),
                                     (16384, 2), (16384, 259)])
def test_depth_product_rounds(cfft, oracle, N, batch):
    """N = 4096, 8192, 16384 on Dif16 / 1024 plans (the persistent depth-pass
    product: 8 / R products per workgroup and round, at most one workgroup
    per CU, operands of the next step loaded during the current one):
    bit-exact against the oracle for one round, and for batches one round
    past a full grid of 256 workgroups."""
    from cfft_amd import products
    from cfft_amd.unordered import FftAlgo, Method, Plan
    a, b = rand_inputs(N + batch, batch, N)
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    c = products.negacyclic_polymul_u64_host(plan, a, b)
    ref = oracle.polymul_f64(a, b, 6, 1024, threads=16)
    assert np.array_equal(c, ref), (N, batch, int(np.sum(c != ref)))
