"""The SIMD restatements of the reference's x86 path (V3: AVX2 + FMA, c64x2;
V4: AVX-512F, c64x4) that the CPU baseline times.  The reference's f64 output
is the same on every ISA (its golden vector is asserted on all of them,
src/unordered.rs:1201-1395), so each series must reproduce that vector bit for
bit and agree bit for bit with the scalar restatement on every DIF base
algorithm, base size and depth, forward and inverse, and on the fused
product.  CPU only; a series the host CPU lacks is skipped."""
import os

import numpy as np
import pytest

from oracle_lib import load_oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "unordered_n2048_dif4_b32.npy")
DIF = [0, 2, 4, 6]  # Dif2, Dif4, Dif8, Dif16


@pytest.fixture(scope="module")
def O():
    return load_oracle()


def _need(O, isa):
    if not O.L.or_cpu_has_isa(isa):
        pytest.skip("host CPU lacks the V%d instruction set" % isa)


@pytest.mark.parametrize("isa", [3, 4])
def test_simd_golden_vector(O, isa):
    _need(O, isa)
    g = np.load(GOLDEN, allow_pickle=False)
    z = O.stdrng_f64(0, 2 * 2048).view(np.complex128).copy()
    got = O.simd_unordered(isa, z, 2, 32)  # Dif4, base_n = 32
    assert np.array_equal(got.view(np.float64), g.view(np.float64)), int(np.sum(got != g))


@pytest.mark.parametrize("isa", [3, 4])
@pytest.mark.parametrize("algo", DIF)
@pytest.mark.parametrize("n,base_n", [(64, 64), (256, 64), (1024, 1024), (1024, 512),
                                      (2048, 1024), (4096, 128), (32768, 1024)])
def test_simd_equals_scalar(O, isa, algo, n, base_n):
    _need(O, isa)
    r = {0: 2, 2: 4, 4: 8, 6: 16}[algo]
    if base_n // r < (2 if isa == 3 else 4):
        pytest.skip("a vector must fit one row of the first pass")
    rng = np.random.default_rng(n + base_n + algo)
    z = (rng.uniform(-1, 1, (3, n)) + 1j * rng.uniform(-1, 1, (3, n)))
    f_simd = O.simd_unordered(isa, z, algo, base_n, fwd=True)
    f_ref = O.unordered_fwd(z, algo, base_n)
    assert np.array_equal(f_simd.view(np.float64), f_ref.view(np.float64))
    i_simd = O.simd_unordered(isa, f_ref, algo, base_n, fwd=False)
    i_ref = O.unordered_inv(f_ref, algo, base_n)
    assert np.array_equal(i_simd.view(np.float64), i_ref.view(np.float64))


@pytest.mark.parametrize("isa", [3, 4])
@pytest.mark.parametrize("N,algo,base_n", [(2048, 6, 1024), (2048, 2, 32), (4096, 4, 256),
                                           (1024, 6, 512)])
def test_simd_product_equals_scalar(O, isa, N, algo, base_n):
    _need(O, isa)
    rng = np.random.default_rng(N + algo)
    a = rng.integers(0, 2**64, (5, N), dtype=np.uint64)
    b = rng.integers(-128, 128, (5, N), dtype=np.int64)
    # torus edge values in A: 0, 2^63, all ones, small magnitudes
    a[0, :8] = [0, 1 << 63, 2**64 - 1, 1, 2, 3, (1 << 63) - 1, (1 << 63) + 1]
    got = O.simd_polymul_f64(isa, a, b, algo, base_n, threads=4)
    want = O.polymul_f64(a, b, algo, base_n, threads=4)
    assert np.array_equal(got, want), int(np.sum(got != want))
