"""CPU pins of the product oracles: the u64 negacyclic schoolbook (against
Python integers and the X^N = -1 shift invariant), the oracle's f64 product
(torus error bound, exactness on a coarse grid), and the gadget
decomposition (recomposition identity and digit range)."""
import numpy as np
import pytest

from oracle_lib import load_oracle

M64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def O():
    return load_oracle()


def schoolbook_py(a, b):
    N = len(a)
    c = [0] * N
    for i in range(N):
        for j in range(N):
            k = i + j
            if k < N:
                c[k] += int(a[i]) * int(b[j])
            else:
                c[k - N] -= int(a[i]) * int(b[j])
    return [x & M64 for x in c]


@pytest.mark.parametrize("N", [1, 2, 4, 8, 32])
def test_schoolbook_vs_python_integers(O, N):
    rng = np.random.default_rng(N)
    a = rng.integers(0, 2**64, N, dtype=np.uint64)
    b = rng.integers(0, 2**64, N, dtype=np.uint64)
    got = O.schoolbook_u64(a[None], b[None])[0]
    assert [int(x) for x in got] == schoolbook_py(a, b)


def test_schoolbook_shift_invariant(O):
    # multiplying by X^k rotates with a sign flip on wrap (X^N = -1)
    N = 256
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2**64, N, dtype=np.uint64)
    for k in (0, 1, 77, N - 1):
        xk = np.zeros(N, dtype=np.uint64)
        xk[k] = 1
        got = O.schoolbook_u64(a[None], xk[None])[0]
        want = np.concatenate([(-(a[N - k:].view(np.int64))).view(np.uint64), a[:N - k]])
        assert np.array_equal(got, want)


def test_f64_product_error_bound_and_grid_exactness(O):
    N, batch = 2048, 4
    rng = np.random.default_rng(4)
    a = rng.integers(0, 2**64, (batch, N), dtype=np.uint64)
    b = rng.integers(-128, 128, (batch, N), dtype=np.int64)
    c = O.polymul_f64(a, b, 6, 1024)
    exact = O.schoolbook_u64(a, b.view(np.uint64))
    err = np.abs((c - exact).view(np.int64)).max()
    assert 0 < err < 2**32  # f64 FFT noise of a 64-bit torus product
    # coarse grid: exact after rounding to the grid
    a2 = (rng.integers(-2**19, 2**19, (batch, N), dtype=np.int64) << 44).view(np.uint64)
    b2 = rng.integers(-4, 4, (batch, N), dtype=np.int64)
    c2 = O.polymul_f64(a2, b2, 2, 32)
    ex2 = O.schoolbook_u64(a2, b2.view(np.uint64))
    assert np.array_equal(((c2 + np.uint64(1 << 43)) >> np.uint64(44)) << np.uint64(44), ex2)


def test_from_torus_closed_form(O):
    for x, want in [(0.0, 0), (0.25, 1 << 62), (-0.25, (1 << 64) - (1 << 62)), (0.5, 1 << 63),
                    (-0.5, 1 << 63), (3.75, 3 << 62), (2.0 ** -64, 1), (-(2.0 ** -64), M64),
                    (2.0 ** -66, 0)]:
        assert O.L.or_from_torus(x) == want, x


@pytest.mark.parametrize("levels,base_log", [(3, 8), (2, 10), (4, 6), (1, 16)])
def test_gadget_decomposition(O, levels, base_log):
    rng = np.random.default_rng(levels * 100 + base_log)
    x = rng.integers(0, 2**64, 4096, dtype=np.uint64)
    x[:4] = [0, M64, 1 << 63, (1 << 63) - 1]
    d = O.gadget_decompose(x, levels, base_log)
    half = 1 << (base_log - 1)
    assert d.min() >= -half and d.max() < half
    rep = levels * base_log
    for j in range(x.size):
        v = int(x[j])
        # closest multiple of 2^(64 - rep), ties up
        rounded = ((v >> (64 - rep - 1)) + 1) >> 1 << (64 - rep) if rep < 64 else v
        recomposed = sum(int(d[i, j]) << (64 - base_log * (i + 1)) for i in range(levels))
        assert recomposed & M64 == rounded & M64, (j, v)


def gadget_key(N, k1, levels, base_log):
    """Trivial GGSW of the message 1: row (p, i), column c holds the constant
    polynomial delta_{p,c} * 2^(64 - (i+1) B)."""
    key = np.zeros((k1 * levels, k1, N), dtype=np.uint64)
    for p in range(k1):
        for i in range(levels):
            key[p * levels + i, p, 0] = np.uint64(1 << (64 - (i + 1) * base_log))
    return key


def test_torus_to_fourier_constant_polynomial(O):
    # a constant polynomial c has the flat spectrum c * 2^-64 in every slot
    N = 512
    for c in (1 << 63, 3 << 60, (1 << 64) - (1 << 40)):
        poly = np.zeros((1, N), dtype=np.uint64)
        poly[0, 0] = np.uint64(c)
        f = O.torus_to_fourier(poly, 6, 256)[0]
        want = float(np.int64(np.uint64(c).view(np.int64))) * 2.0 ** -64
        assert np.all(f.real == want) and np.all(f.imag == 0.0)


@pytest.mark.parametrize("N,k1,levels,base_log", [(256, 2, 3, 8), (1024, 1, 2, 10), (64, 3, 4, 6)])
def test_external_product_trivial_ggsw_recomposes(O, N, k1, levels, base_log):
    # GGSW(1) (x) GLWE = gadget recomposition of the rounded GLWE
    rng = np.random.default_rng(N + k1)
    glwe = rng.integers(0, 2**64, (k1, N), dtype=np.uint64)
    kf = O.torus_to_fourier(gadget_key(N, k1, levels, base_log), 6, min(N // 2, 256))
    out = O.external_product(N, k1, levels, base_log, kf, 6, min(N // 2, 256), glwe)
    rep = levels * base_log
    half = np.uint64(1 << (64 - rep - 1))
    rounded = ((glwe + half) >> np.uint64(64 - rep)) << np.uint64(64 - rep)
    err = np.abs((out - rounded).view(np.int64)).max()
    assert err < 2**16, err


def test_external_product_coarse_grid_matches_schoolbook(O):
    # key on the 2^44 grid: the f64 product is exact after rounding to it
    N, k1, levels, base_log = 256, 2, 3, 8
    rng = np.random.default_rng(11)
    key = (rng.integers(-2**12, 2**12, (k1 * levels, k1, N), dtype=np.int64) << 44).view(np.uint64)
    glwe = rng.integers(0, 2**64, (k1, N), dtype=np.uint64)
    kf = O.torus_to_fourier(key, 2, 32)
    out = O.external_product(N, k1, levels, base_log, kf, 2, 32, glwe)
    digits = np.concatenate([O.gadget_decompose(glwe[p], levels, base_log) for p in range(k1)])
    want = np.zeros((k1, N), dtype=np.uint64)
    for c in range(k1):
        for r in range(k1 * levels):
            want[c] += O.schoolbook_u64(key[r, c][None], digits[r].view(np.uint64)[None])[0]
    got = ((out + np.uint64(1 << 43)) >> np.uint64(44)) << np.uint64(44)
    assert np.array_equal(got, want)


def _device_from_torus(x):
    """The device kernels' branch-free torus rounding (device_math.hpp
    from_torus), restated in numpy: every step is exact or a single
    correctly rounded IEEE operation, as on the GPU."""
    x = np.asarray(x, dtype=np.float64)
    t = x - np.rint(x)
    T = t * 2.0 ** 64
    Tr = np.trunc(T + np.copysign(0.49999999999999994, T))
    hi = np.floor(Tr * 2.0 ** -32)
    lo = Tr - hi * 2.0 ** 32  # exact: the fma of the device form
    hm = (hi + 6755399441055744.0).view(np.uint64) & np.uint64(0xFFFFFFFF)
    return (hm << np.uint64(32)) | lo.astype(np.uint64)


def test_device_torus_rounding_matches_oracle(O):
    """The device's rounding formula gives the oracle's bits (the reference
    rounding: fract = x - round(x), round(fract 2^64), 2^63 wrapped) on
    random values over the whole range products produce and on the ties:
    half-integers, odd multiples of 2^-65, neighbours of 2^-64 multiples and
    +-1/2."""
    rng = np.random.default_rng(11)
    e = rng.integers(-100, 80, 200000)
    xs = [np.ldexp(rng.random(200000), e) * rng.choice([-1.0, 1.0], 200000)]
    k = np.arange(-300, 301, dtype=np.float64)
    for base in (0.0, 0.25, 0.5, 0.75):
        xs.append(k + base)
    for ex in range(-100, 10, 3):
        q = np.arange(-20, 21, dtype=np.float64)
        xs.append(np.ldexp(q + 0.5, ex))
        xs.append(np.ldexp(q, ex))
    x = np.concatenate(xs)
    x = np.concatenate([x, np.nextafter(x, np.inf), np.nextafter(x, -np.inf)])
    got = _device_from_torus(x)
    want = np.array([O.L.or_from_torus(float(v)) for v in x], dtype=np.uint64)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (x[bad[:5]], got[bad[:5]], want[bad[:5]])
