"""fft128 (double-double negacyclic FFT): oracle pins on the CPU and GPU parity.

CPU pins of the oracle follow the reference's own tests with mpmath in place
of rug (src/fft128/f128_ops.rs:1043-1215): the double-double add/sub/mul/div
error bounds, sincospi accuracy, and the negacyclic product bound of
test_wrapper (src/fft128/mod.rs:1972-2065).  GPU tests run the device kernels
on the oracle's own twiddles through the free-function ABI and must match the
oracle bit for bit (scalar and AVX-style products); the plan path uses the
library's binary128 twiddles and is held to the reference's tolerance.
"""
import mpmath
import numpy as np
import pytest

from oracle_lib import load_oracle

mpmath.mp.prec = 240


@pytest.fixture(scope="module")
def O():
    return load_oracle()


def dd_value(hi, lo):
    return mpmath.mpf(hi) + mpmath.mpf(lo)


def rand_dd(rng):
    hi = rng.standard_normal()
    lo = rng.uniform(-1, 1) * abs(hi) * 2.0 ** -53
    return (float(hi), float(lo))


@pytest.mark.parametrize("op,bound", [("add", -104), ("sub", -104), ("mul", -104), ("div", -104),
                                      ("add_estimate", -101), ("sub_estimate", -101),
                                      ("mul_fma", -104)])
def test_f128_ops_error_bounds(O, op, bound):
    # test_add/sub/mul/div (src/fft128/f128_ops.rs:1043-1153), normal inputs:
    # accurate ops within 2^-104 of the result; the *_estimate ops within
    # 2^-101 of |a| + |b| (their bound under cancellation)
    rng = np.random.default_rng(sum(map(ord, op)))
    for _ in range(1000):
        a, b = rand_dd(rng), rand_dd(rng)
        got = dd_value(*O.f128(op, a, b))
        A, B = dd_value(*a), dd_value(*b)
        base = op.split("_")[0]
        exact = {"add": A + B, "sub": A - B, "mul": A * B, "div": A / B}[base]
        scale = abs(A) + abs(B) if "estimate" in op else abs(exact)
        assert abs(got - exact) <= scale * mpmath.mpf(2) ** bound, (op, a, b)


def test_f128_sincospi(O):
    # test_sincos (src/fft128/f128_ops.rs:1155-1215): 2^-103 relative-ish accuracy
    rng = np.random.default_rng(5)
    xs = list(rng.uniform(-1, 1, 2000)) + [0.0, 0.25, 0.5, 0.75, 1.0, -0.5]
    for x in xs:
        (sh, sl), (ch, cl) = O.f128_sincospi((float(x), 0.0))
        s, c = dd_value(sh, sl), dd_value(ch, cl)
        es, ec = mpmath.sinpi(mpmath.mpf(x)), mpmath.cospi(mpmath.mpf(x))
        assert abs(s - es) <= mpmath.mpf(2) ** -103 * max(abs(es), mpmath.mpf(2) ** -60)
        assert abs(c - ec) <= mpmath.mpf(2) ** -103 * max(abs(ec), mpmath.mpf(2) ** -60)


def test_oracle_twiddles_vs_mpmath(O):
    n = 256
    t = O.fft128_twiddles(n)
    m = 1
    while m < n:
        for i in range(m):
            k, pos = 2 * m + i, m + i
            br = int(bin(k)[2:].zfill((2 * n).bit_length() - 1)[::-1], 2)  # bitrev over 2n
            ang = mpmath.mpf(br) / (2 * n)
            assert abs(dd_value(t[0][pos], t[1][pos]) - mpmath.cospi(ang)) < mpmath.mpf(2) ** -104
            assert abs(dd_value(t[2][pos], t[3][pos]) - mpmath.sinpi(ang)) < mpmath.mpf(2) ** -104
        m *= 2


def test_library_twiddles_match_oracle(O):
    # library: binary128 evaluation; oracle: the reference's double-double
    # sincospi.  Same table within 2^-104 absolute.
    import cfft_amd
    for n in (32, 256, 2048):
        lib = cfft_amd.fft128.init_negacyclic_twiddles(n)
        ref = O.fft128_twiddles(n)
        for (h1, l1, h2, l2) in ((lib[0], lib[1], ref[0], ref[1]), (lib[2], lib[3], ref[2], ref[3])):
            for pos in range(1, n):
                d = dd_value(h1[pos], l1[pos]) - dd_value(h2[pos], l2[pos])
                assert abs(d) < mpmath.mpf(2) ** -104, (n, pos)


def negacyclic_exact(a, b):
    """Exact negacyclic convolution of double inputs, as mpmath numbers."""
    N = len(a)
    A = [mpmath.mpf(float(x)) for x in a]
    B = [mpmath.mpf(float(x)) for x in b]

    def coeff(k):
        s = mpmath.mpf(0)
        for i in range(N):
            j = k - i
            if j >= 0:
                s += A[i] * B[j]
            else:
                s -= A[i] * B[j + N]
        return s
    return coeff


def fft128_product(fwd, inv, a, b):
    """test_wrapper's fold / pointwise / unfold around given fwd / inv callables."""
    N = len(a)
    n = N // 2
    fa = [np.array(a[:n]), np.zeros(n), np.array(a[n:]), np.zeros(n)]
    fb = [np.array(b[:n]), np.zeros(n), np.array(b[n:]), np.zeros(n)]
    fa, fb = fwd(fa), fwd(fb)
    factor = 2.0 / N
    O = load_oracle()
    out = [np.zeros(n) for _ in range(4)]
    for i in range(n):
        ar, ai = (fa[0][i], fa[1][i]), (fa[2][i], fa[3][i])
        br, bi = (fb[0][i], fb[1][i]), (fb[2][i], fb[3][i])
        rr, ii = O.f128("mul", ar, br), O.f128("mul", ai, bi)
        ri, ir = O.f128("mul", ar, bi), O.f128("mul", ai, br)
        pre = O.f128("sub_estimate", rr, ii)
        pim = O.f128("add_estimate", ir, ri)
        out[0][i], out[1][i] = pre[0] * factor, pre[1] * factor
        out[2][i], out[3][i] = pim[0] * factor, pim[1] * factor
    out = inv(out)
    res = [dd_value(out[0][i], out[1][i]) for i in range(n)] + \
          [dd_value(out[2][i], out[3][i]) for i in range(n)]
    return res


def test_oracle_fft128_product_bound(O):
    rng = np.random.default_rng(8)
    N = 64
    a, b = rng.random(N), rng.random(N)
    tw = O.fft128_twiddles(N // 2)
    res = fft128_product(lambda x: O.fft128_fwd(x, tw), lambda x: O.fft128_inv(x, tw), a, b)
    coeff = negacyclic_exact(a, b)
    for k in range(N):
        assert abs(res[k] - coeff(k)) < mpmath.mpf(1e-30) * N


# ---------------- GPU ----------------

def dev(arrs):
    from cfft_amd.device import DeviceBuffer
    return [DeviceBuffer.from_host(np.ascontiguousarray(x)) for x in arrs]


@pytest.mark.gpu
# n > 4096 runs the stages above the 4096-point LDS blocks over global
# memory: 8192 (one round of 1 stage), 16384 (2), 32768 (3), 2^17 (3 + 2)
@pytest.mark.parametrize("n", [8, 32, 64, 512, 1024, 2048, 4096, 8192, 16384, 32768, 131072])
@pytest.mark.parametrize("mulv", [0, 1])
def test_gpu_free_functions_bit_exact(cfft, O, n, mulv):
    import cfft_amd
    print("fft128 free functions: n=%d mulv=%d" % (n, mulv), flush=True)
    rng = np.random.default_rng(n + mulv)
    batch = 3
    tw = O.fft128_twiddles(n)
    x = [rng.uniform(-1, 1, (batch, n)) for _ in range(4)]
    x[1] *= 2.0 ** -54
    x[3] *= 2.0 ** -54
    d = dev(x)
    dtw = dev(tw)
    cfft_amd.fft128.negacyclic_fwd_fft(*d, dtw, batch=batch, mul_variant=mulv)
    got = [b.download() for b in d]
    ref = O.fft128_fwd([a.reshape(-1) for a in x], tw, mulv)
    for g, r in zip(got, ref):
        assert np.array_equal(g.reshape(-1), r), (n, mulv)
    cfft_amd.fft128.negacyclic_inv_fft(*d, dtw, batch=batch, mul_variant=mulv)
    got2 = [b.download() for b in d]
    ref2 = O.fft128_inv(ref, tw, mulv)
    for g, r in zip(got2, ref2):
        assert np.array_equal(g.reshape(-1), r), (n, mulv)


@pytest.mark.gpu
# below 512 points a workgroup holds 64 / (n / 8) transforms (one wave); 37
# transforms leave the last workgroup partly idle
@pytest.mark.parametrize("n", [8, 16, 64, 128, 256])
def test_gpu_free_functions_packed_batches(cfft, O, n):
    import cfft_amd
    print("fft128 packed batch: n=%d" % n, flush=True)
    rng = np.random.default_rng(100 + n)
    batch = 37
    tw = O.fft128_twiddles(n)
    x = [rng.uniform(-1, 1, (batch, n)) for _ in range(4)]
    x[1] *= 2.0 ** -54
    x[3] *= 2.0 ** -54
    d = dev(x)
    dtw = dev(tw)
    cfft_amd.fft128.negacyclic_fwd_fft(*d, dtw, batch=batch, mul_variant=1)
    got = [b.download() for b in d]
    ref = O.fft128_fwd([a.reshape(-1) for a in x], tw, 1)
    for g, r in zip(got, ref):
        assert np.array_equal(g.reshape(-1), r), n
    cfft_amd.fft128.negacyclic_inv_fft(*d, dtw, batch=batch, mul_variant=1)
    got2 = [b.download() for b in d]
    ref2 = O.fft128_inv(ref, tw, 1)
    for g, r in zip(got2, ref2):
        assert np.array_equal(g.reshape(-1), r), n


@pytest.mark.gpu
@pytest.mark.parametrize("N", [64, 256, 2048, 16384, 32768])
def test_gpu_plan_product_bound(cfft, N):
    # test_wrapper (src/fft128/mod.rs:1972-2065) through fft128::Plan on the device
    import cfft_amd
    rng = np.random.default_rng(N)
    a, b = rng.random(N), rng.random(N)
    plan = cfft_amd.fft128.Plan(N // 2)
    assert plan.fft_size() == N // 2

    def fwd(x):
        x = [np.ascontiguousarray(v) for v in x]
        plan.fwd(*x)
        return x

    def inv(x):
        x = [np.ascontiguousarray(v) for v in x]
        plan.inv(*x)
        return x

    res = fft128_product(fwd, inv, a, b)
    coeff = negacyclic_exact(a, b)
    ks = range(N) if N <= 256 else np.random.default_rng(1).choice(N, 24, replace=False)
    for k in ks:
        assert abs(res[int(k)] - coeff(int(k))) < mpmath.mpf(1e-30) * N, k


@pytest.mark.gpu
def test_gpu_plan_rejects_bad_sizes(cfft):
    import cfft_amd
    with pytest.raises(cfft_amd.CfftPanic):
        cfft_amd.fft128.Plan(16)
    with pytest.raises(cfft_amd.CfftPanic):
        cfft_amd.fft128.Plan(100)
    with pytest.raises(cfft_amd.CfftPanic):
        cfft_amd.fft128.Plan(1 << 21)
