"""GPU parity of the f64 transforms against the oracle (bit-exact) and the
reference's own test relations."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "unordered_n2048_dif4_b32.npy")
ALGOS = list(range(8))


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.float64), np.asarray(b).view(np.float64)) or \
        np.array_equal(a, b)


def rand_c(rng, shape):
    return rng.random(shape) + 1j * rng.random(shape)


def test_golden_vector_on_gpu(cfft, oracle):
    from cfft_amd.unordered import FftAlgo, Method, Plan
    z = oracle.stdrng_f64(0, 4096).view(np.complex128).copy()
    plan = Plan(2048, Method.UserProvided(FftAlgo.Dif4, 32))
    plan.fwd(z)
    g = np.load(GOLDEN, allow_pickle=False)
    assert np.array_equal(z, g), "GPU differs from the reference golden vector at %d places" % int(
        np.sum(z != g))


@pytest.mark.parametrize("algo", ALGOS)
def test_ordered_bit_exact(cfft, oracle, algo):
    from cfft_amd.ordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(10 + algo)
    for e in range(0, 11):
        n = 1 << e
        plan = Plan(n, Method.UserProvided(FftAlgo(algo)))
        assert plan.fft_size() == n and plan.algo() == FftAlgo(algo)
        x = rand_c(rng, (5, n))
        y = x.copy()
        plan.fwd(y)
        assert np.array_equal(y, oracle.ordered_fwd(x, algo)), (algo, n)
        w = y.copy()
        plan.inv(w)
        assert np.array_equal(w, oracle.ordered_inv(y, algo)), (algo, n)
        if n > 1:
            assert np.max(np.abs(w / n - x)) < 1e-14


@pytest.mark.parametrize("algo", ALGOS)
def test_ordered_large_bit_exact(cfft, oracle, algo):
    # cfft_ordered_plan_create_large: 2^11 <= n <= 2^16 against the oracle's
    # ordered recursion at the same sizes (itself pinned to the closed-form
    # DFT on the CPU)
    from cfft_amd import CfftPanic, check, lib
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.ordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(40 + algo)
    for n in (2048, 4096, 8192):
        plan = Plan(n, Method.UserProvided(FftAlgo(algo)), large=True)
        assert plan.fft_size() == n and plan.algo() == FftAlgo(algo)
        x = rand_c(rng, (3, n))
        y = x.copy()
        plan.fwd(y)
        assert np.array_equal(y, oracle.ordered_fwd(x, algo)), (algo, n)
        w = y.copy()
        plan.inv(w)
        assert np.array_equal(w, oracle.ordered_inv(y, algo)), (algo, n)
        assert np.max(np.abs(w / n - x)) < 1e-14
    # 2^14..2^16: every pass over global memory (ping-pong through a scratch;
    # odd and even numbers of core passes), strided batches included
    for n in (16384, 32768, 65536):
        plan = Plan(n, Method.UserProvided(FftAlgo(algo)), large=True)
        x = rand_c(rng, (2, n))
        y = x.copy()
        plan.fwd(y)
        assert np.array_equal(y, oracle.ordered_fwd(x, algo)), (algo, n)
        w = y.copy()
        plan.inv(w)
        assert np.array_equal(w, oracle.ordered_inv(y, algo)), (algo, n)
        assert np.max(np.abs(w / n - x)) < 1e-13
    n = 16384
    plan = Plan(n, Method.UserProvided(FftAlgo(algo)), large=True)
    x = rand_c(rng, (3, n + 64))
    dx = DeviceBuffer.from_host(x)
    check(lib.cfft_fwd(plan.handle, dx.dptr, 3, n + 64, None))
    got = dx.download()
    assert np.array_equal(got[:, :n], oracle.ordered_fwd(x[:, :n], algo))
    assert np.array_equal(got[:, n:], x[:, n:])  # the gaps between transforms untouched
    with pytest.raises(CfftPanic, match="2\\^16"):
        Plan(1 << 17, Method.UserProvided(FftAlgo(algo)), large=True)
    # below 2^11 the large entry point builds the same plan as the capped one
    small = Plan(512, Method.UserProvided(FftAlgo(algo)), large=True)
    x = rand_c(rng, (2, 512))
    y = x.copy()
    small.fwd(y)
    assert np.array_equal(y, oracle.ordered_fwd(x, algo))


def test_ordered_large_measure(cfft, oracle):
    # the device autotuner over the 8 algorithms at an extended size: the
    # chosen plan computes that algorithm's bits
    from cfft_amd.ordered import Method, Plan
    rng = np.random.default_rng(77)
    plan = Plan(4096, Method.Measure(0.002), large=True)
    algo = int(plan.algo())
    x = rand_c(rng, (2, 4096))
    y = x.copy()
    plan.fwd(y)
    assert np.array_equal(y, oracle.ordered_fwd(x, algo)), algo


@pytest.mark.parametrize("n", [32, 64, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_unordered_bit_exact(cfft, oracle, n):
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(n)
    bases = sorted({b for b in (32, 64, 128, 512, 1024, n) if b <= n and b <= 1024 and (b >= 32 or b == n)})
    for base_n in bases:
        for algo in ALGOS:
            plan = Plan(n, Method.UserProvided(FftAlgo(algo), base_n))
            assert plan.algo() == (FftAlgo(algo), base_n)
            x = rand_c(rng, (3, n))
            y = x.copy()
            plan.fwd(y)
            ref = oracle.unordered_fwd(x, algo, base_n)
            assert np.array_equal(y, ref), (n, base_n, algo, int(np.sum(y != ref)))
            w = y.copy()
            plan.inv(w)
            assert np.array_equal(w, oracle.unordered_inv(y, algo, base_n)), (n, base_n, algo)


@pytest.mark.parametrize("n,base_n", [(32768, 1024), (65536, 512), (65536, 1024),
                                      (131072, 1024)])
def test_unordered_large_bit_exact(cfft, oracle, n, base_n):
    # levels above one LDS chunk run as global depth passes; with Dif16 and
    # base_n = 1024 the chunks (4096, 8192 and 2048 points here) run on the
    # wave-local kernels
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(n)
    for algo in (2, 6, 7):
        plan = Plan(n, Method.UserProvided(FftAlgo(algo), base_n))
        x = rand_c(rng, (2, n))
        y = x.copy()
        plan.fwd(y)
        assert np.array_equal(y, oracle.unordered_fwd(x, algo, base_n)), (n, base_n, algo)
        w = y.copy()
        plan.inv(w)
        assert np.array_equal(w, oracle.unordered_inv(y, algo, base_n))
        assert np.max(np.abs(w / n - x)) < 1e-12


@pytest.mark.parametrize("n,stride,batch,base", [(512, 530, 9, 512), (512, 512, 33, 512),
                                                 (1024, 1100, 7, 1024), (1024, 1030, 7, 512),
                                                 (1024, 1024, 32, 512), (2048, 2050, 11, 1024),
                                                 (4096, 4100, 5, 1024), (8192, 8200, 3, 1024),
                                                 (16384, 16400, 2, 1024), (16384, 16384, 5, 1024),
                                                 (8192, 8192, 1, 1024), (8192, 8192, 6, 1024)])
def test_device_path_with_stride(cfft, oracle, n, stride, batch, base):
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(1)
    x = rand_c(rng, (batch, stride))
    plan = Plan(n, Method.UserProvided(FftAlgo.Dif16, base))
    s = Stream()
    d = DeviceBuffer.from_host(x)
    plan.fwd(d, stream=s, batch=batch)
    # the C ABI takes the stride directly
    import cfft_amd
    d2 = DeviceBuffer.from_host(x)
    cfft_amd.check(cfft_amd.lib.cfft_fwd(plan.handle, d2.dptr, batch, stride, s.handle))
    s.synchronize()
    y = d2.download()
    ref = oracle.unordered_fwd(np.ascontiguousarray(x[:, :n]), 6, base)
    assert np.array_equal(y[:, :n], ref)
    assert np.array_equal(y[:, n:], x[:, n:]), "padding between transforms was touched"
    cfft_amd.check(cfft_amd.lib.cfft_inv(plan.handle, d2.dptr, batch, stride, s.handle))
    s.synchronize()
    w = d2.download()
    assert np.array_equal(w[:, :n], oracle.unordered_inv(ref, 6, base))
    assert np.array_equal(w[:, n:], x[:, n:])


@pytest.mark.parametrize("n,stride,batch,algo,base", [(512, 530, 9, 2, 512), (1024, 1100, 7, 5, 1024),
                                                      (1024, 1040, 5, 7, 512), (2048, 2050, 5, 0, 1024),
                                                      (4096, 4100, 3, 3, 512), (8192, 8200, 2, 5, 1024)])
def test_compile_time_programs_with_stride(cfft, oracle, n, stride, batch, algo, base):
    """Plans run by the compile-time pass programs (Dif4/512, Dit8/1024,
    Dit16/512, Dif2/1024, Dit4/512, Dit8/1024 at n = 8192): strided batches in place through the C
    ABI, bit-exact against the oracle, padding between transforms untouched."""
    import cfft_amd
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(n + algo)
    x = rand_c(rng, (batch, stride))
    plan = Plan(n, Method.UserProvided(FftAlgo(algo), base))
    s = Stream()
    d = DeviceBuffer.from_host(x)
    cfft_amd.check(cfft_amd.lib.cfft_fwd(plan.handle, d.dptr, batch, stride, s.handle))
    s.synchronize()
    y = d.download()
    ref = oracle.unordered_fwd(np.ascontiguousarray(x[:, :n]), algo, base)
    assert np.array_equal(y[:, :n], ref), (n, algo, base)
    assert np.array_equal(y[:, n:], x[:, n:]), "padding between transforms was touched"
    cfft_amd.check(cfft_amd.lib.cfft_inv(plan.handle, d.dptr, batch, stride, s.handle))
    s.synchronize()
    w = d.download()
    assert np.array_equal(w[:, :n], oracle.unordered_inv(ref, algo, base)), (n, algo, base)
    assert np.array_equal(w[:, n:], x[:, n:])


@pytest.mark.parametrize("n", [256, 512, 1024])
def test_fwd_monomial(cfft, oracle, n):
    # test_fwd_monomial (src/unordered.rs:1108-1137): fwd_monomial(d) == fwd(e_d)
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(n)
    for base_n in sorted({32, n, n // 2, n // 4, n // 8}):
        plan = Plan(n, Method.UserProvided(FftAlgo.Dif4, base_n))
        for degree in list(rng.integers(0, n, 6)) + [0, n - 1]:
            z = np.zeros(n, dtype=np.complex128)
            z[degree] = 1.0
            target = z.copy()
            plan.fwd(target)
            m = np.zeros(n, dtype=np.complex128)
            plan.fwd_monomial(int(degree), m)
            assert np.max(np.abs(m - target)) < 1e-12
            # closed form: position i holds T[(freq(i) * d) mod n], T = sincospi64(-2j/n)
            perm = plan.permutation()
            freq = np.empty(n, dtype=np.int64)
            freq[perm] = np.arange(n)
            for i in range(0, n, max(1, n // 16)):
                s, c = oracle.sincospi64(-2.0 * ((int(freq[i]) * int(degree)) % n) / n)
                assert m[i] == complex(c, s)


def test_roundtrip_and_numpy(cfft):
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(9)
    for n in (32, 64, 256, 512, 1024):
        plan = Plan(n, Method.UserProvided(FftAlgo.Dif4, 32))
        x = rand_c(rng, n)
        z = x.copy()
        plan.fwd(z)
        perm = plan.permutation()
        assert np.max(np.abs(z[perm] - np.fft.fft(x))) < 1e-12
        plan.inv(z)
        assert np.max(np.abs(z / n - x)) < 1e-12


def test_serde_cross_plan_bincode(cfft):
    # test_serde (src/unordered.rs:9411-9466) through the bincode byte stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(4)
    for n in (64, 128, 256, 512, 1024):
        x = rand_c(rng, n)
        p1 = Plan(n, Method.UserProvided(FftAlgo.Dif4, 32))
        p2 = Plan(n, Method.UserProvided(FftAlgo.Dif4, 64))
        z = x.copy()
        p1.fwd(z)
        data = p1.serialize_fourier_buffer(z)
        assert len(data) == (1 if n < 251 else 3) + 16 * n
        z2 = np.zeros(n, dtype=np.complex128)
        p2.deserialize_fourier_buffer(data, z2)
        p2.inv(z2)
        assert np.max(np.abs(z2 / n - x)) < 1e-12
        # standard order equals numpy's DFT
        std = np.frombuffer(data[len(data) - 16 * n:], dtype=np.complex128)
        assert np.max(np.abs(std - np.fft.fft(x))) < 1e-12


def test_device_serde_gather_scatter(cfft):
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.unordered import FftAlgo, Method, Plan
    n, batch = 1024, 4
    rng = np.random.default_rng(2)
    x = rand_c(rng, (batch, n))
    p1 = Plan(n, Method.UserProvided(FftAlgo.Dif8, 32))
    p2 = Plan(n, Method.UserProvided(FftAlgo.Dit16, 512))
    d = DeviceBuffer.from_host(x)
    p1.fwd(d)
    std = DeviceBuffer((batch, n), np.complex128)
    p1.fourier_to_standard(d, std, batch)
    s = std.download()
    assert np.max(np.abs(s - np.fft.fft(x, axis=1))) < 1e-11
    d2 = DeviceBuffer((batch, n), np.complex128)
    p2.fourier_from_standard(std, d2, batch)
    p2.inv(d2)
    assert np.max(np.abs(d2.download() / n - x)) < 1e-12


def test_measure_method_picks_a_valid_plan(cfft):
    from cfft_amd import ordered, unordered
    p = unordered.Plan(2048, unordered.Method.Measure(0.002))
    algo, base_n = p.algo()
    assert base_n in (512, 1024)
    q = ordered.Plan(256, ordered.Method.Measure(0.002))
    assert q.fft_size() == 256
    rng = np.random.default_rng(0)
    x = rand_c(rng, 2048)
    z = x.copy()
    p.fwd(z)
    p.inv(z)
    assert np.max(np.abs(z / 2048 - x)) < 1e-12


@pytest.mark.parametrize("batch,stride", [(1, 1024), (9, 1024), (5, 1040)])
def test_dit16_register_form(cfft, oracle, batch, stride):
    """Dit16 with n = base_n = 1024 runs the register-resident Dit16 kernel
    (one transform per wave): bit-exact against the oracle, strided batches
    leave the padding alone."""
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    import cfft_amd
    n = 1024
    rng = np.random.default_rng(batch + stride)
    x = rand_c(rng, (batch, stride))
    plan = Plan(n, Method.UserProvided(FftAlgo.Dit16, 1024))
    s = Stream()
    d = DeviceBuffer.from_host(x)
    cfft_amd.check(cfft_amd.lib.cfft_fwd(plan.handle, d.dptr, batch, stride, s.handle))
    s.synchronize()
    y = d.download()
    ref = oracle.unordered_fwd(np.ascontiguousarray(x[:, :n]), 7, 1024)
    assert np.array_equal(y[:, :n], ref)
    assert np.array_equal(y[:, n:], x[:, n:])
    cfft_amd.check(cfft_amd.lib.cfft_inv(plan.handle, d.dptr, batch, stride, s.handle))
    s.synchronize()
    w = d.download()
    assert np.array_equal(w[:, :n], oracle.unordered_inv(ref, 7, 1024))
    assert np.array_equal(w[:, n:], x[:, n:])
