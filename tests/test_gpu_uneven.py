"""Uneven batch on the GPU: every job's Fourier buffer and round trip are
bit-exact against the oracle with the plan its size class runs; the
persistent graph form and the plain per-class form agree; graph replays with
new inputs between them each match; the full 20 000-job workload keeps its
round-trip error at f64 level."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [9, 10, 11, 12, 13, 14]
WEIGHTS = [0.1, 0.2, 0.35, 0.2, 0.1, 0.05]


def _jobs_with_gaps(log2_N, rng):
    """Jobs in random order with random gaps between their buffers."""
    from cfft_amd.uneven import JOB_DTYPE
    jobs = np.zeros(len(log2_N), dtype=JOB_DTYPE)
    pin = pf = pout = 0
    for i, l in enumerate(log2_N):
        N = 1 << int(l)
        pin += int(rng.integers(0, 5))
        pf += int(rng.integers(0, 3))
        pout += int(rng.integers(0, 7))
        jobs[i] = (pin, pf, pout, l, 0)
        pin += N
        pf += N // 2
        pout += N
    return jobs, pin, pf, pout


def _written(jobs, nf, nout):
    """Masks of the Fourier and output positions the jobs write (the gaps
    between jobs are never written, so fresh buffers differ there)."""
    mf = np.zeros(nf, dtype=bool)
    my = np.zeros(nout, dtype=bool)
    for j in jobs:
        N = 1 << int(j["log2_N"])
        mf[int(j["fourier_offset"]):int(j["fourier_offset"]) + N // 2] = True
        my[int(j["out_offset"]):int(j["out_offset"]) + N] = True
    return mf, my


def _same_written(a, b, mask):
    return np.array_equal(a[mask].view(np.uint64), b[mask].view(np.uint64))


def _check_against_oracle(oracle, sched, jobs, x, F, y):
    for l in np.unique(jobs["log2_N"]):
        algo, base_n = sched.class_algo(int(l))
        idx = np.nonzero(jobs["log2_N"] == l)[0]
        N = 1 << int(l)
        xs = np.stack([x[jobs["in_offset"][i]:jobs["in_offset"][i] + N] for i in idx])
        Fw, yw = oracle.negacyclic_roundtrip(xs, int(algo), base_n)
        for k, i in enumerate(idx):
            f0, o0 = int(jobs["fourier_offset"][i]), int(jobs["out_offset"][i])
            assert np.array_equal(F[f0:f0 + N // 2].view(np.uint64), Fw[k].view(np.uint64)), \
                ("fourier", int(l), int(i))
            assert np.array_equal(y[o0:o0 + N].view(np.uint64), yw[k].view(np.uint64)), \
                ("roundtrip", int(l), int(i))


@pytest.mark.parametrize("measure", [False, True])
def test_uneven_bit_exact(cfft, oracle, measure):
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.uneven import Schedule
    rng = np.random.default_rng(3 + measure)
    log2_N = rng.choice([5, 6, 8] + SIZES, size=400)
    jobs, nin, nf, nout = _jobs_with_gaps(log2_N, rng)
    x = rng.uniform(-1, 1, nin)
    sched = Schedule(jobs, nin, nf, nout, measure=measure)
    info = sched.info()
    assert info["jobs"] == 400 and info["classes"] == len(np.unique(log2_N))
    dx = DeviceBuffer.from_host(x)
    dF = DeviceBuffer((nf,), np.complex128)
    dy = DeviceBuffer((nout,), np.float64)
    sched.run(dx, dF, dy)
    F, y = dF.download(), dy.download()
    _check_against_oracle(oracle, sched, jobs, x, F, y)
    # plain form: same bits
    mf, my = _written(jobs, nf, nout)
    dF2 = DeviceBuffer((nf,), np.complex128)
    dy2 = DeviceBuffer((nout,), np.float64)
    sched.run_plain(dx, dF2, dy2)
    assert _same_written(dF2.download(), F, mf)
    assert _same_written(dy2.download(), y, my)


def test_uneven_dif16_body_bit_exact(cfft, oracle):
    # Dif16 plans: N = 512 .. 16384 run the register-resident Dif16 body
    # (wave-local for N <= 2048, one depth pass around 1024-point base
    # transforms above), N <= 256 the generic body, both in one schedule;
    # partial last tiles in every class
    from cfft_amd.device import DeviceBuffer, synchronize
    from cfft_amd.uneven import Schedule
    from cfft_amd.unordered import FftAlgo
    rng = np.random.default_rng(29)
    log2_N = rng.choice(np.arange(5, 15), size=300)
    jobs, nin, nf, nout = _jobs_with_gaps(log2_N, rng)
    sched = Schedule(jobs, nin, nf, nout, algo=FftAlgo.Dif16)
    for l in range(9, 15):
        assert sched.class_algo(l) == (FftAlgo.Dif16, min(1 << (l - 1), 1024))
    assert sched.tile_schedule() == "device_queue"
    dx = DeviceBuffer((nin,), np.float64)
    dF = DeviceBuffer((nf,), np.complex128)
    dy = DeviceBuffer((nout,), np.float64)
    for rep in range(2):
        x = rng.uniform(-1, 1, nin)
        dx.upload(x)
        sched.run(dx, dF, dy)
        synchronize()
        F, y = dF.download(), dy.download()
        _check_against_oracle(oracle, sched, jobs, x, F, y)
    # every launch form and both tile schedules: the same bits
    mf, my = _written(jobs, nf, nout)
    forms = [("device_queue", sched.run_direct), ("device_queue", sched.run_plain),
             ("static_deal", sched.run), ("static_deal", sched.run_direct)]
    for kind, run in forms:
        sched.set_tile_schedule(kind)
        dF2 = DeviceBuffer((nf,), np.complex128)
        dy2 = DeviceBuffer((nout,), np.float64)
        run(dx, dF2, dy2)
        assert _same_written(dF2.download(), F, mf), kind
        assert _same_written(dy2.download(), y, my), kind


def test_uneven_graph_replays(cfft, oracle):
    from cfft_amd.device import DeviceBuffer, synchronize
    from cfft_amd.uneven import Schedule
    rng = np.random.default_rng(17)
    log2_N = rng.choice(SIZES, size=120, p=WEIGHTS)
    jobs, nin, nf, nout = _jobs_with_gaps(log2_N, rng)
    sched = Schedule(jobs, nin, nf, nout)
    dx = DeviceBuffer((nin,), np.float64)
    dF = DeviceBuffer((nf,), np.complex128)
    dy = DeviceBuffer((nout,), np.float64)
    for rep in range(3):
        x = rng.uniform(-1, 1, nin)
        dx.upload(x)
        sched.run(dx, dF, dy)
        synchronize()
        _check_against_oracle(oracle, sched, jobs, x, dF.download(), dy.download())
    # new buffer addresses: the graph is captured again
    dy2 = DeviceBuffer((nout,), np.float64)
    sched.run(dx, dF, dy2)
    assert _same_written(dy2.download(), dy.download(), _written(jobs, nf, nout)[1])


def test_uneven_full_workload_roundtrip(cfft):
    # 20 000 jobs, N from {512 .. 16384} with the stated weights
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.uneven import Schedule, packed_jobs
    rng = np.random.default_rng(3)
    log2_N = rng.choice(SIZES, size=20000, p=WEIGHTS)
    jobs, nin, nf, nout = packed_jobs(log2_N)
    x = rng.uniform(-1, 1, nin)
    sched = Schedule(jobs, nin, nf, nout, measure=True)
    dx = DeviceBuffer.from_host(x)
    dF = DeviceBuffer((nf,), np.complex128)
    dy = DeviceBuffer((nout,), np.float64)
    sched.run(dx, dF, dy)
    y = dy.download()
    assert np.abs(y - x).max() < 1e-13
    # Parseval per job on a sample: sum |F|^2 = n * sum |z|^2
    F = dF.download()
    for i in rng.choice(20000, 50, replace=False):
        N = 1 << int(jobs["log2_N"][i])
        o, f = int(jobs["in_offset"][i]), int(jobs["fourier_offset"][i])
        e_x = np.sum(x[o:o + N] ** 2)
        e_f = np.sum(np.abs(F[f:f + N // 2]) ** 2) / (N // 2)
        assert abs(e_f - e_x) < 1e-12 * e_x


@pytest.mark.parametrize("log2_N", [[14], [11] * 300, [5, 14], [9]])
def test_uneven_degenerate_schedules(cfft, oracle, log2_N):
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.uneven import Schedule, packed_jobs
    rng = np.random.default_rng(len(log2_N))
    jobs, nin, nf, nout = packed_jobs(log2_N)
    x = rng.uniform(-1, 1, nin)
    sched = Schedule(jobs, nin, nf, nout, measure=True)
    dx = DeviceBuffer.from_host(x)
    dF = DeviceBuffer((nf,), np.complex128)
    dy = DeviceBuffer((nout,), np.float64)
    sched.run(dx, dF, dy)
    _check_against_oracle(oracle, sched, jobs, x, dF.download(), dy.download())
