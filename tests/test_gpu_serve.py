"""Request server on the GPU: every result is bit-identical alone and in a
mixed burst, equals the direct batched entry points, survives admission
pressure (a budget of a few requests), and an open-loop Poisson replay
completes with correct results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M = 2**64
K1, LEVELS, BASE_LOG = 2, 3, 8


def _setup(server, rng, sizes=(1024, 2048, 4096)):
    keys = {}
    for N in sizes:
        key = rng.integers(0, M, (K1 * LEVELS, K1, N), dtype=np.uint64)
        keys[N] = (server.add_key(key, K1, LEVELS, BASE_LOG), key)
    return keys


def _requests(server, keys, rng, count):
    reqs = []
    for i in range(count):
        if i % 4 == 3:
            N = int(rng.choice([1024, 2048]))
            a = rng.integers(0, M, N, dtype=np.uint64)
            b = rng.integers(-128, 128, N, dtype=np.int64)
            reqs.append(("polymul", N, (a, b)))
        else:
            N = int(rng.choice(sorted(keys)))
            reqs.append(("extprod", N, rng.integers(0, M, (K1, N), dtype=np.uint64)))
    return reqs


def _make(server, keys, r):
    kind, N, data = r
    if kind == "polymul":
        return server.polymul_request(*data)
    return server.extprod_request(keys[N][0], data)


def test_alone_equals_mix_equals_direct(cfft):
    from cfft_amd import products
    from cfft_amd.serve import Server
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(5)
    srv = Server()
    keys = _setup(srv, rng)
    reqs = _requests(srv, keys, rng, 48)
    alone = []
    for r in reqs:
        q, keep = _make(srv, keys, r)
        alone.append(srv.wait(srv.submit(q, keep)).copy())
    tickets = []
    for r in reqs:
        q, keep = _make(srv, keys, r)
        tickets.append(srv.submit(q, keep))
    mix = [srv.wait(t) for t in tickets]
    for i, (x, y) in enumerate(zip(alone, mix)):
        assert np.array_equal(x, y), i
    st = srv.stats()
    assert st["completed"] == 96 and st["batches"] < 96
    # direct entry points, same plans (Dif16, base_n = min(N/2, 1024))
    for i, (kind, N, data) in enumerate(reqs):
        plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, min(N // 2, 1024)))
        if kind == "polymul":
            want = products.negacyclic_polymul_u64_host(plan, data[0][None], data[1][None])[0]
        else:
            from cfft_amd.device import DeviceBuffer
            key = keys[N][1]
            dk = DeviceBuffer.from_host(key)
            kf = DeviceBuffer((K1 * LEVELS, K1, N // 2), np.complex128)
            products.ggsw_to_fourier(plan, dk, K1, LEVELS, kf)
            dg = DeviceBuffer.from_host(data[None])
            do = DeviceBuffer((1, K1, N), np.uint64)
            products.external_product(plan, kf, K1, LEVELS, BASE_LOG, dg, do, 1)
            want = do.download()[0]
        assert np.array_equal(alone[i], want), (i, kind, N)
    srv.close()


def test_admission_under_small_budget(cfft):
    from cfft_amd.serve import Server
    rng = np.random.default_rng(6)
    # room for two N = 4096 external products: 64 KiB in + 64 KiB out each in
    # pinned staging; on the device also the split form's Fourier scratch,
    # K1 * LEVELS spectra of 2048 c64 (192 KiB) each
    pinned = 2 * 2 * 65536 + 4096
    budget = 2 * (2 * 65536 + K1 * LEVELS * 2048 * 16) + 4096
    srv = Server(hbm_budget_bytes=budget, pinned_bytes=pinned)
    keys = _setup(srv, rng, sizes=(4096,))
    reqs = [("extprod", 4096, rng.integers(0, M, (K1, 4096), dtype=np.uint64)) for _ in range(40)]
    alone = []
    for r in reqs[:5]:
        q, keep = _make(srv, keys, r)
        alone.append(srv.wait(srv.submit(q, keep)).copy())
    tickets = []
    for r in reqs:
        q, keep = _make(srv, keys, r)
        tickets.append(srv.submit(q, keep))
    outs = [srv.wait(t) for t in tickets]
    for i in range(5):
        assert np.array_equal(outs[i], alone[i])
    st = srv.stats()
    assert st["completed"] == 45
    # a request larger than the pinned budget (N = 16384 product: 384 KiB) is
    # refused at submission
    from cfft_amd import CfftPanic
    big = srv.polymul_request(np.zeros(16384, np.uint64), np.zeros(16384, np.int64))
    with pytest.raises(CfftPanic) as ei:
        srv.submit(*big)
    assert ei.value.status == 14
    srv.close()
    # a device pool that holds one request's input and output but not its
    # Fourier scratch refuses the request instead of allocating beside it
    small = Server(hbm_budget_bytes=2 * 65536 + 4096, pinned_bytes=pinned)
    kid = small.add_key(keys[4096][1], K1, LEVELS, BASE_LOG)
    with pytest.raises(CfftPanic) as ei:
        small.submit(*small.extprod_request(kid, reqs[0][2]))
    assert ei.value.status == 14
    small.close()


def test_poisson_replay(cfft):
    from cfft_amd.serve import Server
    rng = np.random.default_rng(7)
    srv = Server()
    keys = _setup(srv, rng)
    reqs = _requests(srv, keys, rng, 2000)
    made = [_make(srv, keys, r) for r in reqs]
    arrivals = np.cumsum(rng.exponential(1.0 / 20000.0, len(reqs)))
    lat = srv.replay([q for q, _ in made], arrivals)
    assert np.all(lat > 0) and np.all(np.isfinite(lat))
    # spot-check results against the same requests served alone
    for i in rng.choice(len(reqs), 20, replace=False):
        q, keep = _make(srv, keys, reqs[i])
        alone = srv.wait(srv.submit(q, keep))
        assert np.array_equal(made[i][1][-1], alone), i
    print("p50 %.1f us p99 %.1f us" % (np.percentile(lat, 50) * 1e6, np.percentile(lat, 99) * 1e6))
    srv.close()


def test_concurrent_submitters(cfft):
    # four host threads submit and wait at once; every result equals the
    # same request served alone
    import threading
    from cfft_amd.serve import Server
    rng = np.random.default_rng(8)
    srv = Server()
    keys = _setup(srv, rng)
    reqs = _requests(srv, keys, rng, 160)
    alone = []
    for r in reqs:
        q, keep = _make(srv, keys, r)
        alone.append(srv.wait(srv.submit(q, keep)).copy())
    out = [None] * len(reqs)
    errors = []
    lock = threading.Lock()

    def worker(idx):
        try:
            for i in idx:
                q, keep = _make(srv, keys, reqs[i])
                with lock:  # the wrapper's ticket table is a plain dict
                    t = srv.submit(q, keep)
                res = srv.wait(t)
                out[i] = res
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(range(k, len(reqs), 4),)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i in range(len(reqs)):
        assert np.array_equal(out[i], alone[i]), i
    srv.close()


def test_ticket_lifetime(cfft):
    """A polled ticket can be released; a released or waited ticket is
    unknown afterwards; a ticket can be waited on once."""
    import threading
    import time

    from cfft_amd import CfftPanic
    from cfft_amd.serve import Server
    rng = np.random.default_rng(7)
    srv = Server()
    a = rng.integers(0, M, 1024, dtype=np.uint64)
    b = rng.integers(-128, 128, 1024, dtype=np.int64)
    q, keep = srv.polymul_request(a, b)
    t1 = srv.submit(q, keep)
    deadline = time.time() + 10
    while not srv.poll(t1):
        assert time.time() < deadline
        time.sleep(1e-4)
    srv.release(t1)
    with pytest.raises(CfftPanic):
        srv.poll(t1)
    with pytest.raises(CfftPanic):
        srv.wait(t1)
    # released before completion: forgotten when it completes, output written
    q2, keep2 = srv.polymul_request(a, b)
    t2 = srv.submit(q2, keep2)
    srv.release(t2)
    q3, keep3 = srv.polymul_request(a, b)
    out3 = srv.wait(srv.submit(q3, keep3))
    np.testing.assert_array_equal(out3, keep[-1])
    # many tickets in flight while one thread waits (the map rehashes under it)
    q4, keep4 = srv.polymul_request(a, b)
    t4 = srv.submit(q4, keep4)
    got = {}
    th = threading.Thread(target=lambda: got.setdefault("r", srv.wait(t4)))
    th.start()
    more = [srv.submit(*srv.polymul_request(a, b)) for _ in range(300)]
    th.join(30)
    assert not th.is_alive()
    np.testing.assert_array_equal(got["r"], keep[-1])
    for t in more:
        srv.wait(t)
    srv.close()


def test_two_waiters_one_ticket(cfft):
    """Two threads waiting on one ticket: exactly one gets the result, the
    other an error (never a read of a forgotten ticket)."""
    import threading

    from cfft_amd import CfftPanic
    from cfft_amd.serve import Server
    rng = np.random.default_rng(8)
    srv = Server()
    a = rng.integers(0, M, 2048, dtype=np.uint64)
    b = rng.integers(-128, 128, 2048, dtype=np.int64)
    for _ in range(20):
        q, keep = srv.polymul_request(a, b)
        t = srv.submit(q, keep)
        res = []

        def waiter():
            try:
                res.append(("ok", srv.wait(t)))
            except CfftPanic as e:
                res.append(("err", str(e)))
        th = [threading.Thread(target=waiter) for _ in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join(30)
        assert sorted(r[0] for r in res) == ["err", "ok"]
    srv.close()


def _make_device(server, keys, r):
    from cfft_amd.device import DeviceBuffer
    kind, N, data = r
    if kind == "polymul":
        a, b = (DeviceBuffer.from_host(x) for x in data)
        out = DeviceBuffer((N,), np.uint64)
        return server.polymul_request_device(a, b, out)
    g = DeviceBuffer.from_host(data)
    out = DeviceBuffer((K1, N), np.uint64)
    return server.extprod_request_device(keys[N][0], g, out)


def test_device_requests_alone_equals_mix_equals_host(cfft):
    """Device-buffer requests: alone, in a burst mixed with host-buffer
    requests of the same shapes, and against the host-buffer results (which
    equal the direct entry points, test above), bit for bit.  External
    products at N = 1024 ... 4096 read and write the callers' buffers through
    the address table; N = 512 and 8192 and the products are gathered and
    scattered."""
    from cfft_amd.serve import Server
    rng = np.random.default_rng(15)
    srv = Server()
    keys = _setup(srv, rng, sizes=(512, 1024, 2048, 4096, 8192))
    reqs = _requests(srv, keys, rng, 48)
    host = []
    for r in reqs:
        q, keep = _make(srv, keys, r)
        host.append(srv.wait(srv.submit(q, keep)).copy())
    alone = []
    for r in reqs:
        q, keep = _make_device(srv, keys, r)
        alone.append(srv.wait(srv.submit(q, keep)).download().reshape(host[len(alone)].shape))
    tickets = []
    for r in reqs:  # device and host requests interleaved
        for make in (_make_device, _make):
            q, keep = make(srv, keys, r)
            tickets.append((make is _make_device, srv.submit(q, keep)))
    for i, (dev, t) in enumerate(tickets):
        res = srv.wait(t)
        res = res.download().reshape(host[i // 2].shape) if dev else res
        assert np.array_equal(res, host[i // 2]), (i, dev)
    for i in range(len(reqs)):
        assert np.array_equal(alone[i], host[i]), i
    srv.close()


def test_device_chain_in_place(cfft):
    """A CMux-style chain: five external products on one device accumulator,
    each request reading and writing the same buffer, equal to five direct
    external products."""
    from cfft_amd import products
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.serve import Server
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(16)
    N = 2048
    srv = Server()
    keys = _setup(srv, rng, sizes=(N,))
    kid, key = keys[N]
    glwe = rng.integers(0, M, (K1, N), dtype=np.uint64)
    acc = DeviceBuffer.from_host(glwe)
    for _ in range(5):
        q, keep = srv.extprod_request_device(kid, acc, acc)
        srv.wait(srv.submit(q, keep))
    plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, 1024))
    kf = DeviceBuffer((K1 * LEVELS, K1, N // 2), np.complex128)
    products.ggsw_to_fourier(plan, DeviceBuffer.from_host(key), K1, LEVELS, kf)
    cur = DeviceBuffer.from_host(glwe[None])
    for _ in range(5):
        nxt = DeviceBuffer((1, K1, N), np.uint64)
        products.external_product(plan, kf, K1, LEVELS, BASE_LOG, cur, nxt, 1)
        cur = nxt
    assert np.array_equal(acc.download().reshape(K1, N), cur.download()[0])
    srv.close()


def test_device_request_refuses_host_memory(cfft):
    from cfft_amd import CfftPanic
    from cfft_amd._lib import Request
    from cfft_amd.serve import DEVICE_BUFFERS, POLYMUL, Server
    srv = Server()
    a = np.zeros(1024, np.uint64)
    b = np.zeros(1024, np.int64)
    out = np.zeros(1024, np.uint64)
    r = Request(POLYMUL | DEVICE_BUFFERS, -1, 1024, a.ctypes.data, b.ctypes.data, out.ctypes.data)
    with pytest.raises(CfftPanic) as ei:
        srv.submit(r, (a, b, out))
    assert ei.value.status == 10  # CFFT_ERROR_INVALID_ARGUMENT
    srv.close()


def test_large_host_batches(cfft):
    """One launch slot and 96 external products submitted at once (open-loop
    replay, all arrivals at 0): the requests queued behind the first batch
    form batches above 16, which copy through the device; every result
    equals the same request served alone."""
    from cfft_amd.serve import Server
    rng = np.random.default_rng(17)
    srv = Server(streams=1)
    keys = _setup(srv, rng, sizes=(2048,))
    reqs = [("extprod", 2048, rng.integers(0, M, (K1, 2048), dtype=np.uint64)) for _ in range(96)]
    made = [_make(srv, keys, r) for r in reqs]
    before = srv.stats()
    srv.replay([q for q, _ in made], np.zeros(len(made)))
    after = srv.stats()
    batches = after["batches"] - before["batches"]
    print("96 requests in %d batches" % batches)
    assert batches < 96 // 4
    for i, r in enumerate(reqs):
        q, keep = _make(srv, keys, r)
        alone = srv.wait(srv.submit(q, keep))
        assert np.array_equal(made[i][1][-1], alone), i
    srv.close()
