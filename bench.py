"""Benchmark: batched negacyclic polynomial products at N = 2048 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode all|polymul|uneven|copy]

Workload (the headline configuration): C = A * B mod (X^2048 + 1) on the
2^-64 torus for a batch of 8192 independent products; A u64 torus
coefficients uniform over [0, 2^64), B signed digits uniform in [-128, 127]
stored as i64.  Inputs are synthetic and resident in HBM before timing.  One
step = one batched call of cfft_negacyclic_polymul_u64 (one kernel launch).

Printed JSON line: value = poly-mults/s of the whole job; roofline = the
fused kernel's algorithmic HBM bytes (24 N bytes per product: A and B read,
C written) per launch over its average launch time measured with HIP events
on its stream; plain = the same products issued one call per polynomial;
cpu_baseline = the reference f64 path restated on the host (its AVX-512 and
AVX2 SIMD forms and the scalar form, one pthread per host core, the fastest
reported) on a bounded sample.

Beside the headline, `uneven` reports the uneven batch: 20 000 f64
negacyclic round-trip jobs (fold, twist, forward, inverse, untwist) with N
drawn from {512, ..., 16384} with weights {.10, .20, .35, .20, .10, .05},
values uniform in [-1, 1), as one persistent launch replayed from a HIP graph
with autotuned per-size plans, against the fixed-schedule plain form (Dif4,
one launch per size class).  `serve` reports the request server under
open-loop Poisson arrivals of external products (k+1 = 2 GLWE polynomials of
N in {1024, 2048, 4096}, l = 3 levels of base 2^8, resident Fourier-domain
keys) at 5k, 20k and 50k requests/s: p50 / p99 latency from scheduled
arrival to result, achieved requests/s, batching.  `exact` times the exact
u64 product (both operands uniform over [0, 2^64), N = 2048 and 4096, batch
2048) through f64 transforms of 16-bit limbs against the double-double
transforms of 32-bit limbs, with both results checked equal and a sample
checked against the u64 schoolbook.  `ordered` times the ordered c64 plan at
n = 1024, batch 1 (fwd and inv per call) and reports its error against the
oracle and its round trip.  `transforms` reports batched unordered forward
transforms (n = 1024 x 8192 and n = 2048 x 4096) per second for several plans.  `matrix`
restates the reference's bench matrix (unordered fwd / inv / fwd_monomial at n = 2^8 ... 2^16,
fft128 at n = 64 ... 16384) and sweeps the fused product over N = 512 ... 16384 (also
`--mode products`).  `--mode copy` runs one
device-to-device copy of --copy-mib MiB (the reference kernel for scaling HBM
byte counters).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "concrete-fft_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

N = 2048
BATCH = 8192
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
BYTES_PER_PRODUCT = 24 * N  # A 8N + B 8N read, C 8N written
# HBM bytes per 8192-product launch of the persistent Dif16 kernel, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE over 4 launches: FETCH_SIZE x 2
# (the counter reads half of a plain copy kernel's known read size on this
# platform) + WRITE_SIZE (exact for a copy): (131391 x 2 + 131113) KiB.  The
# figure holds for the kernel sources it was measured on (their SHA-256
# below); when they change, `traffic` is reported as null until the counters
# are run again.
TRAFFIC_BYTES = (131391 * 2 + 131113) * 1024
TRAFFIC_SOURCES = ("concrete-fft_amd/csrc/polymul_n2048_kernels.hip", "concrete-fft_amd/csrc/r16_exec.hpp",
                   "concrete-fft_amd/csrc/device_math.hpp")
TRAFFIC_SHA256 = "ac50f118ad451f083dbf608844c3896354b4151d9029465d44a98b7a9a635fa4"


def traffic_sources_sha():
    import hashlib
    h = hashlib.sha256()
    for f in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def make_inputs(seed=2):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2**64, size=(BATCH, N), dtype=np.uint64)
    b = rng.integers(-128, 128, size=(BATCH, N), dtype=np.int64)
    return a, b


CPU_SERIES = (("avx512", "or_simd_time_polymul_v4", 4), ("avx2", "or_simd_time_polymul_v3", 3),
              ("scalar", "or_time_polymul", 0))


def cpu_baseline(a, b, algo, base_n, target_s=4.0):
    """The reference f64 path restated on the host, one plan shared by one
    pthread per host core: the AVX-512 (pulp V4) and AVX2 (V3) series, both
    bit-identical to the reference's golden vector, and the scalar series.
    Each runs the whole batch once (its result is compared with the GPU's),
    then repeated passes until about target_s seconds of wall time."""
    from oracle_lib import load_oracle
    O = load_oracle()
    threads = host_cores()
    series, results = {}, {}
    for name, fn, isa in CPU_SERIES:
        if isa and not O.L.or_cpu_has_isa(isa):
            series[name] = {"value": None, "note": "host CPU lacks this instruction set"}
            continue
        f = getattr(O.L, fn)
        c = np.zeros((BATCH, N), dtype=np.uint64)
        t = f(N, algo, base_n, a.ctypes.data, b.ctypes.data, c.ctypes.data, BATCH, threads)
        if t < 0:
            series[name] = {"value": None, "note": "plan not supported by this series"}
            continue
        done = BATCH
        scratch = np.zeros_like(c)
        while t < target_s and done < 1024 * BATCH:
            t += f(N, algo, base_n, a.ctypes.data, b.ctypes.data, scratch.ctypes.data, BATCH,
                   threads)
            done += BATCH
        series[name] = {"value": round(done / t, 1), "products": done, "seconds": round(t, 2)}
        results[name] = c
    best = max((k for k in results), key=lambda k: series[k]["value"])
    same = all(np.array_equal(results[k], results[best]) for k in results)
    return {"value": series[best]["value"], "unit": "poly-mults/s", "cores": threads,
            "kind": "port", "isa": best,
            "sample": "%d products (%d passes over the N=2048 batch), %.1f s wall; C "
                      "restatement of unordered::Plan's %s path (bit-identical to the "
                      "reference golden vector), one plan shared by %d pthreads"
                      % (series[best]["products"], series[best]["products"] // BATCH,
                         series[best]["seconds"], best, threads),
            "series": series, "series_identical": same}, results[best], BATCH


def cpu_transforms(algo, base_n, n=1024, target_s=2.0):
    """Unordered forward transforms per second on the host: the AVX-512 (V4)
    and AVX2 (V3) restatements of the reference plan (n = 1024, the headline
    product's transform), one Python thread per host core each calling the
    batched C routine on its own 512 transforms (ctypes releases the GIL
    during the call), for about target_s seconds per series."""
    import threading
    from oracle_lib import load_oracle
    O = load_oracle()
    threads = host_cores()
    out = {}
    for name, isa in (("avx512", 4), ("avx2", 3)):
        if not O.L.or_cpu_has_isa(isa):
            out[name] = {"value": None, "note": "host CPU lacks this instruction set"}
            continue
        f = getattr(O.L, "or_simd_unordered_fwd_batch_v%d" % isa)
        per = 512
        bufs = [np.random.default_rng(k).uniform(-1, 1, 2 * n * per) for k in range(threads)]
        counts = [0] * threads
        t_end = [0.0]

        def work(k):
            while time.perf_counter() < t_end[0]:
                if f(n, int(algo), base_n, bufs[k].ctypes.data, per) != 0:
                    return
                counts[k] += per
        t0 = time.perf_counter()
        t_end[0] = t0 + target_s
        ths = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        wall = time.perf_counter() - t0
        out[name] = {"value": round(sum(counts) / wall, 1), "transforms": sum(counts),
                     "seconds": round(wall, 2)}
    return {"unit": "transforms/s", "n": n, "cores": threads, "kind": "port",
            "form": "unordered fwd, plan %d / base %d, one thread per core" % (algo, base_n),
            "series": out}


UNEVEN_JOBS = 20000
UNEVEN_LOG2N = [9, 10, 11, 12, 13, 14]
UNEVEN_WEIGHTS = [0.10, 0.20, 0.35, 0.20, 0.10, 0.05]


def time_region(fn, stream, steps, warmup):
    """Average ms per call of `fn` over `steps` back-to-back calls, one HIP
    event pair on `stream` around them (throughput; no marker between calls)."""
    from cfft_amd.device import Event
    for _ in range(warmup):
        fn()
    stream.synchronize()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_ms(e1) / steps


def time_runs(fn, stream, steps, warmup):
    from cfft_amd.device import Event
    for _ in range(warmup):
        fn()
    stream.synchronize()
    evs = [(Event(), Event()) for _ in range(steps)]
    for i in range(steps):
        evs[i][0].record(stream)
        fn()
        evs[i][1].record(stream)
    stream.synchronize()
    return [e0.elapsed_ms(e1) for e0, e1 in evs]


def bench_uneven(steps, warmup):
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.uneven import Schedule, packed_jobs
    from cfft_amd.unordered import FftAlgo
    rng = np.random.default_rng(3)
    log2_N = rng.choice(UNEVEN_LOG2N, size=UNEVEN_JOBS, p=UNEVEN_WEIGHTS)
    jobs, nin, nf, nout = packed_jobs(log2_N)
    x = rng.uniform(-1.0, 1.0, nin)
    s = Stream()
    dx = DeviceBuffer.from_host(x, s)
    dF = DeviceBuffer((nf,), np.complex128)
    dy = DeviceBuffer((nout,), np.float64)
    s.synchronize()
    t0 = time.perf_counter()
    tuned = Schedule(jobs, nin, nf, nout, measure=True)
    tune_s = time.perf_counter() - t0
    fixed = Schedule(jobs, nin, nf, nout, measure=False, algo=FftAlgo.Dif4)
    forms = {
        "tuned_persistent_graph": lambda: tuned.run(dx, dF, dy, s),
        "tuned_persistent_direct": lambda: tuned.run_direct(dx, dF, dy, s),
        "tuned_plain": lambda: tuned.run_plain(dx, dF, dy, s),
        "fixed_persistent_graph": lambda: fixed.run(dx, dF, dy, s),
        "fixed_plain": lambda: fixed.run_plain(dx, dF, dy, s),
    }
    ms = {k: time_region(f, s, steps, warmup) for k, f in forms.items()}
    tuned.run(dx, dF, dy, s)
    s.synchronize()
    err = float(np.abs(dy.download() - x).max())
    total_N = int(np.sum(np.left_shift(1, log2_N.astype(np.int64))))
    # algorithmic bytes per job: a forward transform (read 8N, write 8N) and an
    # inverse transform (read 8N, write 8N) = 32 N; the fused kernel itself
    # moves 24 N (the inverse starts from the spectrum still in registers)
    bytes_algo = 32 * total_N
    bytes_fused = 24 * total_N
    best_form = min(("tuned_persistent_graph", "tuned_persistent_direct"), key=lambda k: ms[k])
    best = ms[best_form]
    plans = {}
    for l in UNEVEN_LOG2N:
        a, b = tuned.class_algo(l)
        plans["N=%d" % (1 << l)] = "%s/%d" % (a.name, b)
    # each size class of the same mix alone, as its own tuned schedule
    # (persistent direct launch): where the fused fraction is lost
    per_class = {}
    for l in UNEVEN_LOG2N:
        sel = log2_N[log2_N == l]
        cj, cin, cf, cout = packed_jobs(sel)
        cx = DeviceBuffer.from_host(x[:cin], s)  # a class alone packs into fewer words
        cF = DeviceBuffer((cf,), np.complex128)
        cy = DeviceBuffer((cout,), np.float64)
        cs = Schedule(cj, cin, cf, cout, measure=True)
        cms = time_region(lambda: cs.run_direct(cx, cF, cy, s), s, max(steps, 10), warmup)
        cN = 1 << int(l)
        per_class["N=%d" % cN] = {
            "jobs": int(len(sel)), "us": round(cms * 1e3, 1),
            "fused_frac": round(24 * cN * len(sel) / (cms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    return {
        "workload": "%d f64 negacyclic fwd+inv jobs, N in {512..16384} weights %s, %d total "
                    "coefficients" % (UNEVEN_JOBS, UNEVEN_WEIGHTS, total_N),
        "form": best_form, "tile_schedule": tuned.tile_schedule(),
        "jobs_per_s": round(UNEVEN_JOBS / (best * 1e-3), 1),
        "transforms_per_s": round(2 * UNEVEN_JOBS / (best * 1e-3), 1),
        "ms": {k: round(v, 4) for k, v in ms.items()},
        "tuned_over_fixed_plain": round(ms["fixed_plain"] / best, 3),
        "roofline": {"bound": "hbm", "achieved": round(bytes_algo / (best * 1e-3) / 1e9, 2),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(bytes_algo / (best * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                     "bytes_per_launch": bytes_algo, "fused_bytes_per_launch": bytes_fused,
                     "fused_frac": round(bytes_fused / (best * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)},
        "tuned_plans": plans, "per_class_alone": per_class, "autotune_s": round(tune_s, 3),
        "roundtrip_error_max": err, "schedule": tuned.info()}


SERVE_RATES = [5000, 20000, 50000]
SERVE_SIZES = [1024, 2048, 4096]


def bench_serve(duration_s, rates=SERVE_RATES):
    """Open-loop Poisson replay of external-product requests (k+1 = 2, l = 3,
    base 2^8, N uniform over SERVE_SIZES) against resident Fourier keys, in
    two forms: host-buffer requests (inputs and results cross PCIe) and
    device-buffer requests (CFFT_REQUEST_DEVICE_BUFFERS: inputs read from and
    results written to device buffers, the form a CMux chain uses)."""
    from cfft_amd import check, lib
    from cfft_amd._lib import Request
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.serve import DEVICE_BUFFERS, EXTERNAL_PRODUCT, Server
    rng = np.random.default_rng(5)
    srv = Server()
    keys = {}
    for N in SERVE_SIZES:
        key = rng.integers(0, 2**64, (6, 2, N), dtype=np.uint64)
        keys[N] = srv.add_key(key, 2, 3, 8)
    glwes = {N: rng.integers(0, 2**64, (64, 2, N), dtype=np.uint64) for N in SERVE_SIZES}
    dglwes = {N: DeviceBuffer.from_host(glwes[N]) for N in SERVE_SIZES}
    # device results go to an arena of ARENA slots used round-robin: request
    # i and i + ARENA are 80 ms apart at 50k req/s, never in flight together
    ARENA = 4096
    slot_bytes = 2 * max(SERVE_SIZES) * 8
    dout = DeviceBuffer((ARENA, slot_bytes // 8), np.uint64)

    def trace(rate, count, dev):
        Ns = [int(N) for N in rng.choice(SERVE_SIZES, count)]
        if dev:
            made = [(Request(EXTERNAL_PRODUCT | DEVICE_BUFFERS, keys[N], N,
                             dglwes[N].dptr + (i % 64) * 2 * N * 8, None,
                             dout.dptr + (i % ARENA) * slot_bytes), None)
                    for i, N in enumerate(Ns)]
        else:
            made = [srv.extprod_request(keys[N], glwes[N][i % 64]) for i, N in enumerate(Ns)]
        arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
        return made, arrivals

    def result(made, i, dev):
        if not dev:
            return made[i][1][-1]
        N = int(made[i][0].N)
        h = np.empty((2, N), np.uint64)
        check(lib.cfft_memcpy_dtoh(h.ctypes.data, dout.dptr + (i % ARENA) * slot_bytes,
                                   2 * N * 8, None))
        return h

    def run(dev):
        # warm-up at every rate: the first passes of a process show one-time
        # stalls of 10-20 ms (runtime resources growing with the number of
        # batches in flight); steady-state serving is what is measured
        for rate in rates:
            made, arr = trace(rate, max(200, int(rate * 0.2)), dev)
            srv.replay([q for q, _ in made], arr)
        out = []
        for rate in rates:
            count = max(200, int(rate * duration_s))
            made, arr = trace(rate, count, dev)
            before = srv.stats()
            lat = srv.replay([q for q, _ in made], arr)
            after = srv.stats()
            span = float(arr[-1] + lat[-1])
            # results of the mix against the same request alone (host form),
            # on a sample of the requests whose result slot is still theirs
            same = True
            for i in rng.choice(np.arange(max(0, count - ARENA), count), 8, replace=False):
                i = int(i)
                N = int(made[i][0].N)
                q, keep = srv.extprod_request(keys[N], glwes[N][i % 64])
                alone = srv.wait(srv.submit(q, keep))
                same &= bool(np.array_equal(alone, result(made, i, dev)))
            nb = after["batches"] - before["batches"]
            out.append({"offered_req_per_s": rate, "requests": count,
                        "achieved_req_per_s": round(count / span, 1),
                        "p50_us": round(float(np.percentile(lat, 50)) * 1e6, 1),
                        "p99_us": round(float(np.percentile(lat, 99)) * 1e6, 1),
                        "max_us": round(float(lat.max()) * 1e6, 1),
                        "batches": nb, "mean_batch": round(count / max(nb, 1), 2),
                        "max_inflight": after["max_inflight"],
                        "alone_equals_mix": same})
        return out

    host = run(False)
    device = run(True)
    st = srv.stats()
    srv.close()
    return {"workload": "external products, GLWE k+1=2 of N in %s, l=3, base 2^8, resident "
                        "Fourier keys, Poisson arrivals, %.2f s per rate" % (SERVE_SIZES, duration_s),
            "rates": host, "device_buffer_rates": device,
            "hbm_budget_bytes": st["hbm_budget_bytes"], "streams": st["streams"]}


EXACT_SIZES = [2048, 4096]
EXACT_BATCH = 2048


def bench_exact(steps, warmup):
    """Exact u64 products, both operands uniform over [0, 2^64): the f64
    16-bit-limb path against the double-double 32-bit-limb path, same inputs."""
    from oracle_lib import load_oracle
    from cfft_amd import fft128, products
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(4)
    s = Stream()
    out = []
    for N in EXACT_SIZES:
        a = rng.integers(0, 2**64, (EXACT_BATCH, N), dtype=np.uint64)
        b = rng.integers(0, 2**64, (EXACT_BATCH, N), dtype=np.uint64)
        da, db = DeviceBuffer.from_host(a, s), DeviceBuffer.from_host(b, s)
        c1, c2 = DeviceBuffer((EXACT_BATCH, N), np.uint64), DeviceBuffer((EXACT_BATCH, N), np.uint64)
        m1, m2 = DeviceBuffer((EXACT_BATCH,), np.float64), DeviceBuffer((EXACT_BATCH,), np.float64)
        p64 = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, min(N // 2, 1024)))
        p128 = fft128.Plan(N // 2)
        ms64 = time_region(lambda: products.exact_polymul_u64(p64, da, db, c1, EXACT_BATCH, m1, s),
                           s, steps, warmup)
        ms128 = time_region(lambda: products.exact128_polymul_u64(p128, da, db, c2, EXACT_BATCH,
                                                                   m2, s), s, steps, warmup)
        r1, r2 = c1.download(), c2.download()
        k = 4
        exact = load_oracle().schoolbook_u64(a[:k], b[:k], threads=host_cores())
        bytes_algo = 24 * N * EXACT_BATCH
        # FP64 flop model per product (n = N/2): 12 transforms (8 forward limb
        # transforms, 4 inverse, one per weight) at 5 n log2 n, 10 complex
        # multiply-adds per point (8 flop) for the weights, 12 twists (6n)
        n = N // 2
        flop = 12 * 5 * n * int(np.log2(n)) + 80 * n + 72 * n
        out.append({"N": N, "batch": EXACT_BATCH,
                    "f64_limb16": {"products_per_s": round(EXACT_BATCH / (ms64 * 1e-3), 1),
                                   "ms": round(ms64, 4),
                                   "hbm_frac": round(bytes_algo / (ms64 * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                   "fp64_tflops_model": round(flop * EXACT_BATCH / (ms64 * 1e-3) / 1e12, 2),
                                   "fp64_frac_model": round(flop * EXACT_BATCH / (ms64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
                                   "max_margin": float(m1.download().max())},
                    "fft128_limb32": {"products_per_s": round(EXACT_BATCH / (ms128 * 1e-3), 1),
                                      "ms": round(ms128, 4),
                                      "max_margin": float(m2.download().max())},
                    "f64_over_fft128": round(ms128 / ms64, 3),
                    "paths_identical": bool(np.array_equal(r1, r2)),
                    "bit_exact_vs_schoolbook": bool(np.array_equal(r1[:k], exact)),
                    "schoolbook_sample": k})
    return out


def bench_transforms(steps=50):
    """Batched unordered forward transforms in place on device, 8 Mi points
    per call: n = 512 x 16384, n = 1024 (N = 2048) x 8192 and n = 2048
    (N = 4096) x 4096;
    transforms/s and HBM fraction at 32 n bytes each."""
    from cfft_amd import check, lib
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    rng = np.random.default_rng(6)
    s = Stream()
    x = DeviceBuffer.from_host(rng.uniform(-1, 1, (8192, 2048)).view(np.complex128), s)
    out = {}
    for n, algo, base in ((512, FftAlgo.Dif16, 512),
                          (1024, FftAlgo.Dif16, 1024), (1024, FftAlgo.Dif8, 1024),
                          (1024, FftAlgo.Dit16, 1024), (1024, FftAlgo.Dif16, 512),
                          (2048, FftAlgo.Dif16, 1024), (2048, FftAlgo.Dif8, 1024)):
        batch = 8192 * 1024 // n
        plan = Plan(n, Method.UserProvided(algo, base))
        f = lambda: check(lib.cfft_fwd(plan.handle, x.dptr, batch, n, s.handle))
        ms = time_region(f, s, steps, 10)
        gbs = 32 * n * batch / (ms * 1e-3) / 1e9
        out["n=%d %s/%d" % (n, algo.name, base)] = {
            "transforms_per_s": round(batch / (ms * 1e-3), 1), "ms": round(ms, 4),
            "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}
    return {"workload": "unordered fwd in place on device, 8 Mi complex points per call",
            "plans": out}


def bench_ordered(steps=200):
    """The ordered c64 plan at n = 1024, batch 1 (latency-bound): fwd then inv
    then 1/n on re, im uniform in [-1, 1) (seed 1); per-call device time, rel
    L2 error of fwd against the oracle's restatement of the same plan, round
    trip error."""
    from oracle_lib import load_oracle
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.ordered import FftAlgo, Method, Plan
    n = 1024
    rng = np.random.default_rng(1)
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex128)
    s = Stream()
    out = {}
    for algo in (FftAlgo.Dif16, FftAlgo.Dif4):
        plan = Plan(n, Method.UserProvided(algo))
        d = DeviceBuffer.from_host(x, s)
        ms_f = float(np.median(time_runs(lambda: plan.fwd(d, s), s, steps, 10)))
        ms_i = float(np.median(time_runs(lambda: plan.inv(d, s), s, steps, 10)))
        y = x.copy()
        plan.fwd(y)
        ref = load_oracle().ordered_fwd(x, int(algo))
        rel = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
        plan.inv(y)
        rt = float(np.abs(y / n - x).max())
        out[algo.name] = {"fwd_us": round(ms_f * 1e3, 2), "inv_us": round(ms_i * 1e3, 2),
                          "rel_l2_vs_oracle": rel, "bit_exact_vs_oracle": rel == 0.0,
                          "roundtrip_max_err": rt}
    return {"workload": "ordered c64 plan, n = 1024, batch 1, fwd / inv per call (device "
                        "buffer, HIP events)", "plans": out}


def bench_host_path(calls=2000):
    """The drop-in host path: unordered::Plan::fwd / inv on a host buffer of
    one n = 1024 polynomial per call (src/unordered.rs:826-839, the call a
    TFHE-rs caller makes per polynomial), wall-clock latency per call (p50,
    p99) through cfft_fwd_host / cfft_inv_host, checked bit-equal to the
    device-buffer path; beside it the oracle's SIMD restatement of the same
    plan on one host core (per transform, from a batch of 4096)."""
    from oracle_lib import load_oracle
    from cfft_amd import check, lib
    from cfft_amd.device import DeviceBuffer
    from cfft_amd.unordered import FftAlgo, Method, Plan
    n, algo, base = 1024, FftAlgo.Dif16, 1024
    plan = Plan(n, Method.UserProvided(algo, base))
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    out = {}
    for name, f in (("fwd", lib.cfft_fwd_host), ("inv", lib.cfft_inv_host)):
        buf = x.copy()
        ptr_ = buf.ctypes.data
        for _ in range(100):
            buf[:] = x
            check(f(plan.handle, ptr_, 1))
        lat = []
        for _ in range(calls):
            t0 = time.perf_counter()
            check(f(plan.handle, ptr_, 1))
            lat.append(time.perf_counter() - t0)
        buf[:] = x
        check(f(plan.handle, ptr_, 1))
        d = DeviceBuffer.from_host(x)
        (plan.fwd if name == "fwd" else plan.inv)(d)
        lat = np.array(lat) * 1e6
        out[name] = {"p50_us": round(float(np.percentile(lat, 50)), 2),
                     "p99_us": round(float(np.percentile(lat, 99)), 2),
                     "calls": calls,
                     "equals_device_path": bool(np.array_equal(buf.view(np.uint64),
                                                               d.download().view(np.uint64)))}
    o = load_oracle()
    isa = 4 if o.L.or_cpu_has_isa(4) else 3
    cpu = {}
    for name, fwd in (("fwd", True), ("inv", False)):
        z = np.tile(x, (4096, 1))
        t0 = time.perf_counter()
        o.simd_unordered(isa, z, int(algo), base, fwd=fwd)
        cpu[name + "_us"] = round((time.perf_counter() - t0) / 4096 * 1e6, 3)
    return {"workload": "unordered n=1024 Dif16/1024, host buffer, batch 1 per call",
            "gpu": out, "cpu_one_core": dict(cpu, isa="avx512" if isa == 4 else "avx2",
                                             kind="port")}


def bench_matrix(steps=100):
    """The reference's bench matrix (benches/fft.rs:116-207, :211-338) on
    device buffers: for n = 2^8 .. 2^16, the unordered plan chosen by
    Method::Measure(10 ms) -- forward, inverse and fwd_monomial at batch 1
    (device time per call, median of `steps` HIP event pairs) and forward /
    inverse batched over 8 Mi points per call (transforms/s, fraction of
    8 TB/s at 32 n bytes each) -- and the ordered plan (n <= 1024, Measure,
    batch-1 forward); for fft128 n = 64 .. 16384, the negacyclic forward and
    inverse (batch 1 and batched over 4 Mi points)."""
    import cfft_amd
    from cfft_amd import check, fft128, lib, ordered
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    s = Stream()
    rng = np.random.default_rng(9)
    big = DeviceBuffer.from_host(rng.uniform(-1, 1, 1 << 24).view(np.complex128), s)
    rows = {}
    for lg in range(8, 17):
        n = 1 << lg
        plan = Plan(n, Method.Measure(0.01))
        algo, base = plan.algo()
        one = DeviceBuffer.from_host(rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n), s)
        row = {"plan": "%s/%d" % (algo.name, base)}
        for name, f in (("fwd", lib.cfft_fwd), ("inv", lib.cfft_inv)):
            ms = time_runs(lambda: check(f(plan.handle, one.dptr, 1, n, s.handle)), s, steps, 10)
            row[name + "_us"] = round(float(np.median(ms)) * 1e3, 2)
            batch = (1 << 23) // n
            msb = time_region(lambda: check(f(plan.handle, big.dptr, batch, n, s.handle)), s,
                              20, 3)
            row[name + "_batched_per_s"] = round(batch / (msb * 1e-3), 1)
            row[name + "_batched_hbm_frac"] = round(32 * n * batch / (msb * 1e-3) / 1e9
                                                    / PEAK_HBM_GBS, 4)
        deg = [0]

        def mono():
            deg[0] = (deg[0] + 1) % n
            check(lib.cfft_fwd_monomial(plan.handle, deg[0], one.dptr, s.handle))
        ms = time_runs(mono, s, steps, 10)
        row["fwd_monomial_us"] = round(float(np.median(ms)) * 1e3, 2)
        if n <= 1024:
            op = ordered.Plan(n, ordered.Method.Measure(0.01))
            ms = time_runs(lambda: op.fwd(one, s), s, steps, 10)
            row["ordered_fwd_us"] = round(float(np.median(ms)) * 1e3, 2)
            row["ordered_plan"] = op.algo().name
        rows[str(n)] = row
    r128 = {}
    for lg in range(6, 15):
        n = 1 << lg
        row = {}
        tw = [DeviceBuffer.from_host(t, s) for t in fft128.init_negacyclic_twiddles(n)]
        d1 = [DeviceBuffer.from_host(rng.standard_normal(n), s) for _ in range(4)]
        try:
            fft128.negacyclic_fwd_fft(*d1, tw, 1, 0, s)
        except cfft_amd.CfftPanic as e:  # sizes the device transforms do not take
            r128[str(n)] = {"unsupported": str(e).splitlines()[0]}
            continue
        batch = (1 << 22) // n
        db = [DeviceBuffer.from_host(rng.standard_normal(batch * n), s) for _ in range(4)]
        for name, f in (("fwd", fft128.negacyclic_fwd_fft), ("inv", fft128.negacyclic_inv_fft)):
            ms = time_runs(lambda: f(*d1, tw, 1, 0, s), s, steps, 10)
            row[name + "_us"] = round(float(np.median(ms)) * 1e3, 2)
            msb = time_region(lambda: f(*db, tw, batch, 0, s), s, 20, 3)
            row[name + "_batched_per_s"] = round(batch / (msb * 1e-3), 1)
            row[name + "_batched_hbm_frac"] = round(64 * n * batch / (msb * 1e-3) / 1e9
                                                    / PEAK_HBM_GBS, 4)
        r128[str(n)] = row
    return {"workload": "reference bench matrix on device buffers (batch-1 device time per "
                        "call; batched: 8 Mi c64 / 4 Mi f128 complex points per call); fused "
                        "products: 16 Mi coefficients per operand per call",
            "unordered": rows, "fft128": r128, "products": bench_products(),
            "plans": bench_plans()}


def bench_products(steps=20):
    """The fused negacyclic product across N = 512 ... 16384 (the caller-side
    path of src/unordered.rs:826-940 at other sizes): Dif16 plans with base_n =
    min(N/2, 1024), batch = 16 Mi / N (each operand 128 MiB), products/s and
    the fraction of 8 TB/s at 24 N bytes per product."""
    from cfft_amd import check, lib
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    s = Stream()
    rng = np.random.default_rng(10)
    prods = {}
    for lg in range(9, 15):
        N = 1 << lg
        plan = Plan(N // 2, Method.UserProvided(FftAlgo.Dif16, min(N // 2, 1024)))
        batch = (1 << 24) // N
        a = DeviceBuffer.from_host(rng.integers(0, 1 << 64, batch * N, dtype=np.uint64), s)
        b = DeviceBuffer.from_host(rng.integers(-128, 128, batch * N, dtype=np.int64), s)
        c = DeviceBuffer((batch * N,), np.uint64)
        msb = time_region(lambda: check(lib.cfft_negacyclic_polymul_u64(
            plan.handle, a.dptr, b.dptr, c.dptr, batch, s.handle)), s, steps, 3)
        prods[str(N)] = {"batch": batch, "ms": round(msb, 4),
                         "products_per_s": round(batch / (msb * 1e-3), 1),
                         "hbm_frac": round(24 * N * batch / (msb * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
        del a, b, c
    return prods


def bench_plans(steps=20):
    """Every f64 plan the autotuner considers at N = 2048 (8 algorithms x
    base_n 512 / 1024): fused products/s at batch 8192 and batched n = 1024
    forward / inverse transforms/s at batch 8192, each with the kernel kind
    the plan runs on (register-resident, compile-time pass program)."""
    from cfft_amd import check, lib
    from cfft_amd.device import DeviceBuffer, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan
    s = Stream()
    rng = np.random.default_rng(12)
    N, batch = 2048, 8192
    n = N // 2
    a = DeviceBuffer.from_host(rng.integers(0, 1 << 64, batch * N, dtype=np.uint64), s)
    b = DeviceBuffer.from_host(rng.integers(-128, 128, batch * N, dtype=np.int64), s)
    c = DeviceBuffer((batch * N,), np.uint64)
    x = DeviceBuffer.from_host(rng.uniform(-1, 1, (batch, 2 * n)).view(np.complex128), s)
    register = {("Dif16", 1024), ("Dit16", 1024), ("Dif8", 1024), ("Dif16", 512)}
    out = {}
    for base in (1024, 512):
        for algo in FftAlgo:
            plan = Plan(n, Method.UserProvided(algo, base))
            mp = time_region(lambda: check(lib.cfft_negacyclic_polymul_u64(
                plan.handle, a.dptr, b.dptr, c.dptr, batch, s.handle)), s, steps, 3)
            mf = time_region(lambda: check(lib.cfft_fwd(plan.handle, x.dptr, batch, n, s.handle)),
                             s, steps, 3)
            mi = time_region(lambda: check(lib.cfft_inv(plan.handle, x.dptr, batch, n, s.handle)),
                             s, steps, 3)
            out["%s/%d" % (algo.name, base)] = {
                "kernel": "register-resident" if (algo.name, base) in register
                          else "compile-time pass program",
                "products_per_s": round(batch / (mp * 1e-3), 1),
                "fwd_per_s": round(batch / (mf * 1e-3), 1),
                "inv_per_s": round(batch / (mi * 1e-3), 1)}
    return {"workload": "N = 2048 products and n = 1024 transforms, batch 8192 each",
            "plans": out}


def copy_probe(mib):
    from cfft_amd import check, lib
    from cfft_amd.device import DeviceBuffer, Stream
    nbytes = mib << 20
    src = DeviceBuffer((nbytes // 8,), np.uint64)
    dst = DeviceBuffer((nbytes // 8,), np.uint64)
    s = Stream()
    ms = time_runs(lambda: check(lib.cfft_memcpy_dtod(dst.dptr, src.dptr, nbytes, s.handle)), s,
                   1, 1)
    return {"mode": "copy", "bytes": nbytes, "ms": round(ms[0], 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: enough warm-up launches (10 ms) for the GPU clocks to settle
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--algo", type=int, default=6, help="base FftAlgo (default Dif16)")
    ap.add_argument("--base-n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--plain-polys", type=int, default=512,
                    help="products timed in the one-call-per-polynomial form")
    ap.add_argument("--traffic-bytes", type=float, default=TRAFFIC_BYTES,
                    help="HBM bytes per launch of the product kernel (rocprofv3 --pmc)")
    ap.add_argument("--mode", choices=["all", "polymul", "uneven", "serve", "exact", "ordered", "products",
                                       "plans", "transforms", "matrix", "host", "copy"],
                    default="all")
    ap.add_argument("--serve-seconds", type=float, default=0.5)
    ap.add_argument("--uneven-steps", type=int, default=20)
    ap.add_argument("--exact-steps", type=int, default=50)
    ap.add_argument("--exact-warmup", type=int, default=10)
    ap.add_argument("--copy-mib", type=int, default=128)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported (products are "
              "independent, so a batch would shard across GPUs with no exchange)" % args.gpus,
              file=sys.stderr)
        return 2

    import cfft_amd
    from cfft_amd import products
    from cfft_amd.device import DeviceBuffer, Event, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan

    if cfft_amd.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    if args.mode == "copy":
        print(json.dumps(copy_probe(args.copy_mib)))
        return 0
    if args.mode == "uneven":
        print(json.dumps({"mode": "uneven", "uneven": bench_uneven(args.uneven_steps, 2)}))
        return 0
    if args.mode == "serve":
        print(json.dumps({"mode": "serve", "serve": bench_serve(args.serve_seconds)}))
        return 0
    if args.mode == "exact":
        print(json.dumps({"mode": "exact",
                          "exact": bench_exact(args.exact_steps, args.exact_warmup)}))
        return 0
    if args.mode == "ordered":
        print(json.dumps({"mode": "ordered", "ordered": bench_ordered()}))
        return 0
    if args.mode == "transforms":
        print(json.dumps({"mode": "transforms", "transforms": bench_transforms()}))
        return 0
    if args.mode == "matrix":
        print(json.dumps({"mode": "matrix", "matrix": bench_matrix()}))
        return 0
    if args.mode == "products":
        print(json.dumps({"mode": "products", "products": bench_products()}))
        return 0
    if args.mode == "plans":
        print(json.dumps({"mode": "plans", "plans": bench_plans()}))
        return 0
    if args.mode == "host":
        print(json.dumps({"mode": "host", "host_path": bench_host_path()}))
        return 0
    a, b = make_inputs()
    plan = Plan(N // 2, Method.UserProvided(FftAlgo(args.algo), args.base_n))
    s = Stream()
    da, db = DeviceBuffer.from_host(a, s), DeviceBuffer.from_host(b, s)
    dc = DeviceBuffer((BATCH, N), np.uint64)
    s.synchronize()

    for _ in range(args.warmup):
        products.negacyclic_polymul_u64(plan, da, db, dc, BATCH, s)
    s.synchronize()

    # one HIP event pair on the launch stream brackets the timed region (a
    # timing marker between every two launches would serialise them and
    # count its own cost): average launch duration = elapsed / K
    ev0, ev1 = Event(), Event()
    t0 = time.perf_counter()
    ev0.record(s)
    for i in range(args.steps):
        products.negacyclic_polymul_u64(plan, da, db, dc, BATCH, s)
    ev1.record(s)
    s.synchronize()
    wall = time.perf_counter() - t0
    avg_ms = ev0.elapsed_ms(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = BATCH * args.steps / wall
    c_gpu = dc.download()

    # plain form: one call per polynomial, same kernel, same inputs
    npl = min(args.plain_polys, BATCH)
    plain = None
    if npl > 0:
        s.synchronize()
        t1 = time.perf_counter()
        for i in range(npl):
            off = i * N * 8
            cfft_amd.check(cfft_amd.lib.cfft_negacyclic_polymul_u64(
                plan.handle, da.dptr + off, db.dptr + off, dc.dptr + off, 1, s.handle))
        s.synchronize()
        plain_rate = npl / (time.perf_counter() - t1)
        c_plain = dc.download()
        plain = {"value": round(plain_rate, 1), "unit": "poly-mults/s",
                 "form": "one cfft_negacyclic_polymul_u64 call per polynomial (%d polys)" % npl,
                 "batched_over_plain": round(value / plain_rate, 2),
                 "identical_results": bool(np.array_equal(c_plain[:npl], c_gpu[:npl]))}

    achieved = BYTES_PER_PRODUCT * BATCH / (avg_ms * 1e-3) / 1e9
    headline_plan = (args.algo, args.base_n) == (6, 1024)
    fresh = headline_plan and traffic_sources_sha() == TRAFFIC_SHA256
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": args.traffic_bytes if fresh else None,
            "kernel": ("k_r16_polymul_persistent (fused twist + 2 fwd + pointwise + inv + torus)"
                       if headline_plan
                       else "fused product kernel of the %s/%d plan"
                       % (FftAlgo(args.algo).name, args.base_n)),
            "bytes_per_launch": BYTES_PER_PRODUCT * BATCH, "avg_launch_ms": round(avg_ms, 4)}
    if not fresh:
        roof["traffic_note"] = ("kernel sources changed since the HBM counter run: traffic not "
                                "re-measured" if headline_plan else
                                "HBM counters were read for the Dif16/1024 kernel only")

    out = {"metric": "negacyclic poly-mults/s at N=2048, batch 8192 (batched)",
           "value": round(value, 1), "unit": "poly-mults/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "roofline": roof,
           "config": {"workload": "negacyclic poly-mult N=2048 batch 8192, u64 torus x i64 "
                                  "digits in [-128,127], unordered plan %s base_n=%d"
                                  % (FftAlgo(args.algo).name, args.base_n),
                      "N": N, "batch": BATCH},
           "plain": plain}

    if args.mode == "all":
        out["uneven"] = bench_uneven(args.uneven_steps, 2)
        out["serve"] = bench_serve(args.serve_seconds)
        out["exact"] = bench_exact(args.exact_steps, args.exact_warmup)
        out["ordered"] = bench_ordered()
        out["transforms"] = bench_transforms()
        out["matrix"] = bench_matrix()
        out["host_path"] = bench_host_path()

    if not args.no_cpu and args.mode == "all":
        cpu, c_cpu, ns = cpu_baseline(a, b, args.algo, args.base_n)
        cpu["transforms"] = cpu_transforms(args.algo, args.base_n)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
        out["gpu_equals_cpu_bitwise"] = bool(np.array_equal(c_cpu, c_gpu[:ns]))
        # torus error of both against the exact schoolbook on a few products
        from oracle_lib import load_oracle
        k = 16
        exact = load_oracle().schoolbook_u64(a[:k], b[:k].view(np.uint64), threads=host_cores())
        e_gpu = int(np.abs((c_gpu[:k] - exact).view(np.int64)).max())
        e_cpu = int(np.abs((c_cpu[:k] - exact).view(np.int64)).max())
        out["torus_error_max"] = {"gpu": e_gpu, "cpu": e_cpu, "unit": "2^-64",
                                  "log2_gpu": round(float(np.log2(max(e_gpu, 1))), 2),
                                  "sample": "%d products vs u64 schoolbook" % k}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
