"""Benchmark: batched negacyclic polynomial products at N = 2048 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the headline configuration): C = A * B mod (X^2048 + 1) on the
2^-64 torus for a batch of 8192 independent products; A u64 torus
coefficients uniform over [0, 2^64), B signed digits uniform in [-128, 127]
stored as i64.  Inputs are synthetic and resident in HBM before timing.  One
step = one batched call of cfft_negacyclic_polymul_u64 (one kernel launch).

Printed JSON line: value = poly-mults/s of the whole job; roofline = the
fused kernel's algorithmic HBM bytes (24 N bytes per product: A and B read,
C written) per launch over its average launch time measured with HIP events
on its stream; plain = the same products issued one call per polynomial;
cpu_baseline = the oracle's restatement of the reference f64 path (one plan
per thread over the host cores) on a bounded sample.
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
BYTES_PER_PRODUCT = 24 * N  # A 8N + B 8N read, C 8N written


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


def cpu_baseline(a, b, algo, base_n, target_s=5.0):
    """Oracle restatement of the reference f64 path over all host cores: the
    whole batch once (its result is compared with the GPU's), then repeated
    passes until about target_s seconds of wall time."""
    from oracle_lib import load_oracle
    O = load_oracle()
    threads = host_cores()
    c = np.zeros((BATCH, N), dtype=np.uint64)
    total_t = O.L.or_time_polymul(N, algo, base_n, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                  BATCH, threads)
    done = BATCH
    scratch = np.zeros_like(c)
    while total_t < target_s and done < 64 * BATCH:
        total_t += O.L.or_time_polymul(N, algo, base_n, a.ctypes.data, b.ctypes.data,
                                       scratch.ctypes.data, BATCH, threads)
        done += BATCH
    return {"value": done / total_t, "unit": "poly-mults/s", "cores": threads, "kind": "port",
            "sample": "%d products (%d passes over the N=2048 batch), %.1f s wall; oracle C "
                      "restatement of unordered::Plan (scalar, bit-identical to the reference "
                      "f64 path), one plan per thread" % (done, done // BATCH, total_t)}, c, BATCH


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--algo", type=int, default=6, help="base FftAlgo (default Dif16)")
    ap.add_argument("--base-n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--plain-polys", type=int, default=512,
                    help="products timed in the one-call-per-polynomial form")
    ap.add_argument("--traffic-bytes", type=float, default=None,
                    help="HBM bytes per launch measured with rocprofv3 --pmc, if known")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build targets one GPU; --gpus %d is not supported (the batch "
              "shards independently across GPUs, see DESIGN.md)" % args.gpus, file=sys.stderr)
        return 2

    import cfft_amd
    from cfft_amd import products
    from cfft_amd.device import DeviceBuffer, Event, Stream
    from cfft_amd.unordered import FftAlgo, Method, Plan

    if cfft_amd.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    a, b = make_inputs()
    plan = Plan(N // 2, Method.UserProvided(FftAlgo(args.algo), args.base_n))
    s = Stream()
    da, db = DeviceBuffer.from_host(a, s), DeviceBuffer.from_host(b, s)
    dc = DeviceBuffer((BATCH, N), np.uint64)
    s.synchronize()

    for _ in range(args.warmup):
        products.negacyclic_polymul_u64(plan, da, db, dc, BATCH, s)
    s.synchronize()

    evs = [(Event(), Event()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        evs[i][0].record(s)
        products.negacyclic_polymul_u64(plan, da, db, dc, BATCH, s)
        evs[i][1].record(s)
    s.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [e0.elapsed_ms(e1) for e0, e1 in evs]
    avg_ms = float(np.mean(kern_ms))
    ms_per_step = wall * 1e3 / args.steps
    value = BATCH * args.steps / wall
    c_gpu = dc.download()

    # plain form: one call per polynomial, same kernel, same inputs
    npl = min(args.plain_polys, BATCH)
    s.synchronize()
    t1 = time.perf_counter()
    for i in range(npl):
        off = i * N * 8
        cfft_amd.check(cfft_amd.lib.cfft_negacyclic_polymul_u64(
            plan.handle, da.dptr + off, db.dptr + off, dc.dptr + off, 1, s.handle))
    s.synchronize()
    plain_rate = npl / (time.perf_counter() - t1)
    c_plain = dc.download()
    plain_equal = bool(np.array_equal(c_plain[:npl], c_gpu[:npl]))

    achieved = BYTES_PER_PRODUCT * BATCH / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": args.traffic_bytes,
            "kernel": "k_polymul (fused twist+2 fwd+pointwise+inv+torus)",
            "bytes_per_launch": BYTES_PER_PRODUCT * BATCH, "avg_launch_ms": round(avg_ms, 4)}

    out = {"metric": "negacyclic poly-mults/s at N=2048, batch 8192 (batched)",
           "value": round(value, 1), "unit": "poly-mults/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "roofline": roof,
           "config": {"workload": "negacyclic poly-mult N=2048 batch 8192, u64 torus x i64 "
                                  "digits in [-128,127], unordered plan %s base_n=%d"
                                  % (FftAlgo(args.algo).name, args.base_n),
                      "N": N, "batch": BATCH},
           "plain": {"value": round(plain_rate, 1), "unit": "poly-mults/s",
                     "form": "one cfft_negacyclic_polymul_u64 call per polynomial (%d polys)" % npl,
                     "batched_over_plain": round(value / plain_rate, 2),
                     "identical_results": plain_equal}}

    if not args.no_cpu:
        cpu, c_cpu, ns = cpu_baseline(a, b, args.algo, args.base_n)
        out["cpu_baseline"] = cpu
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
