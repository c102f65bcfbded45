"""Build libcfft_amd.so (gfx950 HIP kernels + C ABI) in-tree.

    python concrete-fft_amd/build.py            # build lib/libcfft_amd.so

Device code: hipcc --offload-arch=gfx950 with -ffp-contract=off (the
transforms reproduce the reference's rounding bit for bit, so no implicit
contraction) and packed FP32 disabled.  Host-only translation units that
need binary128 (libquadmath) are compiled with g++.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcfft_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CFFT_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]

HIP_SOURCES = ["fft_kernels.hip", "polymul_kernels.hip", "polymul_n2048_kernels.hip", "fft128_kernels.hip",
               "exact_kernels.hip", "extprod_kernels.hip", "uneven_kernels.hip", "noalloc.hip", "serve_kernels.hip",
               "large_transform_kernels.hip", "dit16_kernels.hip", "fixed_n256.hip", "fixed_n512.hip",
               "fixed_n1024.hip", "fixed_n2048.hip", "fixed_n4096.hip",
               "fixed_n8192.hip"]
HOST_HIP_SOURCES = ["api.cpp", "fft128_api.cpp", "products_api.cpp", "uneven_api.cpp",
                    "serve.cpp", "scratch.cpp", "host_stage.cpp"]
GXX_SOURCES = ["host_tables.cpp", "twiddle128.cpp", "tune_cache.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        p = os.path.join(CSRC, s)
        if os.path.exists(p):
            jobs.append([HIPCC] + COMMON + HIP_FLAGS + ["-c", p, "-o", _obj(s)])
    for s in HOST_HIP_SOURCES:
        p = os.path.join(CSRC, s)
        if os.path.exists(p):
            jobs.append([HIPCC] + COMMON + ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", p,
                                            "-o", _obj(s)])
    for s in GXX_SOURCES:
        p = os.path.join(CSRC, s)
        jobs.append(["g++"] + COMMON + ["-std=gnu++17", "-fext-numeric-literals", "-I/opt/rocm/include", "-c", p, "-o", _obj(s)])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lquadmath", "-lpthread",
          "-Wl,--version-script=" + os.path.join(CSRC, "exports.map")])
    os.replace(tmp, LIB)
    # C++ mirror example (links the library through its in-tree rpath)
    host = os.path.join(HERE, "host")
    ex = os.path.join(LIBDIR, "example_roundtrip")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + host,
          os.path.join(host, "example_roundtrip.cpp"), "-o", ex + ".tmp", "-L" + LIBDIR,
          "-lcfft_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(ex + ".tmp", ex)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
