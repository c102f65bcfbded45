"""ctypes loader of oracle/liboracle.so (the CPU restatement used as checker).

Test infrastructure: only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

c_size = ctypes.c_size_t
c_vp = ctypes.c_void_p
c_int = ctypes.c_int


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        self.L = L
        sig = {
            "or_sincospi64": (None, [ctypes.c_double, c_vp, c_vp]),
            "or_init_wt": (None, [c_int, c_size, c_vp, c_vp]),
            "or_unordered_fwd_batch": (c_int, [c_size, c_int, c_size, c_vp, c_size]),
            "or_unordered_inv_batch": (c_int, [c_size, c_int, c_size, c_vp, c_size]),
            "or_ordered_fwd_batch": (c_int, [c_size, c_int, c_vp, c_size]),
            "or_ordered_inv_batch": (c_int, [c_size, c_int, c_vp, c_size]),
            "or_ordered_fwd_batch_large": (c_int, [c_size, c_int, c_vp, c_size]),
            "or_ordered_inv_batch_large": (c_int, [c_size, c_int, c_vp, c_size]),
            "or_stdrng_fill_f64_01": (None, [ctypes.c_uint64, c_vp, c_size]),
            "or_bit_rev_twice": (c_size, [ctypes.c_uint, ctypes.c_uint, c_size]),
            "or_bit_rev_twice_inv": (c_size, [ctypes.c_uint, ctypes.c_uint, c_size]),
            "or_twist_table": (None, [c_size, c_vp]),
            "or_negacyclic_polymul_f64": (c_int, [c_size, c_int, c_size, c_vp, c_vp, c_vp, c_size, c_int]),
            "or_schoolbook_negacyclic_u64": (None, [c_size, c_vp, c_vp, c_vp]),
            "or_schoolbook_negacyclic_u64_batch": (None, [c_size, c_vp, c_vp, c_vp, c_size, c_int]),
            "or_from_torus": (ctypes.c_uint64, [ctypes.c_double]),
            "or_fft128_init_twiddles": (None, [c_size, c_vp, c_vp, c_vp, c_vp]),
            "or_fft128_fwd": (None, [c_size, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int]),
            "or_fft128_inv": (None, [c_size, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int]),
            "or_gadget_decompose": (None, [c_size, c_vp, c_int, c_int, c_vp]),
            "or_negacyclic_roundtrip_f64": (c_int, [c_size, c_int, c_size, c_vp, c_vp, c_vp, c_size]),
            "or_torus_to_fourier": (c_int, [c_size, c_int, c_size, c_vp, c_vp, c_size]),
            "or_external_product": (c_int, [c_size, c_int, c_int, c_int, c_vp, c_int, c_size, c_vp, c_vp]),
            "or_time_polymul": (ctypes.c_double, [c_size, c_int, c_size, c_vp, c_vp, c_vp, c_size, c_int]),
            "or_cpu_has_isa": (c_int, [c_int]),
            "or_simd_unordered_fwd_batch_v3": (c_int, [c_size, c_int, c_size, c_vp, c_size]),
            "or_simd_unordered_inv_batch_v3": (c_int, [c_size, c_int, c_size, c_vp, c_size]),
            "or_simd_polymul_f64_v3": (c_int, [c_size, c_int, c_size, c_vp, c_vp, c_vp, c_size, c_int]),
            "or_simd_time_polymul_v3": (ctypes.c_double, [c_size, c_int, c_size, c_vp, c_vp, c_vp, c_size, c_int]),
            "or_simd_unordered_fwd_batch_v4": (c_int, [c_size, c_int, c_size, c_vp, c_size]),
            "or_simd_unordered_inv_batch_v4": (c_int, [c_size, c_int, c_size, c_vp, c_size]),
            "or_simd_polymul_f64_v4": (c_int, [c_size, c_int, c_size, c_vp, c_vp, c_vp, c_size, c_int]),
            "or_simd_time_polymul_v4": (ctypes.c_double, [c_size, c_int, c_size, c_vp, c_vp, c_vp, c_size, c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # f128 ops take/return a struct of two doubles
        class F128(ctypes.Structure):
            _fields_ = [("hi", ctypes.c_double), ("lo", ctypes.c_double)]
        self.F128 = F128
        for name in ("or_f128_add", "or_f128_sub", "or_f128_add_estimate", "or_f128_sub_estimate",
                     "or_f128_mul", "or_f128_mul_fma", "or_f128_div"):
            f = getattr(L, name)
            f.restype = F128
            f.argtypes = [F128, F128]
        L.or_f128_sincospi.restype = None
        L.or_f128_sincospi.argtypes = [F128, ctypes.POINTER(F128), ctypes.POINTER(F128)]

    # ---- f64 transforms ----
    def sincospi64(self, x):
        s, c = ctypes.c_double(), ctypes.c_double()
        self.L.or_sincospi64(x, ctypes.addressof(s), ctypes.addressof(c))
        return s.value, c.value

    def unordered_fwd(self, z, base_algo, base_n):
        z = np.ascontiguousarray(z, dtype=np.complex128).copy()
        n = z.shape[-1]
        rc = self.L.or_unordered_fwd_batch(n, int(base_algo), base_n, z.ctypes.data, z.size // n)
        assert rc == 0, rc
        return z

    def unordered_inv(self, z, base_algo, base_n):
        z = np.ascontiguousarray(z, dtype=np.complex128).copy()
        n = z.shape[-1]
        rc = self.L.or_unordered_inv_batch(n, int(base_algo), base_n, z.ctypes.data, z.size // n)
        assert rc == 0, rc
        return z

    def simd_unordered(self, isa, z, base_algo, base_n, fwd=True):
        """The V3 (isa 3) / V4 (isa 4) restatement of the reference's SIMD path."""
        z = np.ascontiguousarray(z, dtype=np.complex128).copy()
        n = z.shape[-1]
        f = getattr(self.L, "or_simd_unordered_%s_batch_v%d" % ("fwd" if fwd else "inv", isa))
        rc = f(n, int(base_algo), base_n, z.ctypes.data, z.size // n)
        assert rc == 0, rc
        return z

    def simd_polymul_f64(self, isa, a, b, base_algo, base_n, threads=8):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        N = a.shape[-1]
        c = np.zeros_like(a)
        rc = getattr(self.L, "or_simd_polymul_f64_v%d" % isa)(
            N, int(base_algo), base_n, a.ctypes.data, b.ctypes.data, c.ctypes.data, a.size // N,
            threads)
        assert rc == 0, rc
        return c

    def ordered_fwd(self, z, algo):
        z = np.ascontiguousarray(z, dtype=np.complex128).copy()
        n = z.shape[-1]
        f = self.L.or_ordered_fwd_batch if n <= 1024 else self.L.or_ordered_fwd_batch_large
        assert f(n, int(algo), z.ctypes.data, z.size // n) == 0
        return z

    def ordered_inv(self, z, algo):
        z = np.ascontiguousarray(z, dtype=np.complex128).copy()
        n = z.shape[-1]
        f = self.L.or_ordered_inv_batch if n <= 1024 else self.L.or_ordered_inv_batch_large
        assert f(n, int(algo), z.ctypes.data, z.size // n) == 0
        return z

    def stdrng_f64(self, seed, count):
        out = np.zeros(count, dtype=np.float64)
        self.L.or_stdrng_fill_f64_01(seed, out.ctypes.data, count)
        return out

    def bit_rev_twice(self, nbits, base_nbits, i):
        return self.L.or_bit_rev_twice(nbits, base_nbits, i)

    # ---- products ----
    def polymul_f64(self, a, b, base_algo, base_n, threads=8):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        N = a.shape[-1]
        c = np.zeros_like(a)
        rc = self.L.or_negacyclic_polymul_f64(N, int(base_algo), base_n, a.ctypes.data,
                                              b.ctypes.data, c.ctypes.data, a.size // N, threads)
        assert rc == 0, rc
        return c

    def schoolbook_u64(self, a, b, threads=8):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b).view(np.uint64) if np.asarray(b).dtype == np.int64 else \
            np.ascontiguousarray(b, dtype=np.uint64)
        b = np.ascontiguousarray(b)
        N = a.shape[-1]
        c = np.zeros_like(a)
        self.L.or_schoolbook_negacyclic_u64_batch(N, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                                  a.size // N, threads)
        return c

    def fft128_twiddles(self, n):
        t = [np.zeros(n) for _ in range(4)]
        self.L.or_fft128_init_twiddles(n, *[x.ctypes.data for x in t])
        return t

    def fft128_fwd(self, bufs, tw, mul_variant=0):
        bufs = [np.ascontiguousarray(x, dtype=np.float64).copy() for x in bufs]
        n = tw[0].size
        for b in range(bufs[0].size // n):
            ptrs = [x.ctypes.data + 8 * n * b for x in bufs]
            self.L.or_fft128_fwd(n, *ptrs, *[x.ctypes.data for x in tw], mul_variant)
        return bufs

    def fft128_inv(self, bufs, tw, mul_variant=0):
        bufs = [np.ascontiguousarray(x, dtype=np.float64).copy() for x in bufs]
        n = tw[0].size
        for b in range(bufs[0].size // n):
            ptrs = [x.ctypes.data + 8 * n * b for x in bufs]
            self.L.or_fft128_inv(n, *ptrs, *[x.ctypes.data for x in tw], mul_variant)
        return bufs

    def f128(self, name, a, b):
        r = getattr(self.L, "or_f128_" + name)(self.F128(*a), self.F128(*b))
        return (r.hi, r.lo)

    def f128_sincospi(self, x):
        s, c = self.F128(), self.F128()
        self.L.or_f128_sincospi(self.F128(*x), ctypes.byref(s), ctypes.byref(c))
        return (s.hi, s.lo), (c.hi, c.lo)

    def gadget_decompose(self, x, levels, base_log):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        out = np.zeros((levels, x.size), dtype=np.int64)
        self.L.or_gadget_decompose(x.size, x.ctypes.data, levels, base_log, out.ctypes.data)
        return out

    def negacyclic_roundtrip(self, x, base_algo, base_n):
        """(count, N) f64 -> (Fourier (count, N/2) c128, round trip (count, N) f64)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        N = x.shape[-1]
        F = np.zeros(x.shape[:-1] + (N // 2,), dtype=np.complex128)
        out = np.zeros_like(x)
        rc = self.L.or_negacyclic_roundtrip_f64(N, int(base_algo), base_n, x.ctypes.data,
                                                F.ctypes.data, out.ctypes.data, x.size // N)
        assert rc == 0
        return F, out

    def torus_to_fourier(self, polys, base_algo, base_n):
        polys = np.ascontiguousarray(polys, dtype=np.uint64)
        N = polys.shape[-1]
        out = np.zeros(polys.shape[:-1] + (N // 2,), dtype=np.complex128)
        rc = self.L.or_torus_to_fourier(N, int(base_algo), base_n, polys.ctypes.data,
                                        out.ctypes.data, polys.size // N)
        assert rc == 0
        return out

    def external_product(self, N, k1, levels, base_log, key_fourier, base_algo, base_n, glwe):
        key_fourier = np.ascontiguousarray(key_fourier, dtype=np.complex128)
        glwe = np.ascontiguousarray(glwe, dtype=np.uint64)
        out = np.zeros((k1, N), dtype=np.uint64)
        rc = self.L.or_external_product(N, k1, levels, base_log, key_fourier.ctypes.data,
                                        int(base_algo), base_n, glwe.ctypes.data, out.ctypes.data)
        assert rc == 0
        return out


_ORACLE = None


def load_oracle():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE
