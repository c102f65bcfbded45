"""Uneven batch: f64 negacyclic round-trip jobs of mixed N in one launch.

No counterpart in the reference (it runs one plan per call); the per-class
plans are ordinary unordered plans, so each job's Fourier buffer is in that
plan's order (`class_algo`).
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .unordered import FftAlgo

JOB_DTYPE = np.dtype([("in_offset", "<u8"), ("fourier_offset", "<u8"), ("out_offset", "<u8"),
                      ("log2_N", "<u4"), ("reserved", "<u4")])


def _h(stream):
    return stream.handle if stream else None


def packed_jobs(log2_N):
    """Jobs laid out back to back in the order given: returns (jobs, in_len,
    fourier_len, out_len)."""
    log2_N = np.asarray(log2_N, dtype=np.uint32)
    N = np.left_shift(np.uint64(1), log2_N.astype(np.uint64))
    jobs = np.zeros(log2_N.size, dtype=JOB_DTYPE)
    off = np.concatenate([[0], np.cumsum(N)[:-1]]).astype(np.uint64)
    jobs["in_offset"] = off
    jobs["out_offset"] = off
    jobs["fourier_offset"] = off // np.uint64(2)
    jobs["log2_N"] = log2_N
    total = int(N.sum())
    return jobs, total, total // 2, total


class Schedule:
    """cfft_uneven: built once per job list; `run` replays the persistent
    launches from a captured graph, `run_direct` issues them on the stream,
    `run_plain` launches once per size class."""

    def __init__(self, jobs, in_len, fourier_len, out_len, measure=False, algo=FftAlgo.Dif4,
                 seconds=0.0):
        jobs = np.ascontiguousarray(jobs, dtype=JOB_DTYPE)
        self._jobs = jobs
        h = ctypes.c_void_p()
        check(lib.cfft_uneven_create(jobs.ctypes.data, jobs.size, in_len, fourier_len, out_len,
                                     1 if measure else 0, int(algo), float(seconds),
                                     ctypes.byref(h)))
        self.handle = h.value

    def run(self, x, fourier, y, stream=None):
        check(lib.cfft_uneven_run(self.handle, x.dptr, fourier.dptr, y.dptr, _h(stream)))

    def run_direct(self, x, fourier, y, stream=None):
        check(lib.cfft_uneven_run_direct(self.handle, x.dptr, fourier.dptr, y.dptr, _h(stream)))

    def run_plain(self, x, fourier, y, stream=None):
        check(lib.cfft_uneven_run_plain(self.handle, x.dptr, fourier.dptr, y.dptr, _h(stream)))

    def tile_schedule(self):
        """"device_queue" or "static_deal" (how the Dif16 tiles reach the
        persistent workgroups)."""
        k = ctypes.c_int()
        check(lib.cfft_uneven_tile_schedule(self.handle, ctypes.byref(k)))
        return "device_queue" if k.value == 0 else "static_deal"

    def set_tile_schedule(self, kind):
        check(lib.cfft_uneven_set_tile_schedule(
            self.handle, {"device_queue": 0, "static_deal": 1}[kind]))

    def class_algo(self, log2_N):
        a, b = ctypes.c_int(), ctypes.c_size_t()
        check(lib.cfft_uneven_class_algo(self.handle, int(log2_N), ctypes.byref(a), ctypes.byref(b)))
        return FftAlgo(a.value), b.value

    def info(self):
        j, t, c, g = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint32()
        check(lib.cfft_uneven_info(self.handle, ctypes.byref(j), ctypes.byref(t), ctypes.byref(c),
                                   ctypes.byref(g)))
        return {"jobs": j.value, "tiles": t.value, "classes": c.value, "grid": g.value}

    def close(self):
        if self.handle:
            lib.cfft_uneven_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
