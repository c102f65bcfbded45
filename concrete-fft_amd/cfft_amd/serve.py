"""Request server: continuous batching of negacyclic products and external
products on one device (cfft_server_* in include/cfft.h).

Requests carry host numpy arrays (the server copies inputs when their batch
launches and writes outputs when it completes) or, with the *_device
request builders, DeviceBuffers of the server's device (gathered and
scattered on the device); either way the buffers must stay alive (this
wrapper keeps references) until `wait` returns.
"""
import ctypes

import numpy as np

from ._lib import Request, ServerConfig, ServerStatistics, check, lib

POLYMUL = 0
EXTERNAL_PRODUCT = 1
DEVICE_BUFFERS = 0x100


class Server:
    def __init__(self, hbm_budget_bytes=0, pinned_bytes=0, max_batch=0, streams=0):
        cfg = ServerConfig(hbm_budget_bytes, pinned_bytes, max_batch, streams)
        h = ctypes.c_void_p()
        check(lib.cfft_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self._live = {}
        self._released = []
        self._keys = {}

    def add_key(self, key, k1, levels, base_log):
        """key: (k1*levels, k1, N) uint64 torus polynomials."""
        key = np.ascontiguousarray(key, dtype=np.uint64)
        N = key.shape[-1]
        assert key.shape == (k1 * levels, k1, N)
        kid = ctypes.c_int()
        check(lib.cfft_server_add_key(self.handle, N, k1, levels, base_log, key.ctypes.data,
                                      ctypes.byref(kid)))
        self._keys[kid.value] = (N, k1)
        return kid.value

    def polymul_request(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        out = np.zeros_like(a)
        r = Request(POLYMUL, -1, a.size, a.ctypes.data, b.ctypes.data, out.ctypes.data)
        return r, (a, b, out)

    def extprod_request(self, key_id, glwe):
        N, k1 = self._keys[key_id]
        glwe = np.ascontiguousarray(glwe, dtype=np.uint64)
        assert glwe.shape == (k1, N)
        out = np.zeros_like(glwe)
        r = Request(EXTERNAL_PRODUCT, key_id, N, glwe.ctypes.data, None, out.ctypes.data)
        return r, (glwe, out)

    def polymul_request_device(self, a, b, out):
        """a (N u64), b (N i64), out (N u64): DeviceBuffers; out is written
        on the device."""
        N = int(np.prod(a.shape))
        r = Request(POLYMUL | DEVICE_BUFFERS, -1, N, a.dptr, b.dptr, out.dptr)
        return r, (a, b, out)

    def extprod_request_device(self, key_id, glwe, out):
        """glwe, out: (k1, N) u64 DeviceBuffers (out may be glwe: the input
        is gathered before the output is written)."""
        N, k1 = self._keys[key_id]
        assert int(np.prod(glwe.shape)) == k1 * N and int(np.prod(out.shape)) == k1 * N
        r = Request(EXTERNAL_PRODUCT | DEVICE_BUFFERS, key_id, N, glwe.dptr, None, out.dptr)
        return r, (glwe, out)

    def submit(self, request, keep):
        t = ctypes.c_uint64()
        check(lib.cfft_server_submit(self.handle, ctypes.byref(request), ctypes.byref(t)))
        self._live[t.value] = keep
        return t.value

    def wait(self, ticket):
        keep = self._live.get(ticket)
        st = lib.cfft_server_wait(self.handle, ticket)
        # an unknown or already-waited ticket (status 10) leaves the buffers
        # to the thread that owns the wait
        if st != 10:
            self._live.pop(ticket, None)
        check(st)
        return keep[-1]

    def poll(self, ticket):
        done = ctypes.c_int()
        check(lib.cfft_server_poll(self.handle, ticket, ctypes.byref(done)))
        return bool(done.value)

    def release(self, ticket):
        """Gives up a ticket without waiting; the request's output is still
        written, so its buffers stay referenced until the server is closed."""
        check(lib.cfft_server_release(self.handle, ticket))
        self._released.append(self._live.pop(ticket, None))

    def replay(self, requests, arrival_s):
        """requests: list of Request structs (their buffers kept alive by the
        caller); returns latencies in seconds from each scheduled arrival."""
        n = len(requests)
        arr = (Request * n)(*requests)
        arrival_s = np.ascontiguousarray(arrival_s, dtype=np.float64)
        lat = np.zeros(n, dtype=np.float64)
        check(lib.cfft_server_replay(self.handle, arr, n, arrival_s.ctypes.data, lat.ctypes.data))
        return lat

    def stats(self):
        st = ServerStatistics()
        check(lib.cfft_server_stats(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStatistics._fields_}

    def close(self):
        if self.handle:
            lib.cfft_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
