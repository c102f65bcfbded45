"""ctypes binding of libcfft_amd.so (the C ABI declared in include/cfft.h).

The shared library is loaded from the package's in-tree build directory
(concrete-fft_amd/lib/libcfft_amd.so).  There is no fallback: if it is
missing, importing this module raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libcfft_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libcfft_amd.so is not built (%s); run concrete-fft_amd/build.py" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

c_size = ctypes.c_size_t
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_dbl = ctypes.c_double
c_u64 = ctypes.c_uint64


class CfftPanic(RuntimeError):
    """Raised where the reference would panic (contract violation), or on a
    device error.  `status` is the cfft_status code."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("pinned_bytes", ctypes.c_size_t),
                ("max_batch", ctypes.c_uint32), ("streams", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("key_id", ctypes.c_int), ("N", ctypes.c_size_t),
                ("in0", ctypes.c_void_p), ("in1", ctypes.c_void_p), ("out", ctypes.c_void_p)]


class ServerStatistics(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("admission_waits", ctypes.c_uint64), ("max_inflight", ctypes.c_uint64),
                ("hbm_budget_bytes", ctypes.c_size_t), ("streams", ctypes.c_uint32)]


# (name, restype, argtypes)
_SIGS = [
    ("cfft_status_string", ctypes.c_char_p, [c_int]),
    ("cfft_last_error_message", ctypes.c_char_p, []),
    ("cfft_abi_version", c_int, []),
    ("cfft_device_count", c_int, [ctypes.POINTER(c_int)]),
    ("cfft_set_device", c_int, [c_int]),
    ("cfft_device_alloc", c_int, [c_size, ctypes.POINTER(c_vp)]),
    ("cfft_device_free", c_int, [c_vp]),
    ("cfft_device_mem_info", c_int, [ctypes.POINTER(c_size), ctypes.POINTER(c_size)]),
    ("cfft_memcpy_htod", c_int, [c_vp, c_vp, c_size, c_vp]),
    ("cfft_memcpy_dtoh", c_int, [c_vp, c_vp, c_size, c_vp]),
    ("cfft_memcpy_dtod", c_int, [c_vp, c_vp, c_size, c_vp]),
    ("cfft_memset", c_int, [c_vp, c_int, c_size, c_vp]),
    ("cfft_stream_create", c_int, [ctypes.POINTER(c_vp)]),
    ("cfft_stream_destroy", c_int, [c_vp]),
    ("cfft_stream_synchronize", c_int, [c_vp]),
    ("cfft_device_synchronize", c_int, []),
    ("cfft_event_create", c_int, [ctypes.POINTER(c_vp)]),
    ("cfft_event_destroy", c_int, [c_vp]),
    ("cfft_event_record", c_int, [c_vp, c_vp]),
    ("cfft_event_synchronize", c_int, [c_vp]),
    ("cfft_event_elapsed_ms", c_int, [c_vp, c_vp, ctypes.POINTER(ctypes.c_float)]),
    ("cfft_ordered_plan_create", c_int, [c_size, c_int, c_int, c_dbl, ctypes.POINTER(c_vp)]),
    ("cfft_ordered_plan_create_large", c_int, [c_size, c_int, c_int, c_dbl, ctypes.POINTER(c_vp)]),
    ("cfft_unordered_plan_create", c_int,
     [c_size, c_int, c_int, c_size, c_dbl, ctypes.POINTER(c_vp)]),
    ("cfft128_plan_create", c_int, [c_size, ctypes.POINTER(c_vp)]),
    ("cfft_plan_destroy", c_int, [c_vp]),
    ("cfft_plan_fft_size", c_int, [c_vp, ctypes.POINTER(c_size)]),
    ("cfft_plan_algo", c_int, [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_size)]),
    ("cfft_plan_fft_scratch", c_int, [c_vp, ctypes.POINTER(c_size), ctypes.POINTER(c_size)]),
    ("cfft_plan_permutation", c_int, [c_vp, c_vp]),
    ("cfft_fwd", c_int, [c_vp, c_vp, c_size, c_size, c_vp]),
    ("cfft_inv", c_int, [c_vp, c_vp, c_size, c_size, c_vp]),
    ("cfft_fwd_host", c_int, [c_vp, c_vp, c_size]),
    ("cfft_inv_host", c_int, [c_vp, c_vp, c_size]),
    ("cfft_fwd_monomial", c_int, [c_vp, c_size, c_vp, c_vp]),
    ("cfft_fwd_monomial_host", c_int, [c_vp, c_size, c_vp]),
    ("cfft_fourier_to_standard", c_int, [c_vp, c_vp, c_vp, c_size, c_vp]),
    ("cfft_fourier_from_standard", c_int, [c_vp, c_vp, c_vp, c_size, c_vp]),
    ("cfft_serialize_fourier_buffer", c_int, [c_vp, c_vp, c_vp, c_size, ctypes.POINTER(c_size)]),
    ("cfft_deserialize_fourier_buffer", c_int, [c_vp, c_vp, c_size, c_vp]),
    ("cfft128_fwd", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    ("cfft128_inv", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    ("cfft128_fwd_host", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_size]),
    ("cfft128_inv_host", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_size]),
    ("cfft128_negacyclic_fwd_fft", c_int,
     [c_size, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_int, c_vp]),
    ("cfft128_negacyclic_inv_fft", c_int,
     [c_size, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_int, c_vp]),
    ("cfft128_init_negacyclic_twiddles", c_int, [c_size, c_vp, c_vp, c_vp, c_vp]),
    ("cfft_negacyclic_polymul_u64", c_int, [c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    ("cfft_exact_polymul_u64", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    ("cfft128_exact_polymul_u64", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    ("cfft_ggsw_to_fourier", c_int, [c_vp, c_vp, c_int, c_int, c_vp, c_vp]),
    ("cfft_external_product", c_int, [c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp, c_size, c_vp]),
    ("cfft_external_product_noalloc", c_int,
     [c_vp, c_vp, c_int, c_int, c_int, c_vp, c_vp, c_size, c_vp]),
    ("cfft_noalloc_scratch_bytes", c_size, []),
    ("cfft_uneven_create", c_int, [c_vp, c_size, c_size, c_size, c_size, c_int, c_int, c_dbl,
                                   ctypes.POINTER(c_vp)]),
    ("cfft_uneven_destroy", c_int, [c_vp]),
    ("cfft_uneven_run", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("cfft_uneven_run_plain", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("cfft_uneven_run_direct", c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("cfft_uneven_tile_schedule", c_int, [c_vp, ctypes.POINTER(c_int)]),
    ("cfft_uneven_set_tile_schedule", c_int, [c_vp, c_int]),
    ("cfft_uneven_class_algo", c_int, [c_vp, ctypes.c_uint32, ctypes.POINTER(c_int),
                                       ctypes.POINTER(c_size)]),
    ("cfft_server_create", c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(c_vp)]),
    ("cfft_server_destroy", c_int, [c_vp]),
    ("cfft_server_add_key", c_int, [c_vp, c_size, c_int, c_int, c_int, c_vp,
                                    ctypes.POINTER(c_int)]),
    ("cfft_server_submit", c_int, [c_vp, ctypes.POINTER(Request), ctypes.POINTER(c_u64)]),
    ("cfft_server_poll", c_int, [c_vp, c_u64, ctypes.POINTER(c_int)]),
    ("cfft_server_wait", c_int, [c_vp, c_u64]),
    ("cfft_server_release", c_int, [c_vp, c_u64]),
    ("cfft_server_replay", c_int, [c_vp, ctypes.POINTER(Request), c_size, c_vp, c_vp]),
    ("cfft_server_stats", c_int, [c_vp, ctypes.POINTER(ServerStatistics)]),
    ("cfft_tune_cache_path", c_int, [ctypes.c_char_p, c_size, ctypes.POINTER(c_size)]),
    ("cfft_tune_device_key", c_int, [ctypes.c_char_p, c_size, ctypes.POINTER(c_size)]),
    ("cfft_tune_cache_import", c_int, [ctypes.c_char_p, ctypes.POINTER(c_size)]),
    ("cfft_tune_cache_export", c_int, [ctypes.c_char_p]),
    ("cfft_tune_cache_clear", c_int, []),
    ("cfft_tune_cache_size", c_int, [ctypes.POINTER(c_size)]),
    ("cfft_tune_cache_lookup", c_int, [ctypes.c_char_p, c_int, c_size, ctypes.POINTER(c_int),
                                       ctypes.POINTER(c_int), ctypes.POINTER(c_size)]),
    ("cfft_uneven_info", c_int, [c_vp, ctypes.POINTER(c_size), ctypes.POINTER(c_size),
                                 ctypes.POINTER(c_size), ctypes.POINTER(ctypes.c_uint32)]),
]

for _name, _res, _args in _SIGS:
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = [s[0] for s in _SIGS]


def last_error():
    m = lib.cfft_last_error_message()
    return m.decode() if m else ""


def check(status):
    if status != 0:
        raise CfftPanic(status, last_error() or lib.cfft_status_string(status).decode())
    return status


def ptr(x):
    """Address of a numpy array's data, a DeviceBuffer, or an int address."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "dptr"):
        return x.dptr
    return x.ctypes.data
