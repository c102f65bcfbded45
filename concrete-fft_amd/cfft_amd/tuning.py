"""The tuning cache behind Method::Measure (no reference counterpart: the
reference measures again in every process, src/unordered.rs:553-640).

Entries map (device key, kind, n) to (algo, base_n); kind 0 is an ordered
plan, 1 an unordered plan.  Measure plans read the default file once and
rewrite it after each new measurement; see `default_path`.
"""
import ctypes

from ._lib import check, lib

ORDERED = 0
UNORDERED = 1


def _string(fn):
    n = ctypes.c_size_t(0)
    check(fn(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    check(fn(buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


def default_path():
    """$CFFT_TUNE_CACHE, else $XDG_CACHE_HOME/cfft_amd/tune-cache.tsv, else
    $HOME/.cache/cfft_amd/tune-cache.tsv; "" when CFFT_TUNE_CACHE is "" or
    "off" (choices stay in the process)."""
    return _string(lib.cfft_tune_cache_path)


def device_key():
    """The current device's key (model|ISA|cuN|hipV|abiV); needs a device."""
    return _string(lib.cfft_tune_device_key)


def import_file(path):
    """Adds the entries of `path`; returns how many were read."""
    n = ctypes.c_size_t(0)
    check(lib.cfft_tune_cache_import(str(path).encode(), ctypes.byref(n)))
    return n.value


def export_file(path):
    check(lib.cfft_tune_cache_export(str(path).encode()))


def clear():
    check(lib.cfft_tune_cache_clear())


def size():
    n = ctypes.c_size_t(0)
    check(lib.cfft_tune_cache_size(ctypes.byref(n)))
    return n.value


def lookup(device_key, kind, n):
    """(algo, base_n) of the process table's entry, or None."""
    found, algo, base = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_size_t(0)
    check(lib.cfft_tune_cache_lookup(device_key.encode(), int(kind), int(n), ctypes.byref(found),
                                     ctypes.byref(algo), ctypes.byref(base)))
    return (algo.value, base.value) if found.value else None
