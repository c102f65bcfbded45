"""cfft_amd -- Python mirror of concrete-fft's public API over libcfft_amd.so.

Modules mirror the reference crate: `ordered`, `unordered`, `fft128`; the
fused products (caller-side in the reference) are in `products`.  All
compute runs in the gfx950 kernels of libcfft_amd.so; there is no CPU path.
"""
from . import device, fft128, ordered, products, serve, tuning, uneven, unordered  # noqa: F401
from ._lib import EXPORTED, LIB_PATH, CfftPanic, check, last_error, lib  # noqa: F401
from .device import DeviceBuffer, Event, Stream, device_count  # noqa: F401
from .ordered import FftAlgo  # noqa: F401

c64 = complex
