"""The C++ mirror of the reference API (concrete-fft_amd/host/concrete_fft.hpp)
runs the reference's doctest / roundtrip relations on the device."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "concrete-fft_amd", "lib", "example_roundtrip")


def test_cpp_mirror_roundtrips(cfft):
    assert os.path.exists(EXE), "build the library first (concrete-fft_amd/build.py)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
    assert "panic ok: assertion failed: n.is_power_of_two()" in r.stdout
    assert "ordered n=4096 (large) roundtrip" in r.stdout
