"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point declared in include/cfft.h, and rejects contract violations with
the reference's panic text before touching the device."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cfft.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(cfft\w*)\s*\(", text)
    return sorted(set(n for n in names if not n.isupper()))


def test_header_parses():
    names = declared_functions()
    assert "cfft_fwd" in names and "cfft_unordered_plan_create" in names
    assert len(names) > 40


def test_library_exports_every_declared_symbol():
    import cfft_amd
    missing = [n for n in declared_functions() if not hasattr(cfft_amd.lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert not [n for n in declared_functions() if n not in cfft_amd.EXPORTED]


def test_library_exports_nothing_else():
    # the dynamic symbol table holds the header's entry points only (the
    # kernels and host internals stay local to the library)
    import shutil
    import subprocess
    from cfft_amd._lib import LIB_PATH
    if not shutil.which("nm"):
        pytest.skip("nm is not available")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert exported == set(declared_functions()), sorted(exported ^ set(declared_functions()))


def test_abi_version():
    import cfft_amd
    assert cfft_amd.lib.cfft_abi_version() == 1


@pytest.mark.parametrize("n", [0, 3, 6, 1000])
def test_ordered_rejects_non_power_of_two(n):
    import cfft_amd
    from cfft_amd.ordered import FftAlgo, Method, Plan
    with pytest.raises(cfft_amd.CfftPanic) as e:
        Plan(n, Method.UserProvided(FftAlgo.Dif4))
    assert "n.is_power_of_two()" in str(e.value)
    assert e.value.status == 1


def test_ordered_rejects_too_large():
    import cfft_amd
    from cfft_amd.ordered import FftAlgo, Method, Plan
    with pytest.raises(cfft_amd.CfftPanic) as e:
        Plan(2048, Method.UserProvided(FftAlgo.Dif4))
    assert "n.trailing_zeros() < 11" in str(e.value)


@pytest.mark.parametrize("n,status", [(131072, 2), (3072, 1)])
def test_ordered_large_limits(n, status):
    # the extended entry point stops at the reference's documented 2^16
    import cfft_amd
    from cfft_amd.ordered import FftAlgo, Method, Plan
    with pytest.raises(cfft_amd.CfftPanic) as e:
        Plan(n, Method.UserProvided(FftAlgo.Dif4), large=True)
    assert e.value.status == status


@pytest.mark.parametrize("n,base_n,msg", [
    (1024, 48, "base_n.is_power_of_two()"),
    (1024, 2048, "base_n <= n"),
    (1024, 16, "base_n >= 32"),
    (4096, 2048, "base_n.trailing_zeros() <= 10"),
])
def test_unordered_rejects_bad_base(n, base_n, msg):
    import cfft_amd
    from cfft_amd.unordered import FftAlgo, Method, Plan
    with pytest.raises(cfft_amd.CfftPanic) as e:
        Plan(n, Method.UserProvided(FftAlgo.Dif4, base_n))
    assert msg in str(e.value)
    assert e.value.status == 3


def test_status_strings():
    import cfft_amd
    for s in range(0, 15):
        assert cfft_amd.lib.cfft_status_string(s)
    assert cfft_amd.lib.cfft_status_string(999) == b"unknown status"


def test_null_plan_is_an_error_not_a_crash():
    import cfft_amd
    assert cfft_amd.lib.cfft_fwd(None, None, 1, 1, None) == 5
    assert cfft_amd.lib.cfft_plan_destroy(None) == 0
