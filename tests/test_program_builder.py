"""CPU checks of the pass-list builder (cfft_prog::fixed_program), which makes
every plan's device program and the compile-time kernels' programs: for each
algorithm, n = 2^0 .. 2^30 and base_n = 2^0 .. min(n, 2^16), the list is the
reference's plan structure -- depth levels of radix 2 / 4 / 8 above base_n
(src/unordered.rs:361-367, :407-413; forward top level first, inverse in
reverse), then the Stockham recursion of the base (src/dif*.rs, src/dit*.rs:
radix-R cores at strides R^0, R^1, ... while more than log2 R levels remain,
then the end butterfly; DIT runs the end first and the cores in reverse) --
and it fits the device program (at most 24 passes) for every size the C ABI
accepts (n <= 2^30 with base_n <= 1024, ordered plans up to 2^16)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "concrete-fft_amd", "csrc")
DEPTH_FWD, DEPTH_INV, DIF_CORE, DIT_CORE, END = range(5)


@pytest.fixture(scope="module")
def programs(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("prog") / "program_dump")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC,
                    os.path.join(HERE, "native", "program_dump.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe, "30"], check=True, capture_output=True, text=True).stdout
    progs = {}
    for line in out.splitlines():
        f = line.split()
        algo, l2n, l2b, fwd, npass = map(int, f[:5])
        passes = [tuple(map(int, p.split(":"))) for p in f[5:]]
        progs[(algo, l2n, l2b, fwd)] = (npass, passes)
    return progs


def expected(algo, l2n, l2b, fwd):
    """The reference's plan structure, restated independently of the C++."""
    radix = 2 << (algo // 2)
    lr = radix.bit_length() - 1
    dit = algo % 2 == 1
    levels, nl = [], l2n
    while nl > l2b:
        r = 2 if nl == l2b + 1 else 4 if nl == l2b + 2 else 8
        levels.append((r, nl))
        nl -= r.bit_length() - 1
    tw = l2n - l2b
    base = []
    if l2b > 0:
        strides, L, s = [], l2b, 0
        while L > lr:
            strides.append(s)
            s += lr
            L -= lr
        end = (END, 1 << L, fwd, l2b, 0, tw)
        if not dit:
            base = [(DIF_CORE, radix, fwd, l2b, st, tw) for st in strides] + [end]
        else:
            base = [end] + [(DIT_CORE, radix, fwd, l2b, st, tw) for st in reversed(strides)]
    if fwd:
        return [(DEPTH_FWD, r, 1, l, 0, l2n - l) for r, l in levels] + base
    return base + [(DEPTH_INV, r, 0, l, 0, l2n - l) for r, l in reversed(levels)]


def test_lists_follow_the_reference_structure(programs):
    checked = 0
    for (algo, l2n, l2b, fwd), (npass, passes) in programs.items():
        want = expected(algo, l2n, l2b, fwd)
        if len(want) > 24:
            assert npass == 25, (algo, l2n, l2b, fwd)  # marked as too long
            continue
        assert npass == len(want) and passes == want, (algo, l2n, l2b, fwd)
        checked += 1
    assert checked > 5000


def test_every_accepted_size_fits_the_device_program(programs):
    for (algo, l2n, l2b, fwd), (npass, _) in programs.items():
        unordered = l2n <= 30 and (l2b == l2n and l2n <= 10 or 5 <= l2b <= 10)
        ordered = l2b == l2n and l2n <= 16
        if unordered or ordered:
            assert npass <= 24, (algo, l2n, l2b, fwd)


def test_depth_radices_multiply_to_n_over_base(programs):
    for (algo, l2n, l2b, fwd), (npass, passes) in programs.items():
        if npass > 24:
            continue
        depth = [p for p in passes if p[0] in (DEPTH_FWD, DEPTH_INV)]
        assert sum(p[1].bit_length() - 1 for p in depth) == l2n - l2b
        base = [p for p in passes if p[0] not in (DEPTH_FWD, DEPTH_INV)]
        assert sum(p[1].bit_length() - 1 for p in base) == l2b
