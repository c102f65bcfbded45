"""Extract the reference's bit-exact golden vector into a fixture.

Source: zama-ai/concrete-fft v0.5.1, src/unordered.rs:1176-1395
(`test_equivalency`): a 2048-point `unordered::Plan` forward FFT with
`Method::UserProvided { base_algo: FftAlgo::Dif4, base_n: 32 }` of inputs
drawn from `StdRng::seed_from_u64(0)` with `gen_range(0.0..1.0)` (re, then
im, per element).  The expected output array `target` (:1201-1394) is read as
text and written, as little-endian float64 pairs, to
`unordered_n2048_dif4_b32.npy` (complex128, shape (2048,)).

Run from the repository root:  python tests/golden/make_unordered_golden.py
"""
import os
import re
import sys

import numpy as np

REF = "/root/reference/src/unordered.rs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "unordered_n2048_dif4_b32.npy")


def main() -> int:
    with open(REF, "r", encoding="utf-8") as fh:
        text = fh.read()
    start = text.index("let target: [Complex<f64>; 2048] = [")
    end = text.index("];", start)
    body = text[start:end]
    pat = re.compile(r"Complex\s*\{\s*re:\s*([-+0-9.eE]+),\s*im:\s*([-+0-9.eE]+),?\s*\}")
    vals = [complex(float(a), float(b)) for a, b in pat.findall(body)]
    if len(vals) != 2048:
        print(f"expected 2048 entries, parsed {len(vals)}", file=sys.stderr)
        return 1
    arr = np.array(vals, dtype=np.complex128)
    # self-check quoted in the source: element 0 is the plain sum of the inputs
    assert arr[0] == complex(1028.0259662650005, 1018.8306666971462)
    np.save(OUT, arr, allow_pickle=False)
    print(f"wrote {OUT}: {arr.shape} {arr.dtype}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
