"""Method::Measure and its tuning cache on the device: a Measure plan in a
fresh process records its choice in the cache file, and a Measure plan in a
later process takes the file's choice without timing (the test plants a
choice the timing would not pick, so only a file read can produce it)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "concrete-fft_amd")

CHILD = r"""
import json, sys, time
sys.path.insert(0, sys.argv[1])
import cfft_amd
from cfft_amd import ordered, tuning, unordered
n = int(sys.argv[2])
# runtime start-up before the clock starts: only plan creation is timed
cfft_amd.check(cfft_amd.lib.cfft_device_synchronize())
key = tuning.device_key()
t0 = time.perf_counter()
p = unordered.Plan(n, unordered.Method.Measure(0.002))
t1 = time.perf_counter()
q = ordered.Plan(256, ordered.Method.Measure(0.002))
t2 = time.perf_counter()
print(json.dumps({"key": key, "unordered": [int(v) for v in p.algo()],
                  "ordered": int(q.algo()), "t_unordered": t1 - t0, "t_ordered": t2 - t1}))
"""


def _child(cache, n):
    env = dict(os.environ, CFFT_TUNE_CACHE=str(cache))
    print("tune child: n=%d cache=%s" % (n, cache), flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, PKG, str(n)], capture_output=True, text=True,
                       timeout=110, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def _entries(path):
    out = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            dev, kind, n, algo, base = line.rstrip("\n").split("\t")
            out[(dev, int(kind), int(n))] = (int(algo), int(base))
    return out


@pytest.mark.gpu
def test_measure_choice_persists_across_processes(cfft, tmp_path):
    cache = tmp_path / "tune" / "cache.tsv"
    first = _child(cache, 1024)
    key = first["key"]
    assert "gfx950" in key
    table = _entries(cache)
    assert table[(key, 1, 1024)] == tuple(first["unordered"])
    assert table[(key, 0, 256)] == (first["ordered"], 256)
    # plant choices the timing would not make; a new process must take them
    planted_u, planted_o = (1, 512), 0  # Dit2 over base 512, Dif2
    if tuple(first["unordered"]) == planted_u:
        planted_u = (3, 512)
    if first["ordered"] == planted_o:
        planted_o = 1
    table[(key, 1, 1024)] = planted_u
    table[(key, 0, 256)] = (planted_o, 256)
    with open(cache, "w") as f:
        f.write("# planted\n")
        for (dev, kind, n), (algo, base) in table.items():
            f.write("%s\t%d\t%d\t%d\t%d\n" % (dev, kind, n, algo, base))
    second = _child(cache, 1024)
    assert tuple(second["unordered"]) == planted_u
    assert second["ordered"] == planted_o
    # a cache hit builds one plan instead of timing 16 candidates
    assert second["t_unordered"] < first["t_unordered"]


@pytest.mark.gpu
def test_measure_in_process_reuses_choice(cfft, tmp_path, monkeypatch):
    from cfft_amd import tuning, unordered
    monkeypatch.setenv("CFFT_TUNE_CACHE", "off")
    tuning.clear()
    try:
        p = unordered.Plan(4096, unordered.Method.Measure(0.002))
        key = tuning.device_key()
        chosen = tuple(int(v) for v in p.algo())
        assert tuning.lookup(key, tuning.UNORDERED, 4096) == chosen
        q = unordered.Plan(4096, unordered.Method.Measure(0.002))
        assert tuple(int(v) for v in q.algo()) == chosen
        # the chosen plan computes a correct round trip
        rng = np.random.default_rng(5)
        x = rng.standard_normal(4096) + 1j * rng.standard_normal(4096)
        z = x.copy()
        q.fwd(z)
        q.inv(z)
        assert np.max(np.abs(z / 4096 - x)) < 1e-12
    finally:
        tuning.clear()
