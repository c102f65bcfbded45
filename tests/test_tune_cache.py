"""The tuning cache behind Method::Measure: file format, import/export,
validation and the default path (host code only, no device needed)."""
import os

import pytest

import cfft_amd
from cfft_amd import tuning

KEY = "AMD Instinct MI355X|gfx950:sramecc+:xnack-|cu256|hip70226015|abi1"
OTHER = "Some Other GPU|gfx942|cu304|hip60342134|abi1"

# (line, valid)
LINES = [
    ("# a comment", False),
    ("%s\t1\t1024\t6\t1024" % KEY, True),
    ("%s\t1\t2048\t6\t512" % KEY, True),
    ("%s\t0\t256\t4\t256" % KEY, True),
    ("%s\t1\t1024\t2\t1024" % OTHER, True),
    ("%s\t1\t65536\t7\t32" % KEY, True),
    ("%s\t1\t4096\t8\t1024" % KEY, False),    # algo out of range
    ("%s\t1\t4096\t6\t2048" % KEY, False),    # base_n above 2^10
    ("%s\t1\t4096\t6\t16" % KEY, False),      # base_n < 32 and != n
    ("%s\t1\t3000\t6\t1024" % KEY, False),    # n not a power of two
    ("%s\t0\t512\t6\t256" % KEY, False),      # ordered: base_n != n
    ("%s\t0\t131072\t6\t131072" % KEY, False),  # ordered above 2^16
    ("%s\t2\t1024\t6\t1024" % KEY, False),    # unknown kind
    ("%s\t1\t1024\t6" % KEY, False),          # missing field
    ("%s\t1\t1024\t6\t1024\t9" % KEY, False),  # extra field
    ("%s\t1\t1024\t-6\t1024" % KEY, False),   # sign
    ("%s\t1\t1024\t6\t1024 " % KEY, False),   # trailing blank
    ("\t1\t1024\t6\t1024", False),            # empty key
    ("", False),
]


@pytest.fixture(autouse=True)
def _no_default_file(monkeypatch):
    monkeypatch.setenv("CFFT_TUNE_CACHE", "off")
    tuning.clear()
    yield
    tuning.clear()


def _write(path, lines):
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _entries(path):
    out = {}
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line.startswith("#"):
                continue
            dev, kind, n, algo, base = line.split("\t")
            out[(dev, int(kind), int(n))] = (int(algo), int(base))
    return out


def test_import_skips_what_no_plan_accepts(tmp_path):
    p = tmp_path / "in.tsv"
    _write(p, [l for l, _ in LINES])
    assert tuning.import_file(p) == sum(v for _, v in LINES)
    assert tuning.size() == sum(v for _, v in LINES)
    assert tuning.lookup(KEY, tuning.UNORDERED, 1024) == (6, 1024)
    assert tuning.lookup(KEY, tuning.UNORDERED, 2048) == (6, 512)
    assert tuning.lookup(KEY, tuning.ORDERED, 256) == (4, 256)
    assert tuning.lookup(OTHER, tuning.UNORDERED, 1024) == (2, 1024)
    assert tuning.lookup(KEY, tuning.UNORDERED, 65536) == (7, 32)
    # keys differ in device, kind and n
    assert tuning.lookup(OTHER, tuning.UNORDERED, 2048) is None
    assert tuning.lookup(KEY, tuning.ORDERED, 1024) is None
    assert tuning.lookup(KEY, tuning.UNORDERED, 4096) is None


def test_export_then_import_round_trip(tmp_path):
    p, q = tmp_path / "in.tsv", tmp_path / "sub" / "out.tsv"
    _write(p, [l for l, v in LINES if v])
    tuning.import_file(p)
    tuning.export_file(q)  # creates the missing directory
    with open(q) as f:
        assert f.readline().startswith("# cfft_amd tuning cache v1")
    assert _entries(q) == _entries(p)
    tuning.clear()
    assert tuning.size() == 0
    assert tuning.import_file(q) == len(_entries(p))
    assert tuning.lookup(KEY, tuning.UNORDERED, 1024) == (6, 1024)
    assert not [n for n in os.listdir(q.parent) if ".tmp." in n]


def test_later_import_replaces_same_key(tmp_path):
    a, b = tmp_path / "a.tsv", tmp_path / "b.tsv"
    _write(a, ["%s\t1\t1024\t6\t1024" % KEY, "%s\t1\t512\t4\t512" % KEY])
    _write(b, ["%s\t1\t1024\t3\t512" % KEY])
    tuning.import_file(a)
    tuning.import_file(b)
    assert tuning.lookup(KEY, tuning.UNORDERED, 1024) == (3, 512)
    assert tuning.lookup(KEY, tuning.UNORDERED, 512) == (4, 512)
    assert tuning.size() == 2


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(cfft_amd.CfftPanic):
        tuning.import_file(tmp_path / "absent.tsv")
    blocker = tmp_path / "plain_file"
    blocker.write_text("")
    with pytest.raises(cfft_amd.CfftPanic):
        tuning.export_file(blocker / "x.tsv")  # parent is not a directory


def test_default_path(monkeypatch, tmp_path):
    monkeypatch.setenv("CFFT_TUNE_CACHE", str(tmp_path / "t.tsv"))
    assert tuning.default_path() == str(tmp_path / "t.tsv")
    for off in ("", "off"):
        monkeypatch.setenv("CFFT_TUNE_CACHE", off)
        assert tuning.default_path() == ""
    monkeypatch.delenv("CFFT_TUNE_CACHE")
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path / "xdg"))
    assert tuning.default_path() == str(tmp_path / "xdg" / "cfft_amd" / "tune-cache.tsv")
    monkeypatch.delenv("XDG_CACHE_HOME")
    monkeypatch.setenv("HOME", str(tmp_path / "home"))
    assert tuning.default_path() == str(tmp_path / "home" / ".cache" / "cfft_amd" /
                                        "tune-cache.tsv")


def test_long_lines_are_dropped(tmp_path):
    p = tmp_path / "in.tsv"
    _write(p, ["x" * 10000 + "\t1\t1024\t6\t1024", "%s\t1\t1024\t6\t1024" % KEY])
    assert tuning.import_file(p) == 1
    assert tuning.lookup(KEY, tuning.UNORDERED, 1024) == (6, 1024)


def test_clear_is_not_undone_by_the_file(tmp_path, monkeypatch):
    # a file written by another process keeps its entries, but a cleared
    # process table is not refilled from it
    p = tmp_path / "shared.tsv"
    _write(p, ["%s\t1\t1024\t6\t1024" % KEY, "%s\t1\t512\t4\t512" % OTHER])
    tuning.import_file(p)
    assert tuning.size() == 2
    tuning.clear()
    assert tuning.size() == 0
    assert tuning.lookup(KEY, tuning.UNORDERED, 1024) is None
    assert _entries(p)[(KEY, 1, 1024)] == (6, 1024)
