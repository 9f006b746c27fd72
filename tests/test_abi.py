"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls; runs without a GPU)."""

import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for header in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(header).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for match in re.finditer(r"\b(rps_[a-z_0-9]+)\s*\(", text):
            names.add(match.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("rps_open", "rps_close", "rps_sieve", "rps_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol(pkg):
    path = pkg.LIB_PATH
    assert os.path.exists(path), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(path)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (rps_\w+)", out))
    assert set(declared_functions()) <= exported


def test_variants_listed(pkg):
    names = pkg.variants()
    assert names and names[0] == "q_u2_f12_w8"
    for n in names:
        assert re.fullmatch(r"([qw]_)?u\d_f\d+_w\d+", n)


def test_tuning_table_names_compiled_variants(pkg):
    # Every variant the autotuner's table names is compiled in, and the
    # choice is defined for any line length.
    import os
    names = pkg.variants()
    table = os.path.join(os.path.dirname(pkg.LIB_PATH), "..", "csrc", "tuning_table.inc")
    rows = re.findall(r'\{([^,]+), "([^"]+)", (\d+), (\d+), "([^"]+)", (\d+)\}',
                      open(table).read())
    assert rows and all(r[1] in names and r[4] in names for r in rows)
    for words in (1, 769, 2048, 10 ** 5, 3 * 10 ** 10):
        assert 0 <= pkg.tuned_variant(words) < len(names)
        for many in (False, True):
            v, run_log2, batch = pkg.tuned(words, many)
            assert 0 <= v < len(names) and 10 <= run_log2 <= 24
            assert batch == 0 or (not many and 1 <= batch <= 1 << 20)
    # The table's rows are what the lookup returns, bucket by bucket.
    for upper, name, run_log2, batch, many_name, many_run in rows:
        words = (1 << 40) if upper.startswith("~") else int(upper.rstrip("ULL"))
        assert pkg.tuned(words, False) == (names.index(name), int(run_log2), int(batch))
        assert pkg.tuned(words, True) == (names.index(many_name), int(many_run), 0)


def test_variant_info(pkg):
    for i, name in enumerate(pkg.variants()):
        info = pkg.variant_info(i)
        m = re.fullmatch(r"([qw]_)?u(\d)_f(\d+)_w(\d+)", name)
        assert info["words_per_lane"] == int(m.group(2))
        assert info["dense_primes"] == int(m.group(3))
        assert info["waves"] == int(m.group(4))
        assert info["wide"] == (1 if name.startswith("w_") else 0)
        assert 10 <= info["run_log2"] <= 24
    with pytest.raises(ValueError):
        pkg.variant_info(len(pkg.variants()))


def test_library_targets_gfx950(pkg):
    out = subprocess.run(["strings", pkg.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_missing_library_fails_loudly(pkg, tmp_path):
    saved = pkg._LIB
    pkg._LIB = None
    try:
        with pytest.raises(RuntimeError, match="not built"):
            pkg.load_library(str(tmp_path / "nope.so"))
    finally:
        pkg._LIB = saved


def test_check_request_limits(pkg):
    # Host-side validation only: the same limits as every sieve entry point.
    ok = (["1", "0", "1"], 1000, 1, 10)
    assert pkg.check_request(*ok) == (0, "")
    assert pkg.check_request(["1", "0", "1"], 2 ** 52 - 1, 1, 10) == (0, "")
    assert pkg.check_request(["1", "0", "1"], 2 ** 52, 1, 10) == (
        1, "maximum height exceeds GPU limit")
    assert pkg.check_request(["1", "0", "1"], 0, 1, 10) == (1, "invalid numerator bound")
    assert pkg.check_request(["1", "0", "1"], 10, 1, 46341, True) == (
        1, "invalid denominator range")
    assert pkg.check_request(["1", "0", "1"], 10, 1, 46340, True) == (0, "")
    assert pkg.check_request(["1"], 10, 1, 5)[0] == 1
    assert pkg.check_request(["1"] * 102, 10, 1, 5)[0] == 1
    assert pkg.check_request(["1", "x"], 10, 1, 5) == (
        1, "coefficients must be decimal integers")
