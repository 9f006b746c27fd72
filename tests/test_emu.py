"""The HIP library's own sources built for the CPU wave emulator
(tests/emu/_build/libratpoints_sieve_emu.so), driven through the same C ABI
as the GPU tests.

Bar: survivor multisets bit-identical to the CPU oracle's (or to rps_sieve's
for the other entry points), with no barrier, cross-lane or watchdog report
from the emulator.  The emulator runs every lane of a workgroup as a fiber,
checks that the lanes of a wave meet at the same cross-lane operation and
that the waves meet at the same barrier, and here also fills the kernels'
LDS with garbage before each workgroup and interleaves the waves in a seeded
random order, so a read of LDS no lane wrote or a race between waves shows
up as a wrong survivor set.  These tests check the kernels' control flow and
indexing on the CPU; the GPU tests check the same cases on the device.
"""

import os
import random
import subprocess

import numpy as np
import pytest

import curves
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_DIR = os.path.join(ROOT, "tests", "emu", "_build")
EMU_LIB = os.path.join(EMU_DIR, "libratpoints_sieve_emu.so")
EMU_SERVE = os.path.join(EMU_DIR, "ratpoints_serve_emu")
CONTRACT = os.path.join(ROOT, "tests", "cpp", "_build", "streaming_contract_emu")

SCHEDULE = {"RPS_EMU_SCHEDULE": "random", "RPS_EMU_POISON_LDS": "1", "RPS_EMU_SEED": "11",
            "RPS_EMU_WATCHDOG_S": "300"}


@pytest.fixture(scope="module")
def emu(pkg):
    if not os.path.exists(EMU_LIB):
        pytest.skip("emulated library not built")
    os.environ.update(SCHEDULE)  # read once by the emulator, at its first launch
    lib = pkg.bind_library(EMU_LIB)
    ctx = pkg.SieveContext(0, lib=lib)
    yield ctx
    ctx.close()


def variants(ctx):
    """Variant names of the library behind `ctx` (the emulated one here)."""
    return [ctx.lib.rps_variant_name(i).decode() for i in range(ctx.lib.rps_variant_count())]


def defaults(ctx):
    for name, value in (("variant", -1), ("run_log2", -1), ("survivor_capacity", 1 << 24),
                        ("batch_size", 65536),
                        ("min_tiles", 2048), ("direct_words", 2048), ("primitive_only", 0)):
        ctx.set_option(name, value)
    return ctx


def same(res, ref):
    count, digest, num, den = ref
    return (res.survivor_count == count and np.array_equal(res.numerators, num)
            and np.array_equal(res.denominators, den))


def test_every_variant_matches_oracle(emu, oracle_lib):
    # One long-line case per compiled variant: lines of 1,251 words in runs
    # (min_tiles 1), denominators divisible by candidate primes included.
    coeffs, _ = curves.partially_dense(random.Random(5), 7, 6)
    ref = oracle_lib.sieve(coeffs, 40000, 437, 446, method="words")
    for v, name in enumerate(variants(emu)):
        defaults(emu).set_option("variant", v)
        emu.set_option("min_tiles", 1)
        assert same(emu.sieve(coeffs, 40000, 437, 446), ref), name


def test_resume_exactly_once(emu, oracle_lib):
    # About a quarter of all sites survive: a 16,384-entry buffer overflows
    # in every launch and the waves resume at their cursors.
    coeffs, _ = curves.partially_dense(random.Random(78), 2, 4)
    ref = oracle_lib.sieve(coeffs, 20000, 3, 6, method="words")
    for name in ("q_u1_f12_w8", "q_u2_f12_w8", "w_u2_f14_w16"):
        defaults(emu).set_option("variant", variants(emu).index(name))
        emu.set_option("survivor_capacity", 16384)
        res = emu.sieve(coeffs, 20000, 3, 6)
        assert res.stats["launches"] > 2
        assert same(res, ref), name


def carry_curve():
    # Nine of the twelve dense primes live and every late prime dead: about
    # an eighth of the words pass the dense rows, so each wave queues a few
    # dozen words per 2,000-word line and its queue spans several lines,
    # while an eighth of a survivor per word fills a 16,384-entry buffer
    # every ~65 lines.
    dense = sorted(oracle.candidate_primes(), reverse=True)[:12]
    return curves.with_live_primes(random.Random(31), dense[:9], 6)


def test_queue_carried_across_lines_resumes_exactly_once(emu, oracle_lib):
    coeffs = carry_curve()
    ref = oracle_lib.sieve(coeffs, 64000, 1, 160, method="words")
    for name in ("q_u1_f12_w8", "q_u2_f12_w8", "w_u2_f14_w16"):
        defaults(emu).set_option("variant", variants(emu).index(name))
        emu.set_option("min_tiles", 1)
        emu.set_option("survivor_capacity", 16384)
        res = emu.sieve(coeffs, 64000, 1, 160)
        assert res.stats["launches"] > 2, name
        assert same(res, ref), name


def test_queue_carried_across_lines_many_path(emu, oracle_lib):
    # The same run through the persistent launch: a full buffer marks the
    # curve failed mid-run and the host searches it again on the plain path.
    reqs = [{"coefficients": carry_curve(), "bound": 64000, "first": 1, "last": 160},
            {"coefficients": ["1", "0", "3", "0", "0", "-7", "2"], "bound": 64000,
             "first": 1, "last": 40}]
    defaults(emu).set_option("variant", variants(emu).index("q_u1_f12_w8"))
    emu.set_option("min_tiles", 1)
    emu.set_option("survivor_capacity", 16384)
    results, stats = emu.sieve_many(reqs)
    assert stats["requeued"] >= 1
    for r, (num, den) in zip(reqs, results):
        count, digest, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                                 r["last"], method="words")
        assert np.array_equal(num, on) and np.array_equal(den, od)


def test_short_lines_and_square_mode(emu, oracle_lib):
    coeffs, _ = curves.partially_dense(random.Random(9), 6, 5)
    defaults(emu)
    assert same(emu.sieve(coeffs, 3000, 1, 40), oracle_lib.sieve(coeffs, 3000, 1, 40,
                                                                  method="words"))
    assert same(emu.sieve(coeffs, 9000, 46300, 46340, square=True),
                oracle_lib.sieve(coeffs, 9000, 46300, 46340, square=True, method="words"))


def test_many_curves_mixed_tiles(emu, oracle_lib):
    # Row runs, chunked lines, wide-row lines, short lines and square
    # denominators in one call: two persistent launches.
    rng = random.Random(21)
    reqs = []
    for bound, first, last, square in ((90000, 1, 3, False), (3000, 5, 60, False),
                                       (40, 1, 200, False), (12000, 7, 9, False),
                                       (6000, 45000, 45010, True), (1, 1, 30, False)):
        coeffs, _ = curves.partially_dense(rng, rng.randint(6, 10), rng.choice([3, 4, 6]))
        reqs.append({"coefficients": coeffs, "bound": bound, "first": first, "last": last,
                     "square": square})
    results, stats = defaults(emu).sieve_many(reqs)
    assert stats["requeued"] == 0
    for r, (num, den) in zip(reqs, results):
        count, digest, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                                 r["last"], r["square"], method="words")
        assert np.array_equal(num, on) and np.array_equal(den, od), r["bound"]


def test_sieve_into_matches_sieve(emu):
    coeffs = ["1", "-3", "0", "2", "0", "0", "4"]
    defaults(emu)
    for bound, first, last, cap in ((300000, 1, 20, 4096), (3000, 1, 30, 4096)):
        ref = emu.sieve(coeffs, bound, first, last)
        num = np.zeros(cap, np.int64)
        den = np.zeros(cap, np.int32)
        rc, n = emu.sieve_into(coeffs, bound, first, last, num.ctypes.data, den.ctypes.data, cap)
        order = np.lexsort((num[:n], den[:n]))
        assert rc == 0 and n == ref.survivor_count
        assert np.array_equal(num[:n][order], ref.numerators)
        assert np.array_equal(den[:n][order], ref.denominators)


def test_streaming_contract_through_reference_signature():
    if not os.path.exists(CONTRACT):
        pytest.skip("streaming_contract_emu not built")
    env = dict(os.environ, **SCHEDULE)
    for bound, last, sites in (("5000", "2", 20002), ("30000", "3", 180003)):
        proc = subprocess.run([CONTRACT, bound, last], capture_output=True, text=True,
                              timeout=600, env=env)
        assert proc.returncode == 0, proc.stderr
        ok, got, flushes = proc.stdout.split()
        assert ok == "ok" and int(got) == sites and int(flushes) >= 2


def test_mode_coverage_trace_through_serving_loop(workloads, tmp_path):
    # Square-denominator cubics and quartics and degree 50-100 curves with
    # 1,024-bit coefficients, intervals and infinity, all at once through
    # the serving loop's many-curve rounds, then one at a time.
    if not os.path.exists(EMU_SERVE):
        pytest.skip("ratpoints_serve_emu not built")
    import test_serve
    reqs = test_serve.mode_trace(workloads, 8, square_k=40, square_bound=60, high_height=12)
    trace = tmp_path / "trace.txt"
    workloads.write_trace(str(trace), reqs)
    env = dict(os.environ, **SCHEDULE)
    results = {}
    for mode in ("serve", "alone"):
        out = tmp_path / f"{mode}.txt"
        proc = subprocess.run([EMU_SERVE, "--trace", str(trace), "--mode", mode, "--out",
                               str(out), "--round-sites", "3000", "--max-active", "8"],
                              capture_output=True, text=True, timeout=900, env=env)
        assert proc.returncode == 0, proc.stderr
        results[mode] = test_serve.parse_results(out)
    test_serve.check(reqs, results, oracle_limit=10 ** 9)


def test_mode_coverage_trace_at_full_line_shapes(workloads, tmp_path):
    # The mode-coverage serving trace with its full line shapes: square-
    # denominator cubics and quartics at A = 1e5 (lines of 3,126 words, row
    # runs) and degree 50-100 curves at height 3,000 (lines of 94 words,
    # short-line tiles) share every many-curve launch, with the tiny curves
    # besides; fewer denominators than the full trace (k <= 200, b <= 300)
    # and several rounds.  Every launch must finish, with the same points in
    # both modes and the oracle's points for the tiny requests.
    if not os.path.exists(EMU_SERVE):
        pytest.skip("ratpoints_serve_emu not built")
    import test_serve
    reqs = test_serve.mode_trace(workloads, 24, square_k=200, square_bound=100_000,
                                 high_height=3000)
    for r in reqs:
        if not r.get("square"):
            r["last"] = 300
    small = test_serve.mode_trace(workloads, 8, square_k=40, square_bound=60, high_height=12)
    for k, r in enumerate(small):
        r["id"] = 100 + k
    trace = tmp_path / "trace.txt"
    workloads.write_trace(str(trace), reqs + small)
    env = dict(os.environ, **SCHEDULE)
    results = {}
    for mode in ("serve", "alone"):
        out = tmp_path / f"{mode}.txt"
        proc = subprocess.run([EMU_SERVE, "--trace", str(trace), "--mode", mode, "--out",
                               str(out), "--round-sites", "1e8", "--max-active", "32",
                               "--progress"],
                              capture_output=True, text=True, timeout=900, env=env)
        assert proc.returncode == 0, proc.stderr
        if mode == "serve":
            assert proc.stderr.count("serve: round") >= 3, proc.stderr
        results[mode] = test_serve.parse_results(out)
    test_serve.check(reqs + small, results, oracle_limit=10_000)


def test_run_lengths_match_oracle(emu, oracle_lib):
    # Whole-line runs of other lengths than a variant's own (the "run_log2"
    # option the autotuner sweeps), plain and many-curve paths.
    coeffs, _ = curves.partially_dense(random.Random(17), 7, 5)
    ref = oracle_lib.sieve(coeffs, 30000, 11, 60, method="words")
    for name in ("q_u1_f12_w8", "w_u2_f14_w16"):
        for run_log2 in (10, 13):
            defaults(emu).set_option("variant", variants(emu).index(name))
            emu.set_option("run_log2", run_log2)
            emu.set_option("min_tiles", 1)
            assert same(emu.sieve(coeffs, 30000, 11, 60), ref), (name, run_log2)
            results, stats = emu.sieve_many([{"coefficients": coeffs, "bound": 30000,
                                              "first": 11, "last": 60}])
            num, den = results[0]
            assert np.array_equal(num, ref[2]) and np.array_equal(den, ref[3]), (name, run_log2)
    defaults(emu)
