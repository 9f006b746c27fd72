"""GPU parity of the HIP sieve against the CPU oracle, through the C ABI.

Bar: the survivor multiset is bit-identical (sorted arrays equal) to the
oracle's on every case the oracle finishes in seconds; at full sizes the
survivor counts equal the counts the reference holds, and size-independent
properties hold (monotonicity in A, soundness of every survivor checked by
Euler's criterion, completeness on windows checked exhaustively).
"""

import json
import math
import os
import random

import numpy as np
import pytest

import conftest

import curves
import oracle

pytestmark = pytest.mark.gpu

VARIANTS = [i for i, _ in conftest.compiled_variants()]
VARIANT_IDS = [name for _, name in conftest.compiled_variants()]

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROWS = {r["name"]: r for r in json.load(open(os.path.join(GOLDEN, "reference_survivor_counts.json")))}


def announce(**case):
    print("case:", " ".join(f"{k}={v}" for k, v in case.items()), flush=True)


def assert_same(gpu_result, oracle_result):
    count, digest, num, den = oracle_result
    assert gpu_result.survivor_count == count
    assert np.array_equal(gpu_result.numerators, num)
    assert np.array_equal(gpu_result.denominators, den)
    assert oracle.site_digest(gpu_result.numerators, gpu_result.denominators) == digest


def test_dense_streaming_exactly_once(fresh_gpu):
    # tests/test_sieve_streaming.cpp: 20,002 survivors through several
    # flushes of a buffer smaller than the survivor set.
    announce(case="dense", capacity=16384)
    fresh_gpu.set_option("survivor_capacity", 16384)
    res = fresh_gpu.sieve(["1", str(curves.dense_product())], 5000, 1, 2)
    assert res.survivor_count == 20002
    expected_num = np.concatenate([np.arange(-5000, 5001)] * 2)
    expected_den = np.repeat(np.array([1, 2], dtype=np.int32), 10001)
    assert np.array_equal(res.numerators, expected_num)
    assert np.array_equal(res.denominators, expected_den)
    assert res.stats["launches"] >= 2


def test_dense_many_flushes(fresh_gpu, oracle_lib):
    announce(case="dense-many-flushes", capacity=16384)
    fresh_gpu.set_option("survivor_capacity", 16384)
    coeffs = ["1", str(curves.dense_product())]
    res = fresh_gpu.sieve(coeffs, 40000, 3, 7)
    assert res.survivor_count == 5 * 80001
    assert res.stats["launches"] >= 5 * 80001 // 16384
    assert_same(res, oracle_lib.sieve(coeffs, 40000, 3, 7, method="words"))


FAST_GOLDEN = ["dense-1-P", "rank31-gpu-fixed-validation-p0", "rank31-gpu-fixed-validation-grid",
               "rank31-gpu-fixed-benchmark-1m-64", "rank31-gpu-fixed-p0-10m-64",
               "rank31-gpu-fixed-pd-100m-256", "rank31-gpu-fixed-s-1b-1024",
               "rank31-gpu-fixed-p0-2b-8192", "record-screen-10000000-1000000"]


@pytest.mark.parametrize("name", FAST_GOLDEN)
def test_reference_survivor_counts(fresh_gpu, name):
    row = ROWS[name]
    announce(case=name, A=row["bound"], b=f"{row['first']}..{row['last']}")
    rc, stats = fresh_gpu.run(row["coefficients"].split(), row["bound"], row["first"],
                              row["last"], row["square"], lambda n, d: False)
    assert stats["survivor_count"] == row["survivors"]


# The reference's remaining published counts, at their full sizes (up to
# 2.6e14 sites, a few seconds each): the record curve over 1e7 x 1e7 and the
# rank-31 cubics at A = 1e8 and A = 2e9 (lines of up to 62.5 million words,
# cut into chunks).
FULL_GOLDEN = sorted(set(ROWS) - set(FAST_GOLDEN))


@pytest.mark.parametrize("name", FULL_GOLDEN)
def test_reference_survivor_counts_full_size(fresh_gpu, name):
    row = ROWS[name]
    announce(case=name, A=row["bound"], b=f"{row['first']}..{row['last']}")
    rc, stats = fresh_gpu.run(row["coefficients"].split(), row["bound"], row["first"],
                              row["last"], row["square"], lambda n, d: False)
    assert stats["survivor_count"] == row["survivors"]


@pytest.mark.parametrize("seed", range(10))
def test_random_curves_match_oracle(fresh_gpu, oracle_lib, seed):
    rng = random.Random(seed)
    degree = rng.choice([1, 2, 3, 4, 5, 6, 7, 8, 10, 12])
    coeffs, live = curves.partially_dense(rng, rng.randint(3, 8), degree,
                                          coeff_bound=rng.choice([10, 10 ** 6, 10 ** 30]))
    bound = rng.choice([1, 2, 31, 63, 64, 65, 127, 128, 1000, rng.randint(1, 50000)])
    first = rng.choice([1, rng.randint(1, 10 ** 5), 2 ** 31 - 300])
    last = min(2 ** 31 - 1, first + rng.randint(0, 200))
    announce(seed=seed, degree=degree, A=bound, b=f"{first}..{last}")
    res = fresh_gpu.sieve(coeffs, bound, first, last)
    assert_same(res, oracle_lib.sieve(coeffs, bound, first, last, method="words"))


@pytest.mark.parametrize("seed", range(3))
def test_high_degree_curves_match_oracle(fresh_gpu, oracle_lib, seed):
    # Degree 50-100 with coefficients of about 1,024 bits (the mode-coverage
    # shapes): the residue tables come from 101 coefficient residues per
    # prime; 4-5 live primes keep the survivor sets large.
    rng = random.Random(300 + seed)
    degree = rng.choice([50, 63, 75, 100])
    coeffs, live = curves.partially_dense(rng, rng.randint(4, 5), degree,
                                          coeff_bound=1 << 700)
    assert len(coeffs) == degree + 1 and max(len(c) for c in coeffs) > 300
    bound = rng.choice([1500, 3000, 4096])
    first = rng.randint(1, 10 ** 6)
    last = first + 80
    announce(seed=seed, degree=degree, A=bound, b=f"{first}..{last}")
    res = fresh_gpu.sieve(coeffs, bound, first, last)
    want = oracle_lib.sieve(coeffs, bound, first, last, method="words")
    assert want[0] > 1000
    assert_same(res, want)


@pytest.mark.parametrize("seed", range(4))
def test_square_denominators_match_oracle(fresh_gpu, oracle_lib, seed):
    rng = random.Random(100 + seed)
    coeffs, _ = curves.partially_dense(rng, rng.randint(3, 7), rng.choice([3, 4]))
    first = rng.choice([1, 46340 - 150, rng.randint(1, 46000)])
    last = min(46340, first + 150)
    bound = rng.randint(1, 20000)
    announce(seed=seed, square=True, A=bound, k=f"{first}..{last}")
    res = fresh_gpu.sieve(coeffs, bound, first, last, square=True)
    assert_same(res, oracle_lib.sieve(coeffs, bound, first, last, square=True, method="words"))


def test_square_mode_golden_point(fresh_gpu):
    # The point (1 : 14 : 4) of y^2 = 1 - x + x^3 at k = 2 must survive.
    announce(case="square-golden")
    res = fresh_gpu.sieve(["1", "-1", "0", "1"], 10, 1, 3, square=True)
    pairs = set(zip(res.numerators.tolist(), res.denominators.tolist()))
    assert (1, 4) in pairs and all(d in (1, 4, 9) for _, d in pairs)


def test_divisible_denominators(fresh_gpu, oracle_lib):
    # Denominators divisible by one to three candidate primes use a
    # different prime set.
    coeffs, _ = curves.partially_dense(random.Random(5), 6, 6)
    for b in (509, 293 * 307, 293 * 307 * 311, 499 * 503 * 509 - 1, 2 * 3 * 467 * 499):
        announce(case="divisible", b=b)
        res = fresh_gpu.sieve(coeffs, 30000, b, b + 2)
        assert_same(res, oracle_lib.sieve(coeffs, 30000, b, b + 2, method="words"))


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_variants_agree(fresh_gpu, oracle_lib, pkg, variant):
    coeffs, _ = curves.partially_dense(random.Random(77), 7, 6)
    announce(case="variant", variant=pkg.variants()[variant])
    fresh_gpu.set_option("variant", variant)
    res = fresh_gpu.sieve(coeffs, 70001, 11, 140)
    assert_same(res, oracle_lib.sieve(coeffs, 70001, 11, 140, method="words"))


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_variants_resume_exactly_once(fresh_gpu, oracle_lib, pkg, variant):
    # About a quarter of all sites survive (two live primes), so every wave
    # overflows a 16,384-survivor buffer many times and resumes from its
    # cursor; the multiset must still be exact.
    coeffs, _ = curves.partially_dense(random.Random(78 + variant), 2, 4)
    announce(case="variant-resume", variant=pkg.variants()[variant], capacity=16384)
    fresh_gpu.set_option("variant", variant)
    fresh_gpu.set_option("survivor_capacity", 16384)
    res = fresh_gpu.sieve(coeffs, 40000, 3, 22)
    assert res.stats["launches"] >= res.survivor_count // 16384
    assert_same(res, oracle_lib.sieve(coeffs, 40000, 3, 22, method="words"))


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_runs_of_lines(fresh_gpu, oracle_lib, pkg, variant):
    # min_tiles = 1 lets lines shorter than 2^17 words share a tile: 2,501-word
    # lines go 52 to a tile, so b = 7..140 is two full runs and a partial
    # one, each line's strings loaded while the previous line is swept.
    coeffs, _ = curves.partially_dense(random.Random(90 + variant), 7, 6)
    announce(case="runs", variant=pkg.variants()[variant])
    fresh_gpu.set_option("variant", variant)
    fresh_gpu.set_option("min_tiles", 1)
    res = fresh_gpu.sieve(coeffs, 80000, 7, 140)
    assert_same(res, oracle_lib.sieve(coeffs, 80000, 7, 140, method="words"))
    # 94-word lines on row tiles (short-line kernel off): all 200 square
    # denominators in one run.
    fresh_gpu.set_option("direct_words", 0)
    res = fresh_gpu.sieve(coeffs, 3000, 1, 200, square=True)
    assert_same(res, oracle_lib.sieve(coeffs, 3000, 1, 200, square=True, method="words"))


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_run_lengths(fresh_gpu, oracle_lib, pkg, variant):
    # Run lengths other than the variant's own (the "run_log2" option the
    # autotuner sweeps): 2,501-word lines in runs of 1, 3 and 13 lines, on
    # the plain and the many-curve path.
    coeffs, _ = curves.partially_dense(random.Random(190 + variant), 7, 6)
    ref = oracle_lib.sieve(coeffs, 80000, 7, 140, method="words")
    fresh_gpu.set_option("variant", variant)
    fresh_gpu.set_option("min_tiles", 1)
    for run_log2 in (11, 13, 15):
        announce(case="run-length", variant=pkg.variants()[variant], run_log2=run_log2)
        fresh_gpu.set_option("run_log2", run_log2)
        assert_same(fresh_gpu.sieve(coeffs, 80000, 7, 140), ref)
        results, _ = fresh_gpu.sieve_many([{"coefficients": coeffs, "bound": 80000,
                                            "first": 7, "last": 140}])
        assert np.array_equal(results[0][0], ref[2]) and np.array_equal(results[0][1], ref[3])


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_runs_resume_exactly_once(fresh_gpu, oracle_lib, pkg, variant):
    # Runs of 104 lines of 1,251 words with about a quarter of all sites
    # surviving: waves stop mid-run on a full 16,384-survivor buffer and
    # resume at their (line, word) cursors.
    coeffs, _ = curves.partially_dense(random.Random(120 + variant), 2, 5)
    announce(case="runs-resume", variant=pkg.variants()[variant], capacity=16384)
    fresh_gpu.set_option("variant", variant)
    fresh_gpu.set_option("min_tiles", 1)
    fresh_gpu.set_option("survivor_capacity", 16384)
    res = fresh_gpu.sieve(coeffs, 40000, 300, 450)
    assert res.stats["launches"] >= res.survivor_count // 16384
    assert_same(res, oracle_lib.sieve(coeffs, 40000, 300, 450, method="words"))


@pytest.mark.parametrize("bound", [3000, 20000])
def test_short_line_runs(fresh_gpu, oracle_lib, bound):
    # min_tiles = 1: a short-line tile takes a run of up to 2^16 words of
    # lines, swept in blocks of 16 (94-word lines: 300 lines in one tile, 19
    # blocks; 626-word lines: 105 lines per tile).
    coeffs, _ = curves.partially_dense(random.Random(140 + bound), 7, 6)
    announce(case="short-runs", A=bound)
    fresh_gpu.set_option("min_tiles", 1)
    res = fresh_gpu.sieve(coeffs, bound, 5, 304)
    assert_same(res, oracle_lib.sieve(coeffs, bound, 5, 304, method="words"))
    res = fresh_gpu.sieve(coeffs, bound, 1, 150, square=True)
    assert_same(res, oracle_lib.sieve(coeffs, bound, 1, 150, square=True, method="words"))


def test_short_line_runs_resume(fresh_gpu, oracle_lib):
    # Dense survivors in runs of short lines: waves stop mid-run on a full
    # 16,384-survivor buffer and resume at their (block, pair) cursors.
    coeffs, _ = curves.partially_dense(random.Random(150), 2, 6)
    announce(case="short-runs-resume", capacity=16384)
    fresh_gpu.set_option("min_tiles", 1)
    fresh_gpu.set_option("survivor_capacity", 16384)
    res = fresh_gpu.sieve(coeffs, 5000, 2, 401)
    assert res.stats["launches"] >= res.survivor_count // 16384
    assert_same(res, oracle_lib.sieve(coeffs, 5000, 2, 401, method="words"))


def test_line_split_into_chunks(fresh_gpu, oracle_lib):
    # Few denominators and a long line: each line is cut into many tiles.
    coeffs, _ = curves.partially_dense(random.Random(8), 9, 4)
    announce(case="chunks", A=3_000_001, b="1..3")
    res = fresh_gpu.sieve(coeffs, 3_000_001, 1, 3)
    assert res.stats["launches"] == 1
    assert_same(res, oracle_lib.sieve(coeffs, 3_000_001, 1, 3, method="words"))


def test_small_batches(fresh_gpu, oracle_lib):
    coeffs, _ = curves.partially_dense(random.Random(9), 6, 5)
    fresh_gpu.set_option("batch_size", 7)
    announce(case="batch7")
    res = fresh_gpu.sieve(coeffs, 5000, 1, 50)
    assert res.stats["batches"] == 8
    assert_same(res, oracle_lib.sieve(coeffs, 5000, 1, 50, method="words"))


def test_huge_bound_properties(fresh_gpu):
    # A = 2^40 + 12345 (3.4e10 words per line, so 64-bit word indices and
    # line chunks are exercised).  Checks: a sample of survivors and every
    # survivor near the line ends pass Euler's criterion (soundness); the
    # windows |a| <= 2000 and |a| >= A - 2000 are exhaustively complete; and
    # a second run with A2 = A - 777 (a different word alignment) gives the
    # same survivors on |a| <= A2.
    rng = random.Random(12)
    coeffs, _ = curves.partially_dense(rng, 22, 3)
    ints = [int(c) for c in coeffs]
    bound = 2 ** 40 + 12345
    announce(case="huge-bound", A=bound)
    res = fresh_gpu.sieve(coeffs, bound, 5, 6)
    num, den = res.numerators, res.denominators
    assert 0 < res.survivor_count < 20_000_000
    assert np.all(np.abs(num) <= bound)

    def passes(a, b):
        v = oracle.homogenised_value(ints, a, b)
        return all(pow(v % p, (p - 1) // 2, p) in (0, 1) for p in oracle.selected_primes(b))

    picks = np.random.default_rng(0).choice(len(num), size=min(2000, len(num)), replace=False)
    ends = np.nonzero(np.abs(num) > bound - 10 ** 7)[0]
    for i in list(picks) + list(ends):
        assert passes(int(num[i]), int(den[i]))
    got = set(zip(num[np.abs(num) <= 2000].tolist(), den[np.abs(num) <= 2000].tolist()))
    got |= set(zip(num[np.abs(num) >= bound - 2000].tolist(),
                   den[np.abs(num) >= bound - 2000].tolist()))
    for b in (5, 6):
        for a in list(range(-2000, 2001)) + list(range(bound - 2000, bound + 1)) \
                + list(range(-bound, -bound + 2001)):
            assert passes(a, b) == ((a, b) in got)
    bound2 = bound - 777
    res2 = fresh_gpu.sieve(coeffs, bound2, 5, 6)
    keep = np.abs(num) <= bound2
    assert np.array_equal(num[keep], res2.numerators)
    assert np.array_equal(den[keep], res2.denominators)


def test_monotone_in_bound(fresh_gpu):
    # The predicate does not depend on A, so S(A') = S(A) restricted to
    # |a| <= A'.
    coeffs, _ = curves.partially_dense(random.Random(13), 10, 6)
    announce(case="monotone")
    big = fresh_gpu.sieve(coeffs, 2_000_000, 100, 130)
    small = fresh_gpu.sieve(coeffs, 777_777, 100, 130)
    keep = np.abs(big.numerators) <= 777_777
    assert np.array_equal(big.numerators[keep], small.numerators)
    assert np.array_equal(big.denominators[keep], small.denominators)


def test_planted_points_survive(fresh_gpu, workloads):
    coeffs, points = workloads.sextic_a5e6_curve()
    announce(case="planted", A=5_000_000)
    res = fresh_gpu.sieve(coeffs, 5_000_000, 1, 2000)
    got = set(zip(res.numerators.tolist(), res.denominators.tolist()))
    for a, b in points:
        if b <= 2000:
            assert (a, b) in got


def test_cancel_stops(fresh_gpu):
    announce(case="cancel")
    fresh_gpu.set_option("survivor_capacity", 16384)
    fresh_gpu.set_option("batch_size", 1)
    calls = []
    rc, stats = fresh_gpu.run(["1", str(curves.dense_product())], 100000, 1, 50, False,
                              lambda n, d: calls.append(len(n)) or True)
    assert rc == 4 and len(calls) == 1


def test_invalid_requests(fresh_gpu, pkg):
    for args in ((["1", "1"], 10, 0, 5), (["1", "1"], 10, 6, 5), (["1", "1"], 0, 1, 5),
                 (["1"], 10, 1, 5), (["1", "x"], 10, 1, 5)):
        with pytest.raises(pkg.SieveError) as err:
            fresh_gpu.run(*args, sink=lambda n, d: False)
        assert err.value.code == 1
    with pytest.raises(pkg.SieveError) as err:
        fresh_gpu.run(["1", "1"], 10, 1, 46341, True, lambda n, d: False)
    assert str(err.value) == "invalid denominator range"
    # Heights whose lines exceed 2^47 words are refused up front, on the
    # plain and the many-curve path alike.
    with pytest.raises(pkg.SieveError) as err:
        fresh_gpu.run(["1", "1"], 2 ** 52, 1, 1, False, lambda n, d: False)
    assert str(err.value) == "maximum height exceeds GPU limit"
    huge = {"coefficients": ["1", "0", "1"], "bound": 2 ** 52, "first": 1, "last": 1}
    small = {"coefficients": ["1", "0", "1"], "bound": 100, "first": 1, "last": 3}
    with pytest.raises(pkg.SieveError) as err:
        fresh_gpu.sieve_many([small, huge])
    assert str(err.value) == "maximum height exceeds GPU limit"


def test_curve_cache_switch(fresh_gpu, oracle_lib):
    # Alternating curves must rebuild the per-curve table each time.
    rng = random.Random(21)
    c1, _ = curves.partially_dense(rng, 6, 4)
    c2, _ = curves.partially_dense(rng, 6, 6)
    for coeffs in (c1, c2, c1):
        res = fresh_gpu.sieve(coeffs, 4000, 1, 40)
        assert_same(res, oracle_lib.sieve(coeffs, 4000, 1, 40, method="words"))


@pytest.mark.parametrize("bound", [24577, 40000, 65503])
def test_short_line_kernel_up_to_2048_words(fresh_gpu, oracle_lib, bound):
    # The short-line kernel takes lines of up to 2,048 words when asked to
    # (direct_words up to its limit): lines of 769-2,048 words against the
    # oracle, and the option's bound.
    coeffs, _ = curves.partially_dense(random.Random(bound), 6, 6)
    first, last = 3, 40
    announce(case="short-line-kernel", A=bound, b=f"{first}..{last}")
    fresh_gpu.set_option("direct_words", 2048)
    assert_same(fresh_gpu.sieve(coeffs, bound, first, last),
                oracle_lib.sieve(coeffs, bound, first, last, method="words"))
    with pytest.raises(Exception):
        fresh_gpu.set_option("direct_words", 2049)


@pytest.mark.parametrize("bound", [1, 31, 32, 100, 1000, 9000, 24543])
def test_short_lines_both_modes_match_oracle(fresh_gpu, oracle_lib, bound):
    # Lines of at most 2,048 words go to the short-line kernel; the same input
    # through the row kernel (direct_words = 0) and the oracle must agree.
    coeffs, _ = curves.partially_dense(random.Random(bound), 6, 6)
    first, last = 100, 100 + max(40, 40000 // (2 * bound + 1))
    announce(case="short-lines", A=bound, b=f"{first}..{last}")
    expected = oracle_lib.sieve(coeffs, bound, first, last, method="words")
    res = fresh_gpu.sieve(coeffs, bound, first, last)
    assert_same(res, expected)
    fresh_gpu.set_option("direct_words", 0)
    assert_same(fresh_gpu.sieve(coeffs, bound, first, last), expected)


def test_short_lines_many_denominators(fresh_gpu, oracle_lib, workloads):
    # The planted sextic on [-1000, 1000] x [1, 1000] in full.
    coeffs, points = workloads.sextic_a1000_curve()
    announce(case="sextic-A1000")
    res = fresh_gpu.sieve(coeffs, 1000, 1, 1000)
    assert_same(res, oracle_lib.sieve(coeffs, 1000, 1, 1000, method="words"))
    got = set(zip(res.numerators.tolist(), res.denominators.tolist()))
    assert all(p in got for p in points)


def test_short_lines_streaming_resume(fresh_gpu, oracle_lib):
    # Dense 1 + P x on short lines over many tiles with a small buffer.  The
    # line is fully dense only where the 32 selected primes are P's; the
    # denominators 331..397 in range are candidate primes themselves, bring
    # in 313 and lose about half their sites, so the oracle is the reference.
    announce(case="short-dense-resume")
    fresh_gpu.set_option("survivor_capacity", 16384)
    coeffs = ["1", str(curves.dense_product())]
    res = fresh_gpu.sieve(coeffs, 300, 1, 400)
    expected = oracle_lib.sieve(coeffs, 300, 1, 400, method="words")
    assert expected[0] > 230_000
    assert res.stats["launches"] >= expected[0] // 16384
    assert_same(res, expected)


def test_short_lines_square_mode(fresh_gpu, oracle_lib):
    coeffs, _ = curves.partially_dense(random.Random(31), 5, 4)
    announce(case="short-square")
    res = fresh_gpu.sieve(coeffs, 700, 46000, 46340, square=True)
    assert_same(res, oracle_lib.sieve(coeffs, 700, 46000, 46340, square=True, method="words"))


def test_streaming_contract_through_reference_signature():
    # ratpoints_gpu::run_modular_sieve called from C++ with the reference's
    # signature (tests/test_sieve_streaming.cpp:45-80 in the reference): every
    # site of 1 + P x delivered exactly once, over a 16,384-entry buffer so
    # the stream comes in several flushes.
    binary = os.path.join(os.path.dirname(__file__), "cpp", "_build", "streaming_contract")
    assert os.path.exists(binary), "tests/cpp/_build/streaming_contract not built"
    for bound, last, sites in ((5000, 2, 20002), (60000, 8, 960008)):
        announce(case="streaming-contract", A=bound, B=last)
        proc = __import__("subprocess").run([binary, str(bound), str(last)], capture_output=True,
                                             text=True, timeout=120)
        assert proc.returncode == 0, proc.stderr
        ok, got, flushes = proc.stdout.split()
        assert ok == "ok" and int(got) == sites and int(flushes) >= 2
