"""Pins the CPU oracle (no GPU needed).

The oracle is trusted only after it reproduces
  * every reference-held survivor count a CPU reaches in seconds
    (tests/golden/reference_survivor_counts.json),
  * an independent Euler-criterion brute force of the survivor predicate on
    tiny boxes (oracle.survivors_legendre shares no code with the C oracle),
  * the reference's own src/sieve_plan.cpp (compiled into oracle/_ref) for
    prime selection, affine masks and modular inverses,
  * the published record-curve points for the exact stage,
and the two C evaluation methods (per-site and packed words) agree on seeded
random curves.
"""

import json
import math
import os
import random
from fractions import Fraction

import numpy as np
import pytest

import curves
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROWS = json.load(open(os.path.join(GOLDEN, "reference_survivor_counts.json")))
CPU_ROWS = [r for r in ROWS if (2 * r["bound"] + 1) * (r["last"] - r["first"] + 1) <= 6e10]


@pytest.mark.parametrize("row", CPU_ROWS, ids=[r["name"] for r in CPU_ROWS])
def test_golden_counts_words(oracle_lib, row):
    count, _, _, _ = oracle_lib.sieve(row["coefficients"].split(), row["bound"], row["first"],
                                      row["last"], row["square"], method="words",
                                      collect=False)
    assert count == row["survivors"]


SMALL_ROWS = [r for r in CPU_ROWS if (2 * r["bound"] + 1) * (r["last"] - r["first"] + 1) <= 2e8]


@pytest.mark.parametrize("row", SMALL_ROWS, ids=[r["name"] for r in SMALL_ROWS])
def test_golden_counts_direct(oracle_lib, row):
    count, _, _, _ = oracle_lib.sieve(row["coefficients"].split(), row["bound"], row["first"],
                                      row["last"], row["square"], method="direct",
                                      collect=False)
    assert count == row["survivors"]


def test_dense_case_every_site_once(oracle_lib):
    # tests/test_sieve_streaming.cpp:73-108: every site of 1 + P x survives.
    count, digest, num, den = oracle_lib.sieve(["1", str(curves.dense_product())], 5000, 1, 2)
    assert count == 20002
    expected = [(a, b) for b in (1, 2) for a in range(-5000, 5001)]
    assert list(zip(num.tolist(), den.tolist())) == expected
    assert digest == oracle.site_digest(num, den)


@pytest.mark.parametrize("seed", range(8))
def test_direct_matches_legendre(oracle_lib, seed):
    rng = random.Random(seed)
    degree = rng.choice([1, 2, 3, 4, 5, 6, 8])
    coeffs, _ = curves.partially_dense(rng, rng.randint(2, 5), degree)
    square = seed % 3 == 0
    first = rng.randint(1, 40) if square else rng.randint(1, 5000)
    last = first + rng.randint(0, 3)
    bound = rng.randint(1, 60)
    expected = oracle.survivors_legendre(coeffs, bound, first, last, square)
    count, digest, num, den = oracle_lib.sieve(coeffs, bound, first, last, square,
                                               method="direct")
    assert list(zip(num.tolist(), den.tolist())) == expected
    assert count == len(expected) > 0


def test_legendre_divisible_denominators(oracle_lib):
    # Denominators divisible by candidate primes change the selection
    # (src/sieve_plan.cpp:79-86).
    rng = random.Random(99)
    coeffs, _ = curves.partially_dense(rng, 4, 4)
    for b in (509, 509 * 503, 293 * 307 * 311, 2 * 467 * 499):
        expected = oracle.survivors_legendre(coeffs, 40, b, b)
        _, _, num, den = oracle_lib.sieve(coeffs, 40, b, b, method="direct")
        assert list(zip(num.tolist(), den.tolist())) == expected


@pytest.mark.parametrize("seed", range(12))
def test_words_matches_direct(oracle_lib, seed):
    rng = random.Random(1000 + seed)
    degree = rng.randint(1, 12)
    coeffs, _ = curves.partially_dense(rng, rng.randint(3, 9), degree, coeff_bound=10 ** 12)
    square = seed % 4 == 0
    if square:
        first = rng.randint(1, 46000)
        last = min(46340, first + rng.randint(0, 30))
    else:
        first = rng.choice([1, rng.randint(1, 10 ** 6), 2 ** 31 - 40])
        last = min(2 ** 31 - 1, first + rng.randint(0, 30))
    bound = rng.choice([1, 31, 32, 33, 63, 64, 65, rng.randint(1, 40000)])
    a = oracle_lib.sieve(coeffs, bound, first, last, square, method="direct")
    b = oracle_lib.sieve(coeffs, bound, first, last, square, method="words")
    assert a[0] == b[0]
    assert a[1] == b[1]
    assert (a[2] == b[2]).all() and (a[3] == b[3]).all()


def test_digest_matches_python(oracle_lib):
    rng = random.Random(5)
    coeffs, _ = curves.partially_dense(rng, 4, 6)
    count, digest, num, den = oracle_lib.sieve(coeffs, 3000, 1, 50, method="words")
    assert count > 1000
    assert digest == oracle.site_digest(num, den)


def test_selection_never_divides(oracle_lib):
    # tests/test_sieve_plan.cpp:10-25: the boundary ranges of the reference.
    for first, last in ((1, 1), (65535, 65538), (2 ** 31 - 3, 2 ** 31 - 1)):
        for b in range(first, last + 1):
            sel = oracle_lib.selected_primes(b)
            assert len(sel) == 32 and all(b % p for p in sel)
            assert sel == oracle.selected_primes(b)
            assert sel == sorted(sel, reverse=True)


def test_candidate_primes():
    primes = oracle.candidate_primes()
    assert primes[0] == 293 and primes[-1] == 509 and len(primes) == 36
    # At most three candidates divide any b < 2^31: 293*307*311*313 > 2^31.
    assert 293 * 307 * 311 * 313 > 2 ** 31


def _ref_lib():
    try:
        return oracle.RefPlanLib()
    except RuntimeError:
        pytest.skip("oracle/_ref not built (reference sources absent)")


def test_selection_matches_reference_plan(oracle_lib):
    ref = _ref_lib()
    rng = random.Random(7)
    ranges = [(1, 1), (65535, 65538), (2 ** 31 - 3, 2 ** 31 - 1), (1, 3000)]
    ranges += [(x, x + 50) for x in (rng.randint(1, 2 ** 31 - 100) for _ in range(20))]
    ranges += [(509 * 503 * 499 - 5, 509 * 503 * 499 + 5)]
    for first, last in ranges:
        got = ref.selected_primes(first, last)
        for i, b in enumerate(range(first, last + 1)):
            assert list(got[i]) == oracle_lib.selected_primes(b)
    sq = ref.selected_primes(46000, 46340, square=True)
    for i, t in enumerate(range(46000, 46341)):
        assert list(sq[i]) == oracle_lib.selected_primes(t * t)


def test_affine_mask_matches_reference_plan(oracle_lib):
    ref = _ref_lib()
    rng = random.Random(11)
    polys = [["1", "0", "1"], ["-3", "5", "-7", "11", "2"], ["+1", "0", "+1"],
             ["1"] + ["0"] * 99 + ["1"],
             [str(rng.randint(-2 ** 1024, 2 ** 1024)) for _ in range(101)],
             ["0", "0", "-1"]]
    for coeffs in polys:
        for p in oracle.candidate_primes():
            assert (ref.affine_mask(p, coeffs) == oracle_lib.affine_mask(p, coeffs)).all()


def test_affine_mask_definition(oracle_lib):
    # chi_p(x) = [f(x) mod p is a square or zero], evaluated with Python ints.
    rng = random.Random(3)
    coeffs = [rng.randint(-10 ** 40, 10 ** 40) for _ in range(7)]
    for p in (293, 401, 509):
        squares = {y * y % p for y in range(p)}
        expected = [int(sum(c * x ** i for i, c in enumerate(coeffs)) % p in squares)
                    for x in range(p)]
        assert oracle_lib.affine_mask(p, coeffs).tolist() == expected


def test_mod_inverse(oracle_lib):
    ref = None
    try:
        ref = oracle.RefPlanLib()
    except RuntimeError:
        pass
    for p in oracle.candidate_primes():
        for v in range(1, p):
            inv = oracle_lib.lib.rpo_mod_inverse(v, p)
            assert v * inv % p == 1
            if ref is not None:
                assert ref.lib.ref_mod_inverse(v, p) == inv


def test_decimal_mod(oracle_lib):
    for text in ("0", "-0", "+17", "-1", "123456789012345678901234567890", "-" + "9" * 300):
        for p in (293, 509):
            assert oracle_lib.lib.rpo_decimal_mod(text.encode(), p) == int(text) % p


def test_exact_points_record_curve_published():
    # tests/verify_record_curve.py + tests/data/record_curve_x.txt: the
    # published x-coordinates of Stoll's record curve within a small box.
    published = {Fraction(line) for line in
                 open(os.path.join(GOLDEN, "record_curve_x.txt")).read().split()}
    coeffs = [247747600, -985905640, 567207969, 2396040466, 52485681, -470135160, 82342800]
    height = 60
    expected = {x for x in published if abs(x.numerator) <= height and x.denominator <= height}
    got = {Fraction(a, b) for a, _, b in oracle.exact_points(coeffs, height, 1, height)}
    assert got == expected and len(expected) >= 10


def test_square_mode_golden_line():
    # experiments/elliptic-curves/test_rank31_multifiber.py:75-81:
    # '1 -1 0 1' A=10, k in 1..3, --square-denominators -i prints "1 14 4".
    out = oracle.render_output(["1", "-1", "0", "1"], 10, 1, 3, square=True,
                               infinity=False, fmt="%x %y %z\\n").decode()
    assert "1 14 4\n" in out


def test_render_output_options():
    coeffs = ["1", "0", "1"]
    base = oracle.render_output(coeffs, 5, 1, 3).decode().splitlines()
    assert base[0] == "(1 : 1 : 0)" and base[1] == "(1 : -1 : 0)"
    assert oracle.render_output(coeffs, 5, 1, 3, first_only=True) == b"(1 : 1 : 0)\n"
    no_y = oracle.render_output(coeffs, 5, 1, 3, print_y=False).decode()
    assert "(0 : 1)" in no_y and ": -" not in no_y
    assert oracle.render_output(coeffs, 5, 1, 3, print_points=False, after="]") == b"]"
    fmt = oracle.render_output(coeffs, 5, 1, 1, infinity=False, fmt="%x,%y,%z",
                               before="[", between=";", after="]").decode()
    assert fmt == "[0,1,1;0,-1,1]"
    inter = oracle.render_output(coeffs, 5, 1, 3, infinity=False,
                                 intervals=[(Fraction(-1), Fraction(0))]).decode()
    assert all(Fraction(*map(int, l[1:-1].split(" : ")[::2])) <= 0 for l in inter.splitlines())


def test_planted_curves(workloads):
    for make in (workloads.sextic_a1000_curve, workloads.sextic_a5e6_curve):
        coeffs, points = make()
        assert len(coeffs) == 7 and workloads.is_squarefree(coeffs)
        for a, b in points:
            v = oracle.homogenised_value(coeffs, a, b)
            assert v >= 0 and math.isqrt(v) ** 2 == v


@pytest.mark.slow
def test_sextic_a1000_planted_points_survive(oracle_lib, workloads):
    coeffs, points = workloads.sextic_a1000_curve()
    c = workloads.SEXTIC_A1000
    _, _, num, den = oracle_lib.sieve(coeffs, c["bound"], c["first"], c["last"], method="words")
    found = set(zip(num.tolist(), den.tolist()))
    assert all(p in found for p in points)
    assert np.all(np.abs(num) <= c["bound"])
