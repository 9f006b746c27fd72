"""The persistent many-curve launch (rps_sieve_many) through the C ABI.

Bar: for every request, the survivor multiset equals what rps_sieve gives
for that request alone, and the CPU oracle's where the oracle is fast; a
request whose survivors overflow the shared buffer is searched again alone
and still comes out exact.
"""

import math
import random

import numpy as np
import pytest

import conftest

import curves
import oracle

pytestmark = pytest.mark.gpu

VARIANTS = [i for i, _ in conftest.compiled_variants()]
VARIANT_IDS = [name for _, name in conftest.compiled_variants()]


def announce(**case):
    print("case:", " ".join(f"{k}={v}" for k, v in case.items()), flush=True)


def random_requests(seed, count, max_bound=40000, max_den=300):
    rng = random.Random(seed)
    reqs = []
    for _ in range(count):
        degree = rng.choice([1, 2, 3, 4, 5, 6, 8, 10])
        coeffs, _ = curves.partially_dense(rng, rng.randint(6, 12), degree,
                                           coeff_bound=rng.choice([10, 10 ** 9, 10 ** 40]))
        square = rng.random() < 0.15
        bound = rng.choice([1, 63, 64, 65, rng.randint(1, 3000), rng.randint(1, max_bound)])
        if square:
            first = rng.randint(1, 46000)
            last = min(46340, first + rng.randint(0, max_den // 4))
        else:
            first = rng.choice([1, rng.randint(1, 10 ** 6), 2 ** 31 - 500])
            last = min(2 ** 31 - 1, first + rng.randint(0, max_den))
        reqs.append({"coefficients": coeffs, "bound": bound, "first": first, "last": last,
                     "square": square})
    return reqs


def plain(ctx, r):
    res = ctx.sieve(r["coefficients"], r["bound"], r["first"], r["last"], r.get("square", False))
    return res.numerators, res.denominators


@pytest.mark.parametrize("seed", range(3))
def test_many_matches_plain_and_oracle(fresh_gpu, oracle_lib, seed):
    reqs = random_requests(seed, 40)
    announce(case="many", seed=seed, n=len(reqs))
    results, stats = fresh_gpu.sieve_many(reqs)
    assert stats["requeued"] == 0
    total = 0
    for r, (num, den) in zip(reqs, results):
        pn, pd = plain(fresh_gpu, r)
        assert np.array_equal(num, pn) and np.array_equal(den, pd)
        if (2 * r["bound"] + 1) * (r["last"] - r["first"] + 1) < 3e7:
            count, digest, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                                     r["last"], r.get("square", False),
                                                     method="words")
            assert np.array_equal(num, on) and np.array_equal(den, od)
        total += len(num)
    assert stats["survivor_count"] == total


def test_many_long_lines_and_chunks(fresh_gpu):
    # Lines longer than 2^18 words are cut into several tiles per line.
    rng = random.Random(4)
    reqs = []
    for _ in range(6):
        coeffs, _ = curves.partially_dense(rng, 14, rng.choice([3, 4, 6]))
        reqs.append({"coefficients": coeffs, "bound": rng.randint(9_000_000, 20_000_000),
                     "first": rng.randint(1, 1000), "last": 0})
        reqs[-1]["last"] = reqs[-1]["first"] + rng.randint(0, 40)
    announce(case="many-long", n=len(reqs))
    results, stats = fresh_gpu.sieve_many(reqs)
    for r, (num, den) in zip(reqs, results):
        pn, pd = plain(fresh_gpu, r)
        assert np.array_equal(num, pn) and np.array_equal(den, pd)


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_many_overflow_requeues_exactly(fresh_gpu, oracle_lib, pkg, variant):
    # Dense requests overflow a 16,384-survivor buffer: they are searched
    # again alone; every request still matches the oracle.
    fresh_gpu.set_option("variant", variant)
    fresh_gpu.set_option("survivor_capacity", 16384)
    dense = ["1", str(curves.dense_product())]
    reqs = random_requests(11, 12, max_bound=3000, max_den=60)
    reqs.insert(3, {"coefficients": dense, "bound": 20000, "first": 1, "last": 2})
    reqs.insert(8, {"coefficients": dense, "bound": 300, "first": 1, "last": 100})
    announce(case="many-overflow", variant=pkg.variants()[variant], n=len(reqs))
    results, stats = fresh_gpu.sieve_many(reqs)
    assert stats["requeued"] >= 1
    for r, (num, den) in zip(reqs, results):
        count, digest, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                                 r["last"], r.get("square", False),
                                                 method="words")
        assert np.array_equal(num, on) and np.array_equal(den, od)


@pytest.mark.parametrize("capacity", [1 << 24, 16384])
def test_many_mixed_tile_families(fresh_gpu, oracle_lib, pkg, capacity):
    # With the variant from the tuning table, long-line requests (wide-row
    # tiles) and short ones run in two launches; survivors still map to the
    # caller's request indices, equal the plain path's and, with a small
    # buffer, overflowing requests are searched again exactly.
    fresh_gpu.set_option("survivor_capacity", capacity)
    rng = random.Random(51)
    reqs = random_requests(52, 10, max_bound=5000, max_den=80)
    for k in (1, 4, 7):
        coeffs, _ = curves.partially_dense(rng, 12, rng.choice([3, 4, 6]))
        first = rng.randint(1, 5000)
        reqs.insert(k, {"coefficients": coeffs, "bound": rng.randint(1_600_000, 3_000_000),
                        "first": first, "last": first + rng.randint(2, 12)})
    dense = ["1", str(curves.dense_product())]
    reqs.insert(5, {"coefficients": dense, "bound": 20000, "first": 1, "last": 2})
    words = [(2 * r["bound"] + 64) // 64 for r in reqs]
    assert {pkg.variants()[pkg.tuned_variant(w)][0] for w in words} >= {"q", "w"}
    announce(case="many-mixed", capacity=capacity, n=len(reqs))
    results, stats = fresh_gpu.sieve_many(reqs)
    if capacity == 16384:
        assert stats["requeued"] >= 1
    for r, (num, den) in zip(reqs, results):
        pn, pd = plain(fresh_gpu, r)
        assert np.array_equal(num, pn) and np.array_equal(den, pd)
    r = reqs[5]
    count, digest, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                             r["last"], False, method="words")
    assert np.array_equal(results[5][0], on) and np.array_equal(results[5][1], od)


def test_many_more_requests_than_one_launch_holds(fresh_gpu, oracle_lib):
    # More than 8,192 requests are cut into several launches (one residue
    # table per curve per launch); survivors still carry the caller's
    # request indices.  Dense requests at both ends make the mapping visible.
    rng = random.Random(61)
    dense = ["1", str(curves.dense_product())]
    reqs = []
    for i in range(8300):
        if i % 1000 == 7:
            reqs.append({"coefficients": dense, "bound": rng.randint(1, 40), "first": 1,
                         "last": rng.randint(1, 3)})
        else:
            coeffs, _ = curves.partially_dense(rng, 3, rng.choice([2, 3, 4]), coeff_bound=50)
            first = rng.randint(1, 500)
            reqs.append({"coefficients": coeffs, "bound": rng.randint(1, 40), "first": first,
                         "last": first + rng.randint(0, 3)})
    announce(case="many-8300", n=len(reqs))
    results, stats = fresh_gpu.sieve_many(reqs)
    assert stats["survivor_count"] == sum(len(n) for n, _ in results)
    for i in list(range(0, 8300, 97)) + [7, 8007, 8299]:
        r = reqs[i]
        count, digest, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                                 r["last"], False, method="words")
        assert np.array_equal(results[i][0], on) and np.array_equal(results[i][1], od), i


def test_many_batched_sample(fresh_gpu, workloads):
    # The batched workload's curves (bounds capped to keep the test short):
    # the shared launch equals the per-curve plain loop, and every planted
    # point survives.
    sample = workloads.batched_curves(count=96)
    for c in sample:
        c["bound"] = min(c["bound"], 30000)
        c["last"] = min(c["last"], 2000)
        c["points"] = [(a, b) for a, b in c["points"] if abs(a) <= c["bound"] and b <= c["last"]]
    announce(case="batched-sample", n=len(sample))
    results, stats = fresh_gpu.sieve_many(sample)
    for c, (num, den) in zip(sample, results):
        pn, pd = plain(fresh_gpu, c)
        assert np.array_equal(num, pn) and np.array_equal(den, pd)
        got = set(zip(num.tolist(), den.tolist()))
        assert all(p in got for p in c["points"])


def test_many_cancel_and_errors(fresh_gpu, pkg):
    reqs = random_requests(21, 5)
    calls = []
    rc, stats = fresh_gpu.run_many(reqs, lambda q, n, d: calls.append(len(n)) or True)
    assert rc == 4 or not calls
    bad = list(reqs) + [{"coefficients": ["1", "1"], "bound": 10, "first": 5, "last": 4}]
    with pytest.raises(pkg.SieveError) as err:
        fresh_gpu.run_many(bad)
    assert err.value.code == 1
    rc, stats = fresh_gpu.run_many([])
    assert rc == 0 and stats["survivor_count"] == 0


def test_many_table_cache_reuse_and_eviction(pkg):
    # Residue tables stay on the device between calls.  A fresh context with
    # 8 table slots: the second call reuses 3 cached curves and builds 6 (the
    # cache grows to the 9 distinct curves of one call), the third reuses
    # some and rebuilds evicted ones; every result equals a plain search.
    with pkg.SieveContext(0) as ctx:
        ctx.set_option("table_slots", 8)
        reqs = random_requests(31, 12)
        for part in (reqs[:6], reqs[3:12], reqs[:4] + reqs[10:12], reqs[:6]):
            announce(case="table-cache", n=len(part))
            results, stats = ctx.sieve_many(part)
            for r, (num, den) in zip(part, results):
                pn, pd = plain(ctx, r)
                assert np.array_equal(num, pn) and np.array_equal(den, pd)


def primitive(num, den):
    keep = np.array([math.gcd(int(a), int(b)) == 1 for a, b in zip(num, den)], dtype=bool)
    return num[keep], den[keep]


@pytest.mark.parametrize("seed", range(2))
def test_primitive_only_filters_on_device(fresh_gpu, oracle_lib, pkg, seed):
    # With primitive_only set, the device drops (a, b) with gcd(a, b) > 1:
    # the survivors equal the full survivor set filtered by gcd on the host,
    # on the plain path for every variant and on the many-curve path.
    reqs = random_requests(40 + seed, 16, max_bound=20000, max_den=200)
    full = [plain(fresh_gpu, r) for r in reqs]
    want = [primitive(n, d) for n, d in full]
    assert sum(len(n) for n, _ in full) > sum(len(n) for n, _ in want)
    variants = pkg.variants()
    fresh_gpu.set_option("primitive_only", 1)
    for v in range(len(variants)):
        fresh_gpu.set_option("variant", v)
        announce(case="primitive-plain", seed=seed, variant=variants[v])
        for r, (wn, wd) in zip(reqs, want):
            num, den = plain(fresh_gpu, r)
            assert np.array_equal(num, wn) and np.array_equal(den, wd)
    fresh_gpu.set_option("variant", -1)
    announce(case="primitive-many", seed=seed, n=len(reqs))
    results, stats = fresh_gpu.sieve_many(reqs)
    for r, (num, den), (wn, wd) in zip(reqs, results, want):
        assert np.array_equal(num, wn) and np.array_equal(den, wd)
    for r, (wn, wd) in zip(reqs[:4], want[:4]):
        if (2 * r["bound"] + 1) * (r["last"] - r["first"] + 1) < 3e7:
            count, digest, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                                     r["last"], r.get("square", False),
                                                     method="words")
            pn, pd = primitive(on, od)
            assert np.array_equal(wn, pn) and np.array_equal(wd, pd)


def test_many_variant_groups(fresh_gpu, oracle_lib, pkg):
    # Requests whose line lengths the many-curve column gives different
    # variants, each with more than the 1.5e11 sites that earn a launch of
    # its own, plus a small request of each length that joins its family's
    # largest group: every request's survivors equal the plain path's and,
    # for the small ones, the oracle's; one launch per distinct variant at
    # least.
    rng = random.Random(61)
    reqs, groups = [], set()
    for words in (1024, 2048, 4096, 16384, 65536):
        # 28 live primes: about 2^-28 of the sites survive, a few hundred per
        # large request, so no launch runs out of survivor buffer.
        coeffs, _ = curves.partially_dense(rng, 28, rng.choice([3, 4, 6]))
        bound = 32 * words - 32
        lines = int(1.6e11 / (2 * bound + 1)) + 1
        first = rng.randint(1, 1000)
        reqs.append({"coefficients": coeffs, "bound": bound, "first": first,
                     "last": first + lines - 1})
        reqs.append({"coefficients": coeffs, "bound": bound, "first": 7, "last": 9})
        groups.add(pkg.tuned(words, many=True)[0])
    assert len(groups) >= 3
    announce(case="many-variant-groups", n=len(reqs), groups=len(groups))
    results, stats = fresh_gpu.sieve_many(reqs)
    _, launches, _ = fresh_gpu.kernel_timing()
    assert launches >= len(groups)
    assert stats["requeued"] == 0
    for r, (num, den) in zip(reqs, results):
        pn, pd = plain(fresh_gpu, r)
        assert np.array_equal(num, pn) and np.array_equal(den, pd)
        if r["last"] - r["first"] < 10:
            count, _, on, od = oracle_lib.sieve(r["coefficients"], r["bound"], r["first"],
                                                r["last"], False, method="words")
            assert np.array_equal(num, on) and np.array_equal(den, od)
