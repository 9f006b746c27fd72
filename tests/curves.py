"""Curve constructions for parity tests (checker-side helpers)."""

import math

import oracle


def dense_product(count=32):
    """Product of the `count` largest primes below 512: f = 1 + P x makes
    every site survive (tests/test_sieve_streaming.cpp:51-63 in the
    reference)."""
    return math.prod(oracle.primes_below(512)[-count:])


def partially_dense(rng, live_primes, degree, coeff_bound=1000):
    """f = g^2 + Q h with Q the product of all candidates except
    `live_primes` of them: f is a square mod every factor of Q, so only the
    live primes reject sites and the survivor density is about
    2^-live.  Returns (coefficients, live prime list)."""
    live = rng.sample(oracle.candidate_primes(), live_primes)
    return with_live_primes(rng, live, degree, coeff_bound), live


def with_live_primes(rng, live, degree, coeff_bound=1000):
    """f = g^2 + Q h with Q the product of the candidates not in `live`:
    only the primes in `live` reject sites.  Returns the coefficients."""
    q = math.prod(p for p in oracle.candidate_primes() if p not in live)
    g_deg = degree // 2
    g = [rng.randint(-coeff_bound, coeff_bound) for _ in range(g_deg + 1)]
    g[-1] = g[-1] or 1
    h = [rng.randint(-coeff_bound, coeff_bound) for _ in range(degree + 1)]
    h[-1] = h[-1] or 1
    g2 = [0] * (2 * g_deg + 1)
    for i, a in enumerate(g):
        for j, b in enumerate(g):
            g2[i + j] += a * b
    f = [q * h[i] + (g2[i] if i < len(g2) else 0) for i in range(degree + 1)]
    if f[-1] == 0:
        f[-1] = q
    return [str(c) for c in f]
