"""Seeded synthetic workloads for the sieve benchmarks and parity tests.

Curves follow the planted construction f(x) = q(x)^2 + r * prod_i (b_i x - a_i):
at x = a_i / b_i the product vanishes, so F(a_i, b_i) = (b_i^m q(a_i/b_i))^2
is a perfect square and (a_i : y : b_i) is a rational point.  Random draws use
SplitMix64 so every workload is reproducible from its seed alone.
"""

import math
from fractions import Fraction


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo, hi):
        """Integer uniform in [lo, hi] (rejection sampling on 64-bit words)."""
        span = hi - lo + 1
        if span <= 0:
            raise ValueError("empty range")
        bits = span.bit_length()
        while True:
            value = 0
            for _ in range((bits + 63) // 64):
                value = (value << 64) | self.next()
            value &= (1 << bits) - 1
            if value < span:
                return lo + value


def poly_mul(p, q):
    out = [0] * (len(p) + len(q) - 1)
    for i, a in enumerate(p):
        if a:
            for j, b in enumerate(q):
                out[i + j] += a * b
    return out


def poly_add(p, q):
    n = max(len(p), len(q))
    return [(p[i] if i < len(p) else 0) + (q[i] if i < len(q) else 0) for i in range(n)]


def _trim(p):
    while p and p[-1] == 0:
        p.pop()
    return p


def is_squarefree(coeffs):
    """gcd(f, f') is constant, by Euclid over Q (the test of
    src/exact_polynomial.cpp:44-71 in the reference)."""
    f = _trim([Fraction(c) for c in coeffs])
    g = _trim([f[i] * i for i in range(1, len(f))])
    while g:
        r = list(f)
        while len(r) >= len(g):
            shift = len(r) - len(g)
            factor = r[-1] / g[-1]
            for i in range(len(g)):
                r[i + shift] -= factor * g[i]
            _trim(r)
            if not r:
                break
        f, g = g, r
    return len(f) == 1


def _poly_mod(p, m):
    return [c % m for c in p]


def _trim_mod(p):
    while p and p[-1] == 0:
        p.pop()
    return p


def _gcd_degree_mod(f, g, m):
    """Degree of gcd(f, g) over GF(m), m prime."""
    f = _trim_mod(_poly_mod(f, m))
    g = _trim_mod(_poly_mod(g, m))
    while g:
        inv = pow(g[-1], m - 2, m)
        r = list(f)
        while len(r) >= len(g) and r:
            shift = len(r) - len(g)
            factor = r[-1] * inv % m
            for i in range(len(g)):
                r[i + shift] = (r[i + shift] - factor * g[i]) % m
            _trim_mod(r)
        f, g = g, r
    return len(f) - 1


def squarefree_fast(coeffs, modulus=(1 << 61) - 1):
    """Squarefree over Q.  If f is squarefree modulo a prime that does not
    divide its leading coefficient, it is squarefree over Q; otherwise the
    exact test decides."""
    if coeffs[-1] % modulus and _gcd_degree_mod(
            coeffs, [coeffs[i] * i for i in range(1, len(coeffs))], modulus) == 0:
        return True
    return is_squarefree(coeffs)


def planted_curve(seed, q_degree, q_bound, point_bound, n_points, r_bound=100,
                  degree=None):
    """Draws f = q^2 + r * prod (b_i x - a_i) with coprime planted points.

    Returns (coefficients c0..cn as ints, points [(a_i, b_i)]).  Redraws
    until f is squarefree and has the expected degree."""
    rng = SplitMix64(seed)
    target = degree if degree is not None else 2 * q_degree
    while True:
        q = [rng.uniform(-q_bound, q_bound) for _ in range(q_degree + 1)]
        if q[-1] == 0:
            continue
        r = 0
        while r == 0:
            r = rng.uniform(-r_bound, r_bound)
        points = []
        while len(points) < n_points:
            a = rng.uniform(-point_bound, point_bound)
            b = rng.uniform(1, point_bound)
            if math.gcd(a, b) == 1 and (a, b) not in points:
                points.append((a, b))
        prod = [1]
        for a, b in points:
            prod = poly_mul(prod, [-a, b])
        f = poly_add(poly_mul(q, q), [r * c for c in prod])
        _trim(f)
        if len(f) - 1 != target:
            continue
        if is_squarefree(f):
            return f, points


def sextic_a1000_curve(seed=1):
    """Plain method, small: sextic, q coefficients in [-2^15, 2^15], six
    planted points with |a_i| <= 1000, 1 <= b_i <= 1000; box A = B = 1000."""
    return planted_curve(seed, 3, 2 ** 15, 1000, 6)


def sextic_a5e6_curve(seed=2):
    """Plain method at full size: q coefficients in [-2^30, 2^30], planted
    |a_i|, b_i <= 5e6 (coefficients of roughly 150 bits); box A = B = 5e6."""
    return planted_curve(seed, 3, 2 ** 30, 5_000_000, 6)


SEXTIC_A1000 = {"name": "planted-sextic-A1000-B1000", "seed": 1, "bound": 1000,
           "first": 1, "last": 1000}
SEXTIC_A5E6 = {"name": "planted-sextic-A5e6-B5e6", "seed": 2, "bound": 5_000_000,
           "first": 1, "last": 5_000_000, "batch": 65536}

RECORD_CURVE = [247747600, -985905640, 567207969, 2396040466, 52485681,
                -470135160, 82342800]


def sites(bound, first, last):
    return (2 * bound + 1) * (last - first + 1)


def _log_uniform_int(rng, lo, hi):
    """Integer log-uniform in [lo, hi] from 53 random bits."""
    u = (rng.next() >> 11) / float(1 << 53)
    return max(lo, min(hi, int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))))


def batched_curve(rng, degree, bits, n_points, bound, den_bound):
    """One curve of the batched workload.  Even degree d: f = q^2 +
    r prod_i (b_i x - a_i) with deg q = d/2 and the n_points planted factors.
    Odd degree d: q has degree (d-1)/2 and the product is padded with random
    linear factors to degree d, so that deg f = d (q^2 alone would have degree
    d + 1).  q's coefficients are drawn with bits/2 bits so q^2 has about
    `bits`-bit coefficients.  Planted points lie inside the search box; for
    odd degree at most d of them fit the construction, so n_points is capped
    at d."""
    if degree % 2 == 1:
        n_points = min(n_points, degree)
    half = max(1, bits // 2)
    while True:
        q_deg = degree // 2 if degree % 2 == 0 else (degree - 1) // 2
        q = [rng.uniform(-(1 << half), 1 << half) for _ in range(q_deg + 1)]
        if q[-1] == 0:
            continue
        r = 0
        while r == 0:
            r = rng.uniform(-100, 100)
        points = []
        while len(points) < n_points:
            a = rng.uniform(-bound, bound)
            b = rng.uniform(1, den_bound)
            if math.gcd(a, b) == 1 and (a, b) not in points:
                points.append((a, b))
        factors = [[-a, b] for a, b in points]
        if degree % 2 == 1:
            while len(factors) < degree:
                factors.append([rng.uniform(-1000, 1000), rng.uniform(1, 1000)])
        prod = [1]
        for fac in factors:
            prod = poly_mul(prod, fac)
        f = poly_add(poly_mul(q, q), [r * c for c in prod])
        _trim(f)
        if len(f) - 1 == degree and squarefree_fast(f):
            return f, points


def batched_curves(seed=3, count=4096):
    """Batched form: degrees uniform over {3, 4, 5, 6, 8, 10}, coefficient
    size uniform in 8..256 bits, 0-4 planted points, numerator and
    denominator bounds independently log-uniform in [1e3, 1e6].
    Returns a list of dicts {coefficients, bound, first, last, points}."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        degree = (3, 4, 5, 6, 8, 10)[rng.uniform(0, 5)]
        bits = rng.uniform(8, 256)
        n_points = rng.uniform(0, 4)
        bound = _log_uniform_int(rng, 1000, 1_000_000)
        den = _log_uniform_int(rng, 1000, 1_000_000)
        coeffs, points = batched_curve(rng, degree, bits, n_points, bound, den)
        out.append({"coefficients": coeffs, "bound": bound, "first": 1, "last": den,
                    "points": points, "degree": degree})
    return out


def _unit(rng):
    return (rng.next() >> 11) / float(1 << 53)


def serving_requests(seed=5, count=64, qps=2.0, min_height=100, max_height=10_000_000):
    """Serving mix: Poisson arrivals at `qps` requests/s; degree uniform in
    2..12, coefficient size uniform in 16..512 bits, height H log-uniform in
    [min_height, max_height] (|a| <= H, 1 <= b <= H), 0-2 planted points
    (the batched construction, so first-point requests have points to stop
    at), 30% in -1 first-point mode, 10% with 1-3 closed x-intervals.
    Returns dicts {id, arrival_ms, coefficients, bound, first, last,
    first_point, intervals [(lo, hi) Fractions, None = infinite], points}."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - _unit(rng)) / qps * 1000.0
        degree = rng.uniform(2, 12)
        bits = rng.uniform(16, 512)
        height = _log_uniform_int(rng, min_height, max_height)
        n_points = rng.uniform(0, 2)
        coeffs, points = batched_curve(rng, degree, bits, n_points, height, height)
        first_point = _unit(rng) < 0.3
        intervals = []
        if _unit(rng) < 0.1:
            for _ in range(rng.uniform(1, 3)):
                lo = Fraction(rng.uniform(-40, 40), rng.uniform(1, 8))
                hi = lo + Fraction(rng.uniform(1, 60), rng.uniform(1, 8))
                intervals.append((None if _unit(rng) < 0.2 else lo,
                                  None if _unit(rng) < 0.2 else hi))
        out.append({"id": i, "arrival_ms": t, "coefficients": coeffs, "bound": height,
                    "first": 1, "last": height, "first_point": first_point,
                    "intervals": intervals, "points": points, "degree": degree})
    return out


def write_trace(path, requests, square=False, no_infinity=False):
    """Writes requests in the trace format of bin/ratpoints_serve."""
    with open(path, "w") as fh:
        for r in requests:
            flags = (1 if r.get("first_point") else 0) | (2 if r.get("square", square) else 0) \
                | (4 if r.get("no_infinity", no_infinity) else 0)
            ivs = r.get("intervals", [])
            parts = [str(r["id"]), "%.3f" % r["arrival_ms"], str(flags), str(r["bound"]),
                     str(r["first"]), str(r["last"]), str(len(ivs))]
            for lo, hi in ivs:
                parts += ["-inf" if lo is None else str(lo), "inf" if hi is None else str(hi)]
            parts.append(str(len(r["coefficients"])))
            parts += [str(c) for c in r["coefficients"]]
            fh.write(" ".join(parts) + "\n")


def _random_intervals(rng, count):
    out = []
    for _ in range(count):
        lo = Fraction(rng.uniform(-60, 60), rng.uniform(1, 9))
        hi = lo + Fraction(rng.uniform(1, 80), rng.uniform(1, 9))
        out.append((lo, hi))
    return out


def mode_coverage_requests(seed=4, count=1024, square_fraction=0.5, square_k=46340,
                     square_bound=10_000_000, high_height=100_000):
    """Mode-coverage batch: `count` curves, all arriving at once.
    * square part: cubic and quartic curves with 64-128-bit coefficients,
      searched with square denominators b = k^2, k <= square_k, numerator
      bound square_bound; 0-2 planted points (a, k^2) with the batched
      construction;
    * high-degree part: degree 50-100, random 1,024-bit coefficients, height
      high_height (|a| <= H, b <= H), squarefree by construction check.
    Every curve carries 1-3 closed x-intervals; points at infinity enabled.
    Returns request dicts (see serving_requests) with key "square"."""
    rng = SplitMix64(seed)
    n_square = int(round(count * square_fraction))
    out = []
    for i in range(count):
        if i < n_square:
            degree = rng.uniform(3, 4)
            bits = rng.uniform(64, 128)
            n_points = rng.uniform(0, 2)
            ks = [rng.uniform(1, square_k) for _ in range(n_points)]
            coeffs, pts = batched_curve(rng, degree, bits, 0, square_bound, 1)
            # Plant (a, k^2): rebuild with factors (k^2 x - a).
            if n_points:
                half = max(1, bits // 2)
                while True:
                    q_deg = degree // 2 if degree % 2 == 0 else (degree - 1) // 2
                    q = [rng.uniform(-(1 << half), 1 << half) for _ in range(q_deg + 1)]
                    if q[-1] == 0:
                        continue
                    r = 0
                    while r == 0:
                        r = rng.uniform(-100, 100)
                    pts = []
                    for k in ks:
                        while True:
                            a = rng.uniform(-square_bound, square_bound)
                            if math.gcd(a, k * k) == 1:
                                break
                        pts.append((a, k * k))
                    factors = [[-a, b] for a, b in pts]
                    while len(factors) < degree and degree % 2 == 1:
                        factors.append([rng.uniform(-1000, 1000), rng.uniform(1, 1000)])
                    prod = [1]
                    for fac in factors:
                        prod = poly_mul(prod, fac)
                    f = poly_add(poly_mul(q, q), [r * c for c in prod])
                    _trim(f)
                    if len(f) - 1 == degree and squarefree_fast(f):
                        coeffs = f
                        break
            else:
                pts = []
            out.append({"id": i, "arrival_ms": 0.0, "coefficients": coeffs,
                        "bound": square_bound, "first": 1, "last": square_k, "square": True,
                        "first_point": False, "intervals": _random_intervals(rng, rng.uniform(1, 3)),
                        "points": pts, "degree": degree})
        else:
            degree = rng.uniform(50, 100)
            while True:
                coeffs = [rng.uniform(-(1 << 1023), 1 << 1023) for _ in range(degree + 1)]
                if coeffs[-1] != 0 and squarefree_fast(coeffs):
                    break
            out.append({"id": i, "arrival_ms": 0.0, "coefficients": coeffs,
                        "bound": high_height, "first": 1, "last": high_height, "square": False,
                        "first_point": False, "intervals": _random_intervals(rng, rng.uniform(1, 3)),
                        "points": [], "degree": degree})
    return out


def main(argv=None):
    """Writes a serving trace: python workloads.py serving OUT [COUNT QPS]
    or python workloads.py modes OUT."""
    import sys
    args = list(sys.argv[1:] if argv is None else argv)
    if len(args) < 2 or args[0] not in ("modes", "serving"):
        raise SystemExit("usage: workloads.py serving OUT [COUNT QPS] | modes OUT")
    if args[0] == "serving":
        count = int(args[2]) if len(args) > 2 else 100
        qps = float(args[3]) if len(args) > 3 else 5.0
        reqs = serving_requests(count=count, qps=qps)
    else:
        reqs = mode_coverage_requests()
    write_trace(args[1], reqs)
    print("wrote %d requests to %s" % (len(reqs), args[1]))


if __name__ == "__main__":
    main()
