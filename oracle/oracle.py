"""CPU oracle for the modular square sieve.  TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only as the checker (or the timed
CPU baseline).  The product path never imports it.

Contents (reference paths are relative to the reference repository):

* ``OracleLib`` -- ctypes binding of ``liboracle_sieve.so`` (sieve_oracle.c):
  the C restatement of the survivor set of src/sieve_plan.cpp + src/sieve.cu.
* ``survivors_legendre`` -- an independent pure-Python definition of the same
  set, by Euler's criterion on the homogenised value F(a, b) mod p.  It shares
  no code or tables with the C oracle and pins it on tiny boxes.
* ``exact_points`` / ``render_output`` -- the exact stage and the CLI output:
  restatement of the brute-force oracle in tests/test_multi_gpu.py:17-40, the
  verification of src/exact_polynomial.cpp:25-42 + src/point_search.cpp:48-70,
  the interval filter of src/point_search.cpp:131-146 and the formatter of
  src/output_formatter.cpp.
* ``site_digest`` -- the order-independent survivor-set digest, identical to
  rpo_digest in sieve_oracle.c.
"""

import ctypes
import math
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_sieve.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_plan.so")

CANDIDATES = 36
SELECTED = 32


# ---------------------------------------------------------------------------
# Independent definitions (pure Python)
# ---------------------------------------------------------------------------

def primes_below(limit):
    """Trial-division primes below ``limit`` (no shared code with the C)."""
    return [n for n in range(2, limit)
            if all(n % d for d in range(2, math.isqrt(n) + 1))]


def candidate_primes():
    """The 36 largest primes below 512, ascending (src/sieve_plan.cpp:60-61)."""
    return primes_below(512)[-CANDIDATES:]


def selected_primes(b):
    """The 32 largest candidates not dividing b, largest first
    (src/sieve_plan.cpp:75-86)."""
    out = [p for p in reversed(candidate_primes()) if b % p][:SELECTED]
    assert len(out) == SELECTED
    return out


def homogenised_value(coeffs, a, b):
    """F(a, b) = sum c_i a^i b^(n-i), times b for odd n
    (src/exact_polynomial.cpp:25-36)."""
    n = len(coeffs) - 1
    value = sum(c * a ** i * b ** (n - i) for i, c in enumerate(coeffs))
    return value * b if n % 2 else value


def survivors_legendre(coeffs, bound, first, last, square=False):
    """Survivor set by Euler's criterion: (a, b) survives iff for every
    selected p, F(a, b) mod p is 0 or a quadratic residue.  Equivalent to the
    reference's chi_p(a * b^-1) test because F(a, b) = b^(2m) f(a/b) and b^(2m)
    is a nonzero square mod p when p does not divide b."""
    coeffs = [int(c) for c in coeffs]
    out = []
    for t in range(first, last + 1):
        b = t * t if square else t
        primes = selected_primes(b)
        for a in range(-bound, bound + 1):
            v = homogenised_value(coeffs, a, b)
            if all(pow(v % p, (p - 1) // 2, p) in (0, 1) for p in primes):
                out.append((a, b))
    return out


def exact_points(coeffs, bound, first, last, square=False):
    """Exact stage over the box: primitive (a, b) with F(a, b) a perfect
    square, sorted by (b, a), as (a, y, b) with y >= 0
    (tests/test_multi_gpu.py:17-40; src/point_search.cpp:48-70)."""
    coeffs = [int(c) for c in coeffs]
    points = []
    for t in range(first, last + 1):
        b = t * t if square else t
        for a in range(-bound, bound + 1):
            if math.gcd(a, b) != 1:
                continue
            v = homogenised_value(coeffs, a, b)
            if v < 0:
                continue
            r = math.isqrt(v)
            if r * r == v:
                points.append((a, r, b))
    return points


def infinity_point(coeffs):
    """(1 : y : 0) when F(1, 0) is a square (src/point_search.cpp:185-189):
    F(1, 0) is c_n for even degree and 0 for odd degree."""
    coeffs = [int(c) for c in coeffs]
    n = len(coeffs) - 1
    v = 0 if n % 2 else coeffs[-1]
    if v < 0:
        return None
    r = math.isqrt(v)
    return (1, r, 0) if r * r == v else None


def decode_escapes(text):
    """Escape decoding of src/output_formatter.cpp:76-98 (\\n, \\t, \\%)."""
    out = []
    i = 0
    while i < len(text):
        ch = text[i]
        if ch != "\\" or i + 1 == len(text):
            out.append(ch)
            i += 1
            continue
        nxt = text[i + 1]
        i += 2
        if nxt == "n":
            out.append("\n")
        elif nxt == "t":
            out.append("\t")
        elif nxt == "%":
            out.append("\x01")
        else:
            out.append("\\" + nxt)
    return "".join(out)


def _format_point(fmt, a, y, b, print_y):
    out = []
    i = 0
    while i < len(fmt):
        ch = fmt[i]
        if ch != "%" or i + 1 == len(fmt):
            out.append(ch)
            i += 1
            continue
        marker = fmt[i + 1]
        i += 2
        if marker == "x":
            out.append(str(a))
        elif marker == "y" and print_y:
            out.append(y)
        elif marker == "z":
            out.append(str(b))
        else:
            out.append("%" + marker)
    return "".join(out).replace("\x01", "%")


def render_output(coeffs, bound, first, last, *, square=False,
                  infinity=True, print_y=True, print_points=True,
                  first_only=False, intervals=(), fmt=None, before="",
                  between="", after=""):
    """Byte-exact stdout of the CLI (src/main.cpp:68-72,
    src/output_formatter.cpp, src/point_search.cpp:131-146, 171-240).

    ``intervals`` is a list of (lower, upper) Fractions, None for an open end.
    """
    if fmt is None:
        fmt = "(%x : %y : %z)\\n" if print_y else "(%x : %z)\\n"
    fmt = decode_escapes(fmt)
    before, between, after = (decode_escapes(s) for s in (before, between, after))
    pieces = []
    emitted = [False]

    def write_point(a, y, b):
        if not print_points:
            return
        if emitted[0]:
            pieces.append(between)
        pieces.append(_format_point(fmt, a, y, b, print_y))
        emitted[0] = True

    def accepts(a, b):
        if not intervals:
            return True
        x = Fraction(a, b)
        return any((lo is None or x >= lo) and (hi is None or x <= hi)
                   for lo, hi in intervals)

    if print_points:
        pieces.append(before)
    points = []
    if infinity:
        inf = infinity_point(coeffs)
        if inf is not None:
            points.append((inf, False))
    points += [(p, True) for p in exact_points(coeffs, bound, first, last, square)]
    for (a, y, b), filtered in points:
        if filtered and not accepts(a, b):
            continue
        write_point(a, str(y), b)
        if first_only:
            break
        if print_y and y != 0:
            write_point(a, "-" + str(y), b)
    pieces.append(after)
    return "".join(pieces).encode()


# ---------------------------------------------------------------------------
# Digest (same definition as rpo_digest in sieve_oracle.c)
# ---------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def site_digest(num, den):
    """Order-independent digest (two wrapping uint64 sums) of a survivor set."""
    num = np.asarray(num, dtype=np.int64).view(np.uint64)
    den = np.asarray(den, dtype=np.int32).astype(np.uint32).astype(np.uint64)
    with np.errstate(over="ignore"):
        h = _splitmix64(num ^ _splitmix64(den))
        h1 = _splitmix64(h ^ np.uint64(0xA5A5A5A5A5A5A5A5))
        return (int(h.sum(dtype=np.uint64)), int(h1.sum(dtype=np.uint64)))


def add_digests(x, y):
    return ((x[0] + y[0]) & 0xFFFFFFFFFFFFFFFF, (x[1] + y[1]) & 0xFFFFFFFFFFFFFFFF)


# ---------------------------------------------------------------------------
# ctypes binding of the C oracle
# ---------------------------------------------------------------------------

def _char_array(strings):
    arr = (ctypes.c_char_p * len(strings))()
    arr[:] = [s.encode() for s in strings]
    return arr


class OracleLib:
    """Binding of liboracle_sieve.so (built by ``make -C oracle``)."""

    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        lib = ctypes.CDLL(path)
        lib.rpo_curve_new.restype = ctypes.c_void_p
        lib.rpo_curve_new.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
        lib.rpo_curve_free.argtypes = [ctypes.c_void_p]
        for name in ("rpo_sieve_direct", "rpo_sieve_words"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_longlong
            fn.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_longlong, ctypes.POINTER(ctypes.c_uint64)]
        lib.rpo_decimal_mod.restype = ctypes.c_int
        lib.rpo_decimal_mod.argtypes = [ctypes.c_char_p, ctypes.c_int]
        lib.rpo_mod_inverse.restype = ctypes.c_longlong
        lib.rpo_mod_inverse.argtypes = [ctypes.c_longlong, ctypes.c_longlong]
        lib.rpo_affine_mask.restype = ctypes.c_int
        lib.rpo_affine_mask.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.c_int, ctypes.c_void_p]
        lib.rpo_select_primes.restype = ctypes.c_int
        lib.rpo_select_primes.argtypes = [ctypes.c_longlong, ctypes.c_void_p]
        lib.rpo_candidate_primes.restype = ctypes.c_int
        lib.rpo_candidate_primes.argtypes = [ctypes.c_void_p]
        self.lib = lib

    def candidate_primes(self):
        out = np.zeros(CANDIDATES, dtype=np.int32)
        self.lib.rpo_candidate_primes(out.ctypes.data)
        return [int(x) for x in out]

    def selected_primes(self, b):
        idx = np.zeros(SELECTED, dtype=np.int32)
        if self.lib.rpo_select_primes(int(b), idx.ctypes.data) != SELECTED:
            raise ValueError("no 32 selectable primes")
        primes = self.candidate_primes()
        return [primes[i] for i in idx]

    def affine_mask(self, p, coeffs):
        out = np.zeros(p, dtype=np.uint8)
        arr = _char_array([str(c) for c in coeffs])
        self.lib.rpo_affine_mask(p, arr, len(coeffs), out.ctypes.data)
        return out

    def sieve(self, coeffs, bound, first, last, square=False, threads=None,
              method="direct", collect=True, cap=None):
        """Returns (count, digest, num, den); num/den sorted by (b, a) when
        collected."""
        # One thread per usable core, at most 16 (a GPU box's CPU share).
        threads = threads or max(1, min(16, len(os.sched_getaffinity(0))))
        arr = _char_array([str(c) for c in coeffs])
        curve = self.lib.rpo_curve_new(arr, len(coeffs))
        if not curve:
            raise RuntimeError("rpo_curve_new failed")
        fn = self.lib.rpo_sieve_direct if method == "direct" else self.lib.rpo_sieve_words
        digest = (ctypes.c_uint64 * 2)()
        try:
            if not collect:
                count = fn(curve, bound, first, last, int(square), threads, None, None, 0, digest)
                if count < 0:
                    raise RuntimeError("oracle sieve failed")
                return count, (digest[0], digest[1]), None, None
            if cap is None:
                count = fn(curve, bound, first, last, int(square), threads, None, None, 0, digest)
                if count < 0:
                    raise RuntimeError("oracle sieve failed")
                cap = count
            num = np.zeros(max(cap, 1), dtype=np.int64)
            den = np.zeros(max(cap, 1), dtype=np.int32)
            count = fn(curve, bound, first, last, int(square), threads,
                       num.ctypes.data, den.ctypes.data, cap, digest)
            if count < 0 or count > cap:
                raise RuntimeError("oracle sieve failed")
            return count, (digest[0], digest[1]), num[:count], den[:count]
        finally:
            self.lib.rpo_curve_free(curve)


class RefPlanLib:
    """Binding of oracle/_ref/libref_plan.so: the reference's own
    src/sieve_plan.cpp.  Present only where the reference sources were."""

    def __init__(self, path=REF_LIB_PATH):
        if not os.path.exists(path):
            raise RuntimeError(f"reference plan library not built: {path}")
        lib = ctypes.CDLL(path)
        lib.ref_plan_selected_primes.restype = ctypes.c_int
        lib.ref_plan_selected_primes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_void_p]
        lib.ref_plan_sizes.restype = ctypes.c_int
        lib.ref_plan_sizes.argtypes = [ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong),
                                       ctypes.POINTER(ctypes.c_longlong)]
        lib.ref_mod_inverse.restype = ctypes.c_longlong
        lib.ref_mod_inverse.argtypes = [ctypes.c_longlong, ctypes.c_longlong]
        lib.ref_affine_mask.restype = ctypes.c_int
        lib.ref_affine_mask.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.c_int, ctypes.c_void_p]
        self.lib = lib

    def selected_primes(self, first, last, square=False):
        out = np.zeros((last - first + 1) * SELECTED, dtype=np.int32)
        if self.lib.ref_plan_selected_primes(first, last, int(square), out.ctypes.data) != 0:
            raise ValueError("reference plan rejected the range")
        return out.reshape(-1, SELECTED)

    def affine_mask(self, p, coeffs):
        out = np.zeros(p, dtype=np.uint8)
        arr = _char_array([str(c) for c in coeffs])
        self.lib.ref_affine_mask(p, arr, len(coeffs), out.ctypes.data)
        return out
