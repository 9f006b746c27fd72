"""Python binding of the gfx950 modular square sieve's C ABI.

The product is the C ABI in ``include/ratpoints_sieve.h`` (implemented by
``lib/libratpoints_sieve.so``) and the C++ drop-in host built on it
(``bin/ratpoints_gpu``).  This module is plumbing for tests and benchmarks:
a ctypes mirror of ``run_modular_sieve`` (include/sieve.hpp:43-47 in the
reference) that streams survivors into numpy arrays.

It never falls back to a CPU path: when the shared library is missing or no
HIP device can be opened, it raises.
"""

import ctypes
import os
import time

import numpy as np

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PACKAGE_DIR, "lib", "libratpoints_sieve.so")
CLI_PATH = os.path.join(PACKAGE_DIR, "bin", "ratpoints_gpu")
HEADER_PATH = os.path.join(os.path.dirname(PACKAGE_DIR), "include", "ratpoints_sieve.h")

RPS_OK = 0
RPS_EINVAL = 1
RPS_EDEVICE = 2
RPS_ENOMEM = 3
RPS_ECANCELLED = 4
RPS_EOVERFLOW = 5
RPS_EINTERNAL = 6


class RpsRequest(ctypes.Structure):
    _fields_ = [
        ("coefficients", ctypes.POINTER(ctypes.c_char_p)),
        ("n_coefficients", ctypes.c_int32),
        ("numerator_bound", ctypes.c_int64),
        ("den_first", ctypes.c_int32),
        ("den_last", ctypes.c_int32),
        ("square_denominators", ctypes.c_int32),
    ]


class RpsStats(ctypes.Structure):
    _fields_ = [
        ("survivor_count", ctypes.c_uint64),
        ("word_count", ctypes.c_uint64),
        ("basis_ms", ctypes.c_double),
        ("sieve_ms", ctypes.c_double),
        ("plan_ms", ctypes.c_double),
        ("workspace_ms", ctypes.c_double),
        ("basis_wall_ms", ctypes.c_double),
        ("survivor_setup_ms", ctypes.c_double),
        ("verification_ms", ctypes.c_double),
        ("sieve_wall_ms", ctypes.c_double),
        ("initial_mask_bytes", ctypes.c_double),
        ("launches", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("requeued", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                        ctypes.POINTER(ctypes.c_int32), ctypes.c_size_t)
MULTI_SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                              ctypes.c_size_t)

_LIB = None


class SieveError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def load_library(path=LIB_PATH):
    """Loads the HIP sieve library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP sieve library not built: {path}; run __graft_entry__.build()")
    _LIB = bind_library(path)
    return _LIB


def bind_library(path):
    """Declares the C ABI's prototypes on the shared library at `path`
    (another build of the same ABI, for tests) without making it the
    module's default library."""
    lib = ctypes.CDLL(path)
    lib.rps_open.restype = ctypes.c_int
    lib.rps_open.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.rps_close.restype = None
    lib.rps_close.argtypes = [ctypes.c_void_p]
    lib.rps_last_error.restype = ctypes.c_char_p
    lib.rps_last_error.argtypes = [ctypes.c_void_p]
    lib.rps_sieve.restype = ctypes.c_int
    lib.rps_sieve.argtypes = [ctypes.c_void_p, ctypes.POINTER(RpsRequest), SINK,
                              ctypes.c_void_p, ctypes.POINTER(RpsStats)]
    lib.rps_sieve_many.restype = ctypes.c_int
    lib.rps_sieve_many.argtypes = [ctypes.c_void_p, ctypes.POINTER(RpsRequest), ctypes.c_size_t,
                                   MULTI_SINK, ctypes.c_void_p, ctypes.POINTER(RpsStats)]
    lib.rps_sieve_into.restype = ctypes.c_int
    lib.rps_sieve_into.argtypes = [ctypes.c_void_p, ctypes.POINTER(RpsRequest), ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_uint64,
                                   ctypes.POINTER(ctypes.c_uint64)]
    lib.rps_set_option.restype = ctypes.c_int
    lib.rps_set_option.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]
    lib.rps_variant_count.restype = ctypes.c_int
    lib.rps_variant_count.argtypes = []
    lib.rps_variant_name.restype = ctypes.c_char_p
    lib.rps_variant_name.argtypes = [ctypes.c_int]
    lib.rps_tuned_variant.restype = ctypes.c_int
    lib.rps_tuned_variant.argtypes = [ctypes.c_uint64]
    lib.rps_tuned.restype = ctypes.c_int
    lib.rps_tuned.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.rps_variant_info.restype = ctypes.c_int
    lib.rps_variant_info.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 5
    lib.rps_device_count.restype = ctypes.c_int
    lib.rps_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.rps_check_request.restype = ctypes.c_int
    lib.rps_check_request.argtypes = [ctypes.POINTER(RpsRequest), ctypes.c_char_p,
                                      ctypes.c_size_t]
    lib.rps_last_kernel_timing.restype = ctypes.c_int
    lib.rps_last_kernel_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_double)]
    return lib


def check_request(coefficients, bound, first, last, square=False):
    """rps_check_request: (status code, reason) without touching a device."""
    lib = load_library()
    coef = (ctypes.c_char_p * len(coefficients))(*[str(c).encode() for c in coefficients])
    req = RpsRequest(coef, len(coefficients), bound, first, last, 1 if square else 0)
    message = ctypes.create_string_buffer(256)
    rc = lib.rps_check_request(ctypes.byref(req), message, len(message))
    return rc, message.value.decode()


def variants():
    lib = load_library()
    return [lib.rps_variant_name(i).decode() for i in range(lib.rps_variant_count())]


def tuned_variant(words):
    """Index of the variant the tuning table picks for lines of `words`
    64-bit words."""
    return load_library().rps_tuned_variant(int(words))


def tuned(words, many=False):
    """The tuning table's (variant index, run_log2, batch) for lines of
    `words` words on the plain path or in many-curve launches."""
    v, r, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    load_library().rps_tuned(int(words), 1 if many else 0, ctypes.byref(v), ctypes.byref(r),
                             ctypes.byref(b))
    return v.value, r.value, b.value


def variant_info(index, lib=None):
    """Shape of a compiled variant: dict of words_per_lane, dense_primes,
    waves, run_log2, wide."""
    lib = lib or load_library()
    vals = [ctypes.c_int() for _ in range(5)]
    if lib.rps_variant_info(int(index), *[ctypes.byref(v) for v in vals]) != 0:
        raise ValueError("no variant %d" % index)
    return dict(zip(("words_per_lane", "dense_primes", "waves", "run_log2", "wide"),
                    (v.value for v in vals)))


class SieveResult:
    def __init__(self, numerators, denominators, stats, status):
        self.numerators = numerators
        self.denominators = denominators
        self.stats = stats
        self.status = status

    @property
    def survivor_count(self):
        return self.stats["survivor_count"]


class SieveContext:
    """One device context (rps_open/rps_close)."""

    def __init__(self, device=0, lib=None, **options):
        self.lib = lib if lib is not None else load_library()
        handle = ctypes.c_void_p()
        rc = self.lib.rps_open(device, ctypes.byref(handle))
        if rc != RPS_OK:
            raise SieveError(rc, self.lib.rps_last_error(None).decode())
        self.handle = handle
        for name, value in options.items():
            self.set_option(name, value)

    def close(self):
        if self.handle:
            self.lib.rps_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return self.lib.rps_last_error(self.handle).decode()

    def set_option(self, name, value):
        rc = self.lib.rps_set_option(self.handle, name.encode(), int(value))
        if rc != RPS_OK:
            raise SieveError(rc, self.error())

    def kernel_timing(self):
        ms = ctypes.c_double()
        launches = ctypes.c_uint64()
        sites = ctypes.c_double()
        self.lib.rps_last_kernel_timing(self.handle, ctypes.byref(ms), ctypes.byref(launches),
                                        ctypes.byref(sites))
        return ms.value, launches.value, sites.value

    def run(self, coefficients, bound, first, last, square=False, sink=None):
        """Runs rps_sieve.  ``sink(num, den)`` receives numpy views valid only
        during the call; return True to cancel.  Returns (status, stats)."""
        coeffs = [str(c) for c in coefficients]
        arr = (ctypes.c_char_p * len(coeffs))()
        arr[:] = [c.encode() for c in coeffs]
        req = RpsRequest(arr, len(coeffs), int(bound), int(first), int(last), int(bool(square)))
        stats = RpsStats()
        error = []

        def trampoline(_user, num_ptr, den_ptr, n):
            try:
                num = np.ctypeslib.as_array(num_ptr, shape=(n,))
                den = np.ctypeslib.as_array(den_ptr, shape=(n,))
                return 1 if (sink(num, den) if sink else False) else 0
            except BaseException as exc:  # surface sink failures to the caller
                error.append(exc)
                return 1

        cb = SINK(trampoline)
        rc = self.lib.rps_sieve(self.handle, ctypes.byref(req), cb, None, ctypes.byref(stats))
        if error:
            raise error[0]
        if rc not in (RPS_OK, RPS_ECANCELLED):
            raise SieveError(rc, self.error())
        return rc, stats.as_dict()

    def sieve_into(self, coefficients, bound, first, last, numerators, denominators,
                   capacity, square=False):
        """rps_sieve_into: survivors written to the device buffers at the
        addresses `numerators` (int64) and `denominators` (int32) of
        `capacity` entries each.  Returns (status, survivor count); raises on
        errors other than RPS_EOVERFLOW."""
        coeffs = [str(c) for c in coefficients]
        arr = (ctypes.c_char_p * len(coeffs))()
        arr[:] = [c.encode() for c in coeffs]
        req = RpsRequest(arr, len(coeffs), int(bound), int(first), int(last), int(bool(square)))
        count = ctypes.c_uint64()
        rc = self.lib.rps_sieve_into(self.handle, ctypes.byref(req), ctypes.c_void_p(numerators),
                                     ctypes.c_void_p(denominators), int(capacity),
                                     ctypes.byref(count))
        if rc not in (RPS_OK, RPS_EOVERFLOW):
            raise SieveError(rc, self.error())
        return rc, count.value

    def run_many(self, requests, sink=None):
        """Runs rps_sieve_many.  ``requests`` is a list of dicts with keys
        coefficients, bound, first, last and optional square.  ``sink(req,
        num, den)`` gets numpy views valid only during the call; return True
        to cancel.  Returns (status, stats)."""
        keep = []
        reqs = (RpsRequest * len(requests))()
        for i, r in enumerate(requests):
            coeffs = [str(c) for c in r["coefficients"]]
            arr = (ctypes.c_char_p * len(coeffs))()
            arr[:] = [c.encode() for c in coeffs]
            keep.append(arr)
            reqs[i] = RpsRequest(arr, len(coeffs), int(r["bound"]), int(r["first"]),
                                 int(r["last"]), int(bool(r.get("square", False))))
        stats = RpsStats()
        error = []

        def trampoline(_user, req_ptr, num_ptr, den_ptr, n):
            try:
                req = np.ctypeslib.as_array(req_ptr, shape=(n,))
                num = np.ctypeslib.as_array(num_ptr, shape=(n,))
                den = np.ctypeslib.as_array(den_ptr, shape=(n,))
                return 1 if (sink(req, num, den) if sink else False) else 0
            except BaseException as exc:
                error.append(exc)
                return 1

        cb = MULTI_SINK(trampoline)
        rc = self.lib.rps_sieve_many(self.handle, reqs, len(requests), cb, None,
                                     ctypes.byref(stats))
        if error:
            raise error[0]
        if rc not in (RPS_OK, RPS_ECANCELLED):
            raise SieveError(rc, self.error())
        return rc, stats.as_dict()

    def sieve_many(self, requests):
        """Collects every survivor of every request: returns (results, stats)
        with results[i] = (num, den) sorted by (b, a)."""
        parts = {}

        def sink(req, num, den):
            parts.setdefault("req", []).append(req.copy())
            parts.setdefault("num", []).append(num.copy())
            parts.setdefault("den", []).append(den.copy())
            return False

        rc, stats = self.run_many(requests, sink)
        if parts:
            req = np.concatenate(parts["req"])
            num = np.concatenate(parts["num"])
            den = np.concatenate(parts["den"])
        else:
            req = np.zeros(0, dtype=np.int32)
            num = np.zeros(0, dtype=np.int64)
            den = np.zeros(0, dtype=np.int32)
        order = np.lexsort((num, den, req))
        req, num, den = req[order], num[order], den[order]
        bounds = np.searchsorted(req, np.arange(len(requests) + 1))
        results = [(num[bounds[i]:bounds[i + 1]], den[bounds[i]:bounds[i + 1]])
                   for i in range(len(requests))]
        return results, stats

    def sieve(self, coefficients, bound, first, last, square=False):
        """Collects every survivor; returns SieveResult with arrays sorted by
        (b, a)."""
        nums, dens = [], []

        def sink(num, den):
            nums.append(num.copy())
            dens.append(den.copy())
            return False

        rc, stats = self.run(coefficients, bound, first, last, square, sink)
        num = np.concatenate(nums) if nums else np.zeros(0, dtype=np.int64)
        den = np.concatenate(dens) if dens else np.zeros(0, dtype=np.int32)
        order = np.lexsort((num, den))
        return SieveResult(num[order], den[order], stats, rc)


def run_modular_sieve(coefficients, numerator_bound, denominators, callback,
                      square_denominators=False, device=0):
    """Mirror of ratpoints_gpu::run_modular_sieve (include/sieve.hpp:43-47):
    ``denominators`` is (first, last) inclusive; ``callback(num, den)`` gets
    each flushed batch.  Returns the stats dict (survivor_count + metrics)."""
    with SieveContext(device) as ctx:
        first, last = denominators
        _, stats = ctx.run(coefficients, numerator_bound, first, last, square_denominators,
                           lambda n, d: (callback(n, d), False)[1])
        return stats
