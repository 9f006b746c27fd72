"""rps_sieve_into, the allocation-free entry point, against rps_sieve.

Bar: the survivor multiset written to the caller's device buffers is
bit-identical to what rps_sieve streams for the same request, on two calls
in a row (the second with a smaller input, so stale scratch contents from the
first call would show), and an output buffer too small for the survivors is
reported as RPS_EOVERFLOW with a subset of the survivors written.  Device
buffers come from the HIP runtime the library itself links
(libamdhip64.so.7), through ctypes.
"""

import ctypes

import numpy as np
import pytest

import curves

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    """int64 numerators and int32 denominators of `capacity` entries in device
    memory (hipMalloc of the library's own HIP runtime)."""

    def __init__(self, capacity):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.capacity = capacity
        self.num = ctypes.c_void_p()
        self.den = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.num), capacity * 8) == 0
        assert self.hip.hipMalloc(ctypes.byref(self.den), capacity * 4) == 0

    def read(self, n):
        num = np.zeros(n, dtype=np.int64)
        den = np.zeros(n, dtype=np.int32)
        if n:
            assert self.hip.hipMemcpy(num.ctypes.data, self.num, n * 8, 2) == 0  # device to host
            assert self.hip.hipMemcpy(den.ctypes.data, self.den, n * 4, 2) == 0
        order = np.lexsort((num, den))
        return num[order], den[order]

    def free(self):
        self.hip.hipFree(self.num)
        self.hip.hipFree(self.den)


def check_against_sieve(ctx, bufs, coeffs, bound, first, last, square=False):
    print(f"case: sieve_into A={bound} b={first}..{last} square={square}", flush=True)
    ref = ctx.sieve(coeffs, bound, first, last, square=square)
    rc, n = ctx.sieve_into(coeffs, bound, first, last, bufs.num.value, bufs.den.value,
                           bufs.capacity, square=square)
    assert rc == 0
    assert n == ref.survivor_count
    num, den = bufs.read(n)
    assert np.array_equal(num, ref.numerators)
    assert np.array_equal(den, ref.denominators)
    return n


def test_sieve_into_matches_sieve_twice(fresh_gpu, oracle_lib):
    bufs = DeviceBuffers(1 << 20)
    try:
        coeffs = ["1", "-3", "0", "2", "0", "0", "4"]
        # Long lines on row tiles, then a smaller request: short lines on
        # the short-line kernel over fewer denominators.
        n1 = check_against_sieve(fresh_gpu, bufs, coeffs, 2_000_000, 1, 700)
        n2 = check_against_sieve(fresh_gpu, bufs, coeffs, 20000, 1, 300)
        assert n2 < n1
        count, digest, num, den = oracle_lib.sieve(coeffs, 20000, 1, 300, method="words")
        assert n2 == count
    finally:
        bufs.free()


def test_sieve_into_modes_and_overflow(fresh_gpu):
    bufs = DeviceBuffers(1 << 18)
    try:
        coeffs, _ = curves.partially_dense(__import__("random").Random(7), 7, 6)
        check_against_sieve(fresh_gpu, bufs, coeffs, 20000, 1, 300, square=True)
        fresh_gpu.set_option("primitive_only", 1)
        check_against_sieve(fresh_gpu, bufs, coeffs, 30000, 5, 400)
        fresh_gpu.set_option("primitive_only", 0)
        # Dense curve: far more survivors than 4,096 entries.
        dense = ["1", str(curves.dense_product())]
        print("case: sieve_into overflow", flush=True)
        rc, n = fresh_gpu.sieve_into(dense, 5000, 1, 2, bufs.num.value, bufs.den.value, 4096)
        assert rc == 5 and 0 < n <= 4096
        num, den = bufs.read(n)
        full = fresh_gpu.sieve(dense, 5000, 1, 2)
        written = set(zip(num.tolist(), den.tolist()))
        assert len(written) == n
        assert written <= set(zip(full.numerators.tolist(), full.denominators.tolist()))
        # The next call starts from a clean state.
        check_against_sieve(fresh_gpu, bufs, coeffs, 3000, 1, 50)
    finally:
        bufs.free()
