// gfx950 kernels of the modular square sieve.
//
//   chi_kernel     chi_p bitmaps of a batch of curves (src/sieve_plan.cpp:
//                  115-138 is the host routine it replaces)
//   resid_kernel   per (p, b mod p) residue strings; replaces the reference's
//                  p x p 32-bit basis table (basis_kernel, src/sieve.cu:141-172)
//                  with a 72-byte string per (p, b mod p)
//   sieve_kernel   the hot loop (src/sieve.cu:221-388 is the kernel it
//                  replaces): one workgroup per row tile (a chunk of one line
//                  or a run of whole lines); dense primes ANDed from LDS rows
//                  with wave-uniform offsets (8-byte row entries, or 16-byte
//                  elements read with one ds_read_b128 in the wide tiles);
//                  remaining primes applied to surviving words only, through
//                  a per-wave queue; survivors compacted with a wave prefix
//                  sum and one atomic reservation per drain, with a resumable
//                  per-wave cursor
//   direct_kernel  short lines: tiles of whole lines, windows evaluated
//                  directly from the residue strings
//   multi_kernel   many curves in one persistent launch: workgroups pull row
//                  and short-line tiles of all curves from a device counter
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "sieve_kernels.hpp"
#include "sieve_layout.hpp"

namespace rps {
namespace {

// x mod p for x < 2^32 with magic = floor(2^32 / p): the quotient estimate is
// at most one too small, so one conditional subtraction finishes.
__device__ __forceinline__ uint32_t mod_small(uint32_t x, uint32_t p,
                                              uint32_t magic) {
    uint32_t q = __umulhi(x, magic);
    uint32_t r = x - q * p;
    return r >= p ? r - p : r;
}

// 64 bits of the residue string `c` starting at bit o (o + 63 < 576).
__device__ __forceinline__ uint64_t window64(const uint64_t *c, uint32_t o) {
    uint32_t q = o >> 6;
    uint32_t sh = o & 63;
    uint64_t lo = c[q];
    uint64_t hi = c[q + 1];
    return sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
}

// The same window from the string read as 32-bit words: two funnel shifts
// of the three words from o / 32 (o + 63 < 576).
__device__ __forceinline__ uint64_t window64_32(const uint32_t *c, uint32_t o) {
    const uint32_t *d = c + (o >> 5);
    const uint32_t lo = __builtin_amdgcn_alignbit(d[1], d[0], o & 31);
    const uint32_t hi = __builtin_amdgcn_alignbit(d[2], d[1], o & 31);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), lane);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// The wave's issue priority (0..3; the SIMD arbiter prefers higher
// priority, then older waves).
__device__ __forceinline__ void set_priority(int level) {
    switch (level) {
    case 3:
        __builtin_amdgcn_s_setprio(3);
        break;
    case 2:
        __builtin_amdgcn_s_setprio(2);
        break;
    case 1:
        __builtin_amdgcn_s_setprio(1);
        break;
    default:
        __builtin_amdgcn_s_setprio(0);
        break;
    }
}

// Orders one wave's LDS accesses before and after this point (lanes read
// what other lanes of the wave wrote).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int denominator(const LineDesc &L, long long t) {
    return static_cast<int>(L.square ? t * t : t);
}

// gcd(|a|, b) == 1 for b >= 1 (Euclid on a mod b, then 32-bit binary gcd).
__device__ __forceinline__ bool coprime(long long a, uint32_t b) {
    const uint64_t ua = a < 0 ? 0ULL - static_cast<uint64_t>(a) : static_cast<uint64_t>(a);
    uint32_t x = static_cast<uint32_t>(ua % b);
    uint32_t y = b;
    if (x == 0) {
        return y == 1;
    }
    const int shift = __builtin_ctz(x | y);
    x >>= __builtin_ctz(x);
    do {
        y >>= __builtin_ctz(y);
        if (x > y) {
            const uint32_t t = x;
            x = y;
            y = t;
        }
        y -= x;
    } while (y);
    return (x << shift) == 1;
}

// The primitive survivors of a word: bit j of m kept when gcd(a0 + j, b) = 1
// (src/point_search.cpp:48-60 applies the same test on the host; points
// need a primitive (a, b)).
__device__ __forceinline__ uint64_t primitive_bits(uint64_t m, uint64_t a0, int b) {
    uint64_t keep = 0;
    while (m) {
        const int j = __ffsll(static_cast<long long>(m)) - 1;
        m &= m - 1;
        if (coprime(static_cast<long long>(a0 + j), static_cast<uint32_t>(b))) {
            keep |= 1ULL << j;
        }
    }
    return keep;
}

// Wave-level survivor emission.  Lane l holds N masks; mask k covers the 64
// numerators of word word[k] on denominator den[k].  A prefix sum of per-lane
// counts gives every lane its slots under one atomic reservation per wave,
// all or nothing: when the reservation does not fit, first_fail records its
// start, nothing is written and false is returned, so the wave can stop and
// resume at the same point after the host drains the buffer.
// With O.primitive set, only survivors with gcd(a, b) = 1 are emitted.
template <int N>
__device__ __forceinline__ bool emit_words(const OutDesc &O, long long bound,
                                           const uint64_t (&m_in)[N],
                                           const uint64_t (&word)[N], const int (&den)[N],
                                           int request, int lane) {
    uint64_t m[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        m[k] = m_in[k];
    }
    if (O.primitive) {
        const uint64_t neg_a = static_cast<uint64_t>(-bound);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            m[k] = m[k] ? primitive_bits(m[k], neg_a + 64 * word[k], den[k]) : 0ULL;
        }
    }
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        cnt += __popcll(m[k]);
    }
    uint32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d, 64);
        if (lane >= d) {
            incl += v;
        }
    }
    const uint32_t total = __shfl(incl, 63, 64);
    if (total == 0) {
        return true;  // every survivor filtered out: nothing to reserve
    }
    unsigned long long start = 0;
    if (lane == 0) {
        start = atomicAdd(&O.state->count, static_cast<unsigned long long>(total));
    }
    start = __shfl(start, 0, 64);
    if (start + total > O.capacity) {
        if (lane == 0) {
            atomicMin(&O.state->first_fail, start);
            atomicOr(&O.state->full, 1u);
        }
        return false;
    }
    unsigned long long pos = start + incl - cnt;
    const uint64_t neg_a = static_cast<uint64_t>(-bound);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        uint64_t mm = m[k];
        const uint64_t a0 = neg_a + 64 * word[k];
        while (mm) {
            const int j = __ffsll(static_cast<long long>(mm)) - 1;
            mm &= mm - 1;
            O.num[pos] = static_cast<long long>(a0 + j);
            O.den[pos] = den[k];
            if (O.req) {
                O.req[pos] = request;
            }
            ++pos;
        }
    }
    return true;
}

// emit_words for U masks on the consecutive words first_word + 64 u.
template <int U>
__device__ __forceinline__ bool emit_survivors(const OutDesc &O, long long bound,
                                               const uint64_t (&m)[U],
                                               uint64_t first_word, int b,
                                               int request, int lane) {
    uint64_t word[U];
    int den[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        word[u] = first_word + 64 * u;
        den[u] = b;
    }
    return emit_words<U>(O, bound, m, word, den, request, lane);
}

}  // namespace

// Where a wave stops when survivors do not fit: the line of its run and the
// position within it (a word, or W + u for the queued tiles).
struct Cursor {
    uint32_t line;
    uint64_t pos;
};

// ---------------------------------------------------------------------------
// Row tile: one denominator, one chunk [ts, te) of its line.
//   U      64-bit words per lane per iteration (a wave covers 64 U words)
//   FAST   selected primes ANDed densely for every word; the other
//          32 - FAST are applied only to words with bits left
//   WAVES  waves per workgroup; each wave sweeps a contiguous word range
//   WIDE   (U = 2 only) rows of 16-byte elements: element e holds entries e
//          and e + 64, the two words a lane takes per iteration, so the
//          dense stage is one aligned ds_read_b128 per row.  gfx950 serves
//          16-byte LDS reads at about 1.6x the bytes per clock of paired
//          8-byte reads; the rows take twice the LDS, so such a tile runs
//          one workgroup per CU.
// ---------------------------------------------------------------------------
using u64x2 = unsigned long long __attribute__((ext_vector_type(2)));

// The s-th largest prime below 512 (s = 0: 509): the largest prime a dense
// row s can hold, since the selection is the candidates in descending order.
__host__ __device__ constexpr int largest_prime_below_512(int s) {
    constexpr int kDesc[kCandidates] = {
        509, 503, 499, 491, 487, 479, 467, 463, 461, 457, 449, 443,
        439, 433, 431, 421, 419, 409, 401, 397, 389, 383, 379, 373,
        367, 359, 353, 349, 347, 337, 331, 317, 313, 311, 307, 293};
    return kDesc[s];
}
constexpr bool is_prime_below_512(int n) {
    for (int d = 2; d * d <= n; ++d) {
        if (n % d == 0) {
            return false;
        }
    }
    return n > 1;
}
constexpr bool descending_primes_ok() {
    int s = 0;
    for (int n = 511; n > 1 && s < kCandidates; --n) {
        if (is_prime_below_512(n) && largest_prime_below_512(s++) != n) {
            return false;
        }
    }
    return s == kCandidates;
}
static_assert(descending_primes_ok(), "the 36 largest primes below 512, descending");

template <int U, int FAST, int WAVES, bool WIDE = false>
struct RowTile {
    using Base = RowTile;
    static constexpr int kWaves = WAVES;
    static constexpr int kThreads = WAVES * 64;
    static constexpr int kStride = 64 * U;
    // Row elements per slot, sized per row by the largest prime P_s it can
    // hold: P_s + 64 U entries (>= p + 64 U - 1, the most a sweep reads), or
    // P_s + 64 wide elements.  kSlot bounds them all.
    static constexpr int kSlot = WIDE ? kMaxPrime + 64 : kMaxPrime + 64 * U;
    __host__ __device__ static constexpr int slot_elems(int s) {
        return largest_prime_below_512(s) + (WIDE ? 64 : 64 * U);
    }
    // First 64-bit word of row s.
    __host__ __device__ static constexpr int row_base(int s) {
        int w = 0;
        for (int j = 0; j < s; ++j) {
            w += (WIDE ? 2 : 1) * slot_elems(j);
        }
        return w;
    }
    // Entries computed per row (a wide element also holds entry e + 64),
    // rounded up to whole waves.
    static constexpr int kEntries = WIDE ? (slot_elems(0) + 64 + 63) / 64 * 64 : kSlot;
    static constexpr int kFast = FAST;
    static constexpr bool kWide = WIDE;
    // LDS layout, in 64-bit words: rows, the dense primes' residue strings,
    // then two sets (by line parity) of the per-line arrays: six arrays of
    // 32 uint32 (p, (-A) mod p, b mod p, candidate index, ts mod p, magic)
    // and a flag, 1 when the line's dense primes are the FAST largest
    // candidates.
    static constexpr int kCext = row_base(FAST);
    static constexpr int kSmall = kCext + FAST * 10;
    static constexpr int kSmallWords = 3 * kSelected + 1;
    static constexpr int kDefaultFlag = 6 * kSelected;  // u32 index in a set
    static constexpr int kLdsWords = kSmall + 2 * kSmallWords;
    static_assert(kStride < 293, "a row step must stay below every prime");
    static_assert(FAST >= 1 && FAST <= kSelected, "dense prime count");
    static_assert(!WIDE || U == 2, "wide rows pair the two words of U = 2");

    __device__ __forceinline__ static uint32_t *small(uint64_t *pool, int par) {
        return reinterpret_cast<uint32_t *>(pool + kSmall + par * kSmallWords);
    }
    __device__ __forceinline__ static const uint32_t *small(const uint64_t *pool, int par) {
        return reinterpret_cast<const uint32_t *>(pool + kSmall + par * kSmallWords);
    }

    // Word range of wave `wave` in [ts, te), in whole strides.
    __device__ __forceinline__ static void wave_range(uint64_t ts, uint64_t te, int wave,
                                                      uint64_t &ws, uint64_t &we) {
        const uint64_t per =
            ((te - ts + static_cast<uint64_t>(kStride) * WAVES - 1)
             / (static_cast<uint64_t>(kStride) * WAVES)) * kStride;
        ws = ts + per * wave;
        we = ws + per < te ? ws + per : te;
        if (ws > we) {
            ws = we;
        }
    }

    // Dense rows from the strings and selection in LDS: entry e of row s is
    // word e of the line mod p, the window of the residue string at
    // (-A + 64 e) mod p; entries e >= p repeat the row so that reads never
    // wrap.
    __device__ __forceinline__ static void build_rows(uint64_t *pool, int tid, int par) {
        const uint32_t *s_p = small(pool, par);
        const uint32_t *s_c = s_p + kSelected;
        const uint32_t *s_magic = s_p + 5 * kSelected;
        const uint32_t *cext = reinterpret_cast<const uint32_t *>(pool + kCext);
        // Row constants as wave-uniform (scalar) values.
        uint32_t P[FAST], C[FAST], M[FAST];
#pragma unroll
        for (int s = 0; s < FAST; ++s) {
            P[s] = uniform(s_p[s]);
            C[s] = uniform(s_c[s]);
            M[s] = uniform(s_magic[s]);
        }
        // Thread t computes entry t of every row, and entry t + kThreads
        // while that is below kEntries (a wave-uniform test, as kEntries and
        // kThreads are multiples of 64).  Narrow rows: entry e goes into row
        // s when the row's slot has it (a per-row bound).  Wide rows: each
        // entry goes into the two elements that hold it, when row s has them
        // (a per-row bound).  All string reads of a thread are independent and
        // can be in flight together; the window at bit o is two funnel
        // shifts of three 32-bit string words.
        static_assert(kSlot % 64 == 0 && kEntries % 64 == 0 && kThreads % 64 == 0,
                      "wave-uniform slot tests");
        static_assert(kEntries <= 2 * kThreads, "two entries per thread at most");
        // Offsets first, then every string read, then the funnel shifts and
        // the row stores, so a thread's reads are all in flight together and
        // none waits behind a store to the same LDS array.
        const uint32_t e0 = static_cast<uint32_t>(tid);
        if (e0 < static_cast<uint32_t>(kEntries)) {
            build_entries(pool, cext, P, C, M, e0);
        }
        if (kThreads < kEntries && e0 < static_cast<uint32_t>(kEntries - kThreads)) {
            build_entries(pool, cext, P, C, M, e0 + kThreads);
        }
    }

    // Wide rows: the build as (row, 64-entry column) units dealt to the
    // waves in turn, so that every wave of the workgroup takes part (with
    // one thread per entry only the first kEntries threads had work, and
    // the line waited for them: 10 of 16 waves building 14 entries each);
    // a unit's row is wave-uniform.  Units are taken two at a time, their
    // string reads in flight together.
    static constexpr int kCols = kEntries / 64;
    static constexpr int kUnits = FAST * kCols;
    static constexpr int kUnitsPerWave = (kUnits + WAVES - 1) / WAVES;
    static constexpr int kUnitGroup = 2;

    __device__ __forceinline__ static void build_rows_dealt(uint64_t *pool, int lane, int wave,
                                                            int par) {
        const uint32_t *s_p = small(pool, par);
        const uint32_t *s_c = s_p + kSelected;
        const uint32_t *s_magic = s_p + 5 * kSelected;
        const char *cext = reinterpret_cast<const char *>(pool + kCext);
#pragma unroll
        for (int j0 = 0; j0 < kUnitsPerWave; j0 += kUnitGroup) {
            uint32_t w_lo[kUnitGroup], w_hi[kUnitGroup], e[kUnitGroup], elems[kUnitGroup];
            int base[kUnitGroup];
            bool live[kUnitGroup];
#pragma unroll
            for (int g = 0; g < kUnitGroup; ++g) {
                const int unit = wave + WAVES * (j0 + g);
                const int s = unit / kCols;
                const int col = unit - s * kCols;
                // Row s's base and element count, wave-uniform.
                int b = 0, n = 0;
#pragma unroll
                for (int k = 0; k < FAST; ++k) {
                    b += k < s ? 2 * slot_elems(k) : 0;
                    n = k == s ? slot_elems(k) : n;
                }
                base[g] = b;
                elems[g] = static_cast<uint32_t>(n);
                live[g] = j0 + g < kUnitsPerWave && unit < kUnits && 64 * col < n + 64;
                const int sr = live[g] ? s : 0;
                const uint32_t pp = uniform(s_p[sr]);
                const uint32_t cc = uniform(s_c[sr]);
                const uint32_t mg = uniform(s_magic[sr]);
                e[g] = 64u * static_cast<uint32_t>(col) + static_cast<uint32_t>(lane);
                const uint32_t o = live[g] ? mod_small(cc + 64u * e[g], pp, mg) : 0u;
                const uint32_t *d = reinterpret_cast<const uint32_t *>(cext + sr * 80) + (o >> 5);
                const uint32_t d0 = d[0];
                const uint32_t d1 = d[1];
                const uint32_t d2 = d[2];
                w_lo[g] = __builtin_amdgcn_alignbit(d1, d0, o & 31);
                w_hi[g] = __builtin_amdgcn_alignbit(d2, d1, o & 31);
            }
#pragma unroll
            for (int g = 0; g < kUnitGroup; ++g) {
                if (live[g]) {
                    const uint64_t w = (static_cast<uint64_t>(w_hi[g]) << 32) | w_lo[g];
                    if (e[g] < elems[g]) {
                        pool[base[g] + 2 * e[g]] = w;
                    }
                    if (e[g] >= 64u && e[g] < elems[g] + 64u) {
                        pool[base[g] + 2 * (e[g] - 64) + 1] = w;
                    }
                }
            }
        }
    }

    // Entry e of every dense row.
    __device__ __forceinline__ static void build_entries(uint64_t *pool, const uint32_t *cext,
                                                         const uint32_t (&P)[FAST],
                                                         const uint32_t (&C)[FAST],
                                                         const uint32_t (&M)[FAST], uint32_t e) {
        uint32_t o[FAST];
#pragma unroll
        for (int s = 0; s < FAST; ++s) {
            o[s] = mod_small(C[s] + 64u * e, P[s], M[s]);
        }
        uint32_t d0[FAST], d1[FAST], d2[FAST];
#pragma unroll
        for (int s = 0; s < FAST; ++s) {
            const uint32_t *d = cext + s * 20 + (o[s] >> 5);
            d0[s] = d[0];
            d1[s] = d[1];
            d2[s] = d[2];
        }
#pragma unroll
        for (int s = 0; s < FAST; ++s) {
            const uint32_t lo = __builtin_amdgcn_alignbit(d1[s], d0[s], o[s] & 31);
            const uint32_t hi = __builtin_amdgcn_alignbit(d2[s], d1[s], o[s] & 31);
            const uint64_t w = (static_cast<uint64_t>(hi) << 32) | lo;
            if (WIDE) {
                // Entry e is the low half of element e and the high half of
                // element e - 64, for the elements row s has.
                const uint32_t n = static_cast<uint32_t>(slot_elems(s));
                if (e < n) {
                    pool[row_base(s) + 2 * e] = w;
                }
                if (e >= 64u && e < n + 64u) {
                    pool[row_base(s) + 2 * (e - 64) + 1] = w;
                }
            } else if (e < static_cast<uint32_t>(slot_elems(s))) {
                pool[row_base(s) + e] = w;
            }
        }
    }

    // Lines whose dense primes are the FAST largest candidates (none of
    // them divides b: nearly every line) have dense rows whose entry e is
    // the window at o = (c_s + 64 e) mod p_s of the line's string, and that
    // offset depends on the curve only.  A Plan holds, per thread, the LDS
    // byte address of the string word o / 32 shifted left by 5 with o mod 32
    // in the low bits (the funnel shift reads those), for each entry the
    // thread builds: entry tid of every row, and its share of the tail
    // entries kThreads <= e < kEntries, which come in columns of 64 whose
    // FAST x kTailCols (row, column) units are dealt to the waves in turn.
    // Such a line's rows cost three LDS reads and two funnel shifts per
    // entry.  The plan stays in registers through the sweep, so only the
    // one-word tiles (U = 1), whose sweep leaves room, carry one; the
    // others build every line with the reductions.
    static constexpr bool kPlanned = !WIDE && U == 1;
    static constexpr int kTailCols = kEntries > kThreads ? (kEntries - kThreads) / 64 : 0;
    static constexpr int kTailUnits = FAST * kTailCols;
    static constexpr int kTailPerWave = (kTailUnits + WAVES - 1) / WAVES;
    // Tail unit of wave `wave` in its i-th turn.  The units start at wave
    // 1, so the waves with the most of them are not wave 0 and the last
    // wave, which also select the primes of the next two lines during the
    // build.
    __device__ __forceinline__ static int tail_unit(int wave, int i) {
        return (wave + WAVES - 1) % WAVES + WAVES * i;
    }
    struct Plan {
        uint32_t main[kPlanned ? FAST : 1];
        uint32_t tail[kPlanned && kTailPerWave > 0 ? kTailPerWave : 1];
    };

    // Packed geometry of entry e of row s (prime p, c = (-A) mod p).
    __device__ __forceinline__ static uint32_t plan_entry(int s, uint32_t p, uint32_t magic,
                                                          uint32_t c, uint32_t e) {
        const uint32_t o = mod_small(c + 64u * e, p, magic);
        const uint32_t byte = static_cast<uint32_t>(kCext * 8 + s * 80) + (o >> 5) * 4u;
        return (byte << 5) | (o & 31u);
    }

    // The plan of a curve from its candidate arrays (u32 p, magic, c of the
    // 36 candidates, ascending: rank s of a default line is candidate 35 - s).
    __device__ __forceinline__ static Plan make_plan(const uint32_t *cp, int tid, int lane,
                                                     int wave) {
        Plan pl{};
        if (kPlanned) {
#pragma unroll
            for (int s = 0; s < FAST; ++s) {
                const int k = kCandidates - 1 - s;
                pl.main[s] = plan_entry(s, cp[k], cp[kCandidates + k], cp[2 * kCandidates + k],
                                        static_cast<uint32_t>(tid));
            }
#pragma unroll
            for (int i = 0; i < kTailPerWave; ++i) {
                const int j = tail_unit(wave, i);
                if (j < kTailUnits) {
                    const int s = j % FAST;
                    const int k = kCandidates - 1 - s;
                    const uint32_t e = static_cast<uint32_t>(kThreads + 64 * (j / FAST) + lane);
                    pl.tail[i] = plan_entry(s, cp[k], cp[kCandidates + k],
                                            cp[2 * kCandidates + k], e);
                }
            }
        }
        return pl;
    }

    // The window a plan entry names.
    __device__ __forceinline__ static uint64_t plan_window(const uint64_t *pool, uint32_t g) {
        const uint32_t *d =
            reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(pool) + (g >> 5));
        const uint32_t d0 = d[0];
        const uint32_t d1 = d[1];
        const uint32_t d2 = d[2];
        const uint32_t lo = __builtin_amdgcn_alignbit(d1, d0, g & 31u);
        const uint32_t hi = __builtin_amdgcn_alignbit(d2, d1, g & 31u);
        return (static_cast<uint64_t>(hi) << 32) | lo;
    }

    // Dense rows of a line with the default selection, from the plan.
    static constexpr int kBuildGroup = 6;
    __device__ __forceinline__ static void build_planned(uint64_t *pool, const Plan &pl, int tid,
                                                         int lane, int wave) {
        if (!kPlanned) {
            return;
        }
        if (tid < kEntries) {
            // kBuildGroup rows' reads in flight at a time (all FAST at once
            // fits the same registers but measured no faster than six).
#pragma unroll
            for (int s0 = 0; s0 < FAST; s0 += kBuildGroup) {
                uint64_t w[kBuildGroup];
#pragma unroll
                for (int j = 0; j < kBuildGroup; ++j) {
                    if (s0 + j < FAST) {
                        w[j] = plan_window(pool, pl.main[s0 + j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < kBuildGroup; ++j) {
                    if (s0 + j < FAST && tid < slot_elems(s0 + j)) {
                        pool[row_base(s0 + j) + tid] = w[j];
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kTailPerWave; ++i) {
            const int j = tail_unit(wave, i);
            if (j < kTailUnits) {
                const int s = j % FAST;
                const int e = kThreads + 64 * (j / FAST) + lane;
                // Row s's base and size (s is wave-uniform).
                int base = 0, size = 0;
#pragma unroll
                for (int k = 0; k < FAST; ++k) {
                    base += k < s ? slot_elems(k) : 0;
                    size = k == s ? slot_elems(k) : size;
                }
                if (e < size) {
                    pool[base + e] = plan_window(pool, pl.tail[i]);
                }
            }
        }
    }

    // Row offsets of a wave, in bytes (entries of 8 bytes, wide elements of
    // 16): lane l holds dense row s = l mod 16, when s < FAST, with its
    // prime and the offset W mod p_s of the wave's next iteration, so each
    // 16-lane row of the wave holds every offset.  A DPP row broadcast
    // (row_newbcast:s) hands offset s to all lanes inside the address add,
    // one vector instruction per dense read; with v_readlane into an SGPR
    // and an add it took two, and with offsets kept in SGPRs three scalar
    // instructions per row and iteration.  Lanes with s >= FAST hold 0 and
    // a modulus of one stride, so their offset stays 0.
    static constexpr uint32_t kEntryBytes = WIDE ? 16u : 8u;
    static_assert(FAST <= 16, "one dense row per lane of a 16-lane DPP row");
    struct Offsets {
        uint32_t off;
        uint32_t p;
    };

    // Lanes of row s get prime s and the byte offset of word `rel` (from
    // the tile start) in its row.
    __device__ __forceinline__ static Offsets row_offsets(const uint32_t *s_p,
                                                          const uint32_t *s_base,
                                                          const uint32_t *s_magic, uint32_t rel,
                                                          int lane) {
        Offsets o{};
        o.off = 0u;
        o.p = static_cast<uint32_t>(kStride) * kEntryBytes;
        const int s = lane & 15;
        if (s < FAST) {
            const uint32_t p = s_p[s];
            const uint32_t v = mod_small(rel, p, s_magic[s]) + s_base[s];
            o.off = (v >= p ? v - p : v) * kEntryBytes;
            o.p = p * kEntryBytes;
        }
        return o;
    }

    // Byte address of this lane's element in dense row S: the row's offset,
    // broadcast from lane S of the lane's 16-lane row, plus the lane's own
    // element.  The DPP move folds into the add (v_add_u32_dpp).
    template <int S>
    __device__ __forceinline__ static uint32_t row_address(uint32_t off, uint32_t lane_bytes) {
        const uint32_t o = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(
            static_cast<int>(off), 0x150 + S /* row_newbcast:S */, 0xf, 0xf, true));
        return static_cast<uint32_t>(row_base(S)) * 8u + o + lane_bytes;
    }

    // Loads dense rows S.. of one iteration into x.
    template <int S>
    __device__ __forceinline__ static void load_rows(const uint64_t *pool, uint32_t off,
                                                     uint32_t lane_bytes, uint64_t (&x)[FAST][U]) {
        if constexpr (S < FAST) {
            const char *at = reinterpret_cast<const char *>(pool) + row_address<S>(off, lane_bytes);
            if (WIDE) {
                const u64x2 v = *reinterpret_cast<const u64x2 *>(at);
                x[S][0] = v.x;
                x[S][U - 1] = v.y;
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    x[S][u] = reinterpret_cast<const uint64_t *>(at)[64 * u];
                }
            }
            load_rows<S + 1>(pool, off, lane_bytes, x);
        }
    }

    // Dense stage of one iteration: m[u] = AND over the FAST rows of entry
    // off_s + lane + 64 u, then every offset advances one stride (mod p,
    // three vector instructions for all rows).  Rows are combined two at a
    // time on 32-bit halves with the three-input AND (v_bitop3_b32), one
    // instruction per half for two rows.
    __device__ __forceinline__ static void dense(const uint64_t *pool, Offsets &o, int lane,
                                                 uint64_t (&m)[U]) {
        // All loads first, so that each wave keeps FAST reads in flight.
        uint64_t x[FAST][U];
        load_rows<0>(pool, o.off, static_cast<uint32_t>(lane) * kEntryBytes, x);
        {
            const uint32_t next = o.off + static_cast<uint32_t>(kStride) * kEntryBytes;
            o.off = next >= o.p ? next - o.p : next;
        }
        uint32_t lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            lo[u] = ~0u;
            hi[u] = ~0u;
        }
#pragma unroll
        for (int s = 0; s < FAST; s += 2) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (s + 1 < FAST) {
                    lo[u] = __builtin_amdgcn_bitop3_b32(lo[u], static_cast<uint32_t>(x[s][u]),
                                                        static_cast<uint32_t>(x[s + 1][u]), 0x80);
                    hi[u] = __builtin_amdgcn_bitop3_b32(hi[u],
                                                        static_cast<uint32_t>(x[s][u] >> 32),
                                                        static_cast<uint32_t>(x[s + 1][u] >> 32),
                                                        0x80);
                } else {
                    lo[u] &= static_cast<uint32_t>(x[s][u]);
                    hi[u] &= static_cast<uint32_t>(x[s][u] >> 32);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            m[u] = (static_cast<uint64_t>(hi[u]) << 32) | lo[u];
        }
    }
};

// ---------------------------------------------------------------------------
// Row tile with a compaction queue for the late primes.  The dense stage is
// RowTile's; words that still have bits set are appended (ballot + mbcnt)
// to a per-wave queue of 128 entries in LDS, and whenever 64 or more are
// queued the wave drains them with every lane busy -- up to two entries per
// lane, each late prime evaluated from the residue string, a wave-uniform
// exit once all entries are empty -- and commits all of them in one
// reservation.
//
// The queue carries over from one line of a run to the next: the late
// primes' strings, constants and the denominator of each line go into a
// slot of a ring of kRing line slots in LDS, and a queue entry names its
// slot, so one drain can hold the words of several lines.  A wave drains
// early only before a line would reuse the slot of its oldest entry, and
// once at the end of the run.  On a line of a few thousand words each wave
// queues only a few words per line; draining them per line cost as much
// as the dense sweep.
//
// The resume cursor is (line of the oldest uncommitted entry, W + u of that
// entry): a drain commits the whole queue or nothing, so every entry after
// the oldest uncommitted one is uncommitted too, and a sweep that restarts
// there queues exactly those entries again.
// ---------------------------------------------------------------------------
template <int U, int FAST, int WAVES, bool WIDE = false>
struct RowTileQ {
    using Base = RowTile<U, FAST, WAVES, WIDE>;
    static constexpr int kWaves = WAVES;
    static constexpr int kStride = Base::kStride;
    static constexpr int kSlot = Base::kSlot;
    static constexpr int kLate = kSelected - FAST;
    static constexpr int kQueue = 128;
    static constexpr int kGroup = 3;  // late primes per drain step
    // Per wave: 128 masks (128 words) and 128 uint32 entry tags (64 words).
    static constexpr int kQueueWords = kQueue + kQueue / 2;
    static constexpr int kQueueBase = Base::kLdsWords;
    // Line slot, in 64-bit words: late strings [kLate][10], constants
    // [kLate] (p | cb << 16 | magic << 32, cb = the string offset of the
    // tile's first word), the denominator.
    static constexpr int kRingConst = kLate * 10;
    static constexpr int kRingDen = kRingConst + kLate;
    static constexpr int kRingSlotWords = kRingDen + 1;
    static constexpr int kRingBase = kQueueBase + WAVES * kQueueWords;
    // As many slots (up to 8) as keep the workgroups per CU that the tile
    // had without the ring: 160 KiB of LDS per CU, less the run's own arrays
    // and the kernels' scalars (kRunExtra words).
    static constexpr int kRunExtra = 96;
    static constexpr int kTwoPerCu = 10200;
    static constexpr int kOnePerCu = 20400;
    static constexpr int kBudget =
        kRingBase + kRunExtra + kRingSlotWords <= kTwoPerCu ? kTwoPerCu : kOnePerCu;
    static constexpr int kRingFit = (kBudget - kRingBase - kRunExtra) / kRingSlotWords;
    static constexpr int kRing = kRingFit < 1 ? 1 : (kRingFit > 8 ? 8 : kRingFit);
    static constexpr int kLdsWords = kRingBase + kRing * kRingSlotWords;
    // Entry tag: word offset from the tile start below bit 26 (tiles of
    // several lines are whole lines, and lines stay below 2^25 words), ring
    // slot from bit 28.
    static constexpr int kSlotShift = 28;
    static constexpr uint32_t kRelMask = (1u << 26) - 1;
    static_assert((kStride & (kStride - 1)) == 0, "the cursor encodes u below the stride");
    static_assert(kLate >= 1 && kLate <= 32, "late primes");

    __device__ __forceinline__ static void wave_range(uint64_t ts, uint64_t te, int wave,
                                                      uint64_t &ws, uint64_t &we) {
        Base::wave_range(ts, te, wave, ws, we);
    }

    __device__ __forceinline__ static uint64_t *ring(uint64_t *pool, int slot) {
        return pool + kRingBase + slot * kRingSlotWords;
    }
    __device__ __forceinline__ static const uint64_t *ring(const uint64_t *pool, uint32_t slot) {
        return pool + kRingBase + slot * kRingSlotWords;
    }

    // Where word 0 of string s (selection rank) of the line in `slot` goes.
    __device__ __forceinline__ static uint64_t *string(uint64_t *pool, int s, int slot) {
        return s < FAST ? pool + Base::kCext + s * 10 : ring(pool, slot) + (s - FAST) * 10;
    }

    // The line in `slot` selected prime p (its magic) at `rank`, with string
    // offset cb at the tile's first word; denominator b.
    __device__ __forceinline__ static void note_line(uint64_t *pool, int slot, int rank, uint32_t p,
                                                     uint32_t magic, uint32_t cb, uint32_t b) {
        uint64_t *r = ring(pool, slot);
        if (rank >= FAST) {
            r[kRingConst + rank - FAST] =
                (static_cast<uint64_t>(magic) << 32) | p | (cb << 16);
        }
        if (rank == 0) {
            r[kRingDen] = b;
        }
    }

    // A wave's queue between lines: entries waiting and the run line of the
    // oldest.
    struct Carry {
        uint32_t queued = 0;
        uint32_t first_line = 0;
    };

    __device__ __forceinline__ static uint64_t *queue_masks(uint64_t *pool, int wave) {
        return pool + kQueueBase + wave * kQueueWords;
    }
    __device__ __forceinline__ static const uint64_t *queue_masks(const uint64_t *pool, int wave) {
        return pool + kQueueBase + wave * kQueueWords;
    }

    // Applies the late primes to the n queued entries (n <= 128) and emits
    // the survivors; false when they did not fit the buffer.  Entry word
    // ts + rel of the line in slot k has offset (cb + 64 rel) mod p in that
    // line's string (rel < 2^26, so the sum stays below 2^32), and its
    // window is two funnel shifts of three 32-bit string words.
    __device__ __forceinline__ static bool drain(const uint64_t *pool, int wave, uint32_t n,
                                                 uint64_t ts, const LineDesc &L, const OutDesc &O,
                                                 int request, int lane) {
        const uint64_t *qm = queue_masks(pool, wave);
        const uint32_t *qw = reinterpret_cast<const uint32_t *>(qm + kQueue);
        uint64_t m[2];
        uint32_t rel6[2];
        uint32_t rs[2];  // first word of the entry's ring slot, from the ring base
        wave_sync();
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint32_t idx = lane + 64u * k;
            const bool valid = idx < n;
            const uint32_t tag = valid ? qw[idx] : 0u;
            m[k] = valid ? qm[idx] : 0ULL;
            rel6[k] = (tag & kRelMask) << 6;
            rs[k] = (tag >> kSlotShift) * static_cast<uint32_t>(kRingSlotWords);
        }
        const uint64_t *rb = pool + kRingBase;
        wave_sync();
        // Late primes in groups of kGroup: the group's reads are all in
        // flight together, and the wave-uniform exit is tested once per group
        // (a drain typically ends within a few primes).
        for (int s0 = FAST; s0 < kSelected; s0 += kGroup) {
            if (!__ballot((m[0] | m[1]) != 0)) {
                break;
            }
            uint64_t win[kGroup][2];
#pragma unroll
            for (int j = 0; j < kGroup; ++j) {
                const int s = s0 + j;
                win[j][0] = ~0ULL;
                win[j][1] = ~0ULL;
                if (s < kSelected) {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        if (m[k]) {
                            const uint64_t cst = rb[rs[k] + kRingConst + s - FAST];
                            const uint32_t pc = static_cast<uint32_t>(cst);
                            const uint32_t o = mod_small((pc >> 16) + rel6[k], pc & 0xffffu,
                                                         static_cast<uint32_t>(cst >> 32));
                            const uint32_t *d =
                                reinterpret_cast<const uint32_t *>(rb + rs[k] + (s - FAST) * 10)
                                + (o >> 5);
                            const uint32_t d0 = d[0];
                            const uint32_t d1 = d[1];
                            const uint32_t d2 = d[2];
                            const uint32_t lo = __builtin_amdgcn_alignbit(d1, d0, o & 31);
                            const uint32_t hi = __builtin_amdgcn_alignbit(d2, d1, o & 31);
                            win[j][k] = (static_cast<uint64_t>(hi) << 32) | lo;
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kGroup; ++j) {
                m[0] &= win[j][0];
                m[1] &= win[j][1];
            }
        }
        if (!__ballot((m[0] | m[1]) != 0)) {
            return true;
        }
        const uint64_t word[2] = {ts + (rel6[0] >> 6), ts + (rel6[1] >> 6)};
        const int den[2] = {static_cast<int>(rb[rs[0] + kRingDen]),
                            static_cast<int>(rb[rs[1] + kRingDen])};
        // The outcome is wave-uniform; saying so keeps the caller's loop
        // (and its row offsets) in scalar registers.
        return uniform(emit_words<2>(O, L.bound, m, word, den, request, lane) ? 1u : 0u) != 0;
    }

    // Cursor of the oldest queued entry: its line, and W + u from its word
    // offset r = (W - ts) + lane + 64 u (iterations start at multiples of
    // kStride from ts, and lane + 64 u < kStride).
    __device__ __forceinline__ static Cursor oldest(const uint64_t *pool, int wave, const Carry &q,
                                                    uint64_t ts) {
        wave_sync();
        const uint32_t *qw = reinterpret_cast<const uint32_t *>(queue_masks(pool, wave) + kQueue);
        const uint32_t r = uniform(qw[0]) & kRelMask;
        return Cursor{q.first_line,
                      ts + (r & ~static_cast<uint32_t>(kStride - 1)) + ((r & (kStride - 1)) >> 6)};
    }

    // Before run line next_line takes its ring slot (or, with `force`, at
    // the end of the run): drains the queue when its oldest entry's slot is
    // the one next_line takes, or when forced.  False (and `fail` set) when
    // the survivors did not fit.
    __device__ __forceinline__ static bool settle(const uint64_t *pool, uint64_t ts,
                                                  const LineDesc &L, const OutDesc &O, int request,
                                                  int lane, int wave, Carry &q, uint32_t next_line,
                                                  bool force, Cursor &fail) {
        if (q.queued == 0 || (!force && next_line - q.first_line < static_cast<uint32_t>(kRing))) {
            return true;
        }
        if (!drain(pool, wave, q.queued, ts, L, O, request, lane)) {
            fail = oldest(pool, wave, q, ts);
            return false;
        }
        q.queued = 0;
        return true;
    }

    // One wave sweeps [from, we) of run line `line` (ring slot `slot`);
    // `from` is the resume cursor (W + u) or the wave's range start.  False
    // (and `fail` set to the oldest uncommitted entry) when a drain's
    // survivors did not fit the buffer; words left in the queue at the end
    // stay there for a later drain.
    __device__ __forceinline__ static bool sweep_line(uint64_t *pool, int /*b*/, uint64_t ts,
                                                      uint64_t from, uint64_t we, int wave,
                                                      const LineDesc &L, const OutDesc &O,
                                                      int request, int lane, Carry &q,
                                                      uint32_t line, int slot, Cursor &fail) {
        const uint32_t *s_p = Base::small(pool, static_cast<int>(line & 1u));
        const uint32_t *s_base = s_p + 4 * kSelected;
        const uint32_t *s_magic = s_base + kSelected;
        // Iterations start at ts + k kStride, so the sub-word is the cursor's
        // offset from the tile start modulo the stride.
        const uint32_t skip0 = static_cast<uint32_t>((from - ts) & (kStride - 1));
        const uint64_t ws = from - skip0;
        typename Base::Offsets o =
            Base::row_offsets(s_p, s_base, s_magic, static_cast<uint32_t>(ws - ts), lane);
        // Iterations whose words all lie below min(we, last word) need no
        // masking; they run the unmasked body.  The first iteration of a
        // resumed sweep and the range's tail run the masked one.
        // The bounds are wave-uniform; saying so keeps the loops scalar
        // when the line descriptor was read from global memory (the
        // many-curve kernel), where the compiler cannot prove it.
        const uint64_t last_word = uniform64(L.words - 1);
        const uint32_t n_iter =
            uniform(static_cast<uint32_t>((we - ws + kStride - 1) / kStride));
        const uint64_t clean_end = we < last_word ? we : last_word;
        uint32_t n_clean = clean_end > ws
            ? static_cast<uint32_t>((clean_end - ws) / kStride) : 0u;
        n_clean = uniform(n_clean < n_iter ? n_clean : n_iter);
        const Line ln{line, static_cast<uint32_t>(slot) << kSlotShift};
        // Iteration it starts at word ws + it kStride, rel0 + it kStride from ts.
        const uint32_t rel0 = static_cast<uint32_t>(ws - ts);
        uint32_t it = 0;
        if (skip0 && n_iter) {
            if (!step<true>(pool, ws, rel0, we, last_word, skip0, o, q, ln, wave, ts, L, O,
                            request, lane)) {
                fail = oldest(pool, wave, q, ts);
                return false;
            }
            it = 1;
        }
        // The iteration count is wave-uniform; saying so keeps it (and the
        // loop control below) in SGPRs, where the select above would
        // otherwise have left it in a VGPR and cost vector instructions
        // every iteration.
        it = uniform(it);
        // Wave priority falls with the wave's progress through its range:
        // 3 in its first quarter down to 0 in its last.  The waves of a
        // workgroup sweep equal ranges, but the SIMDs issue the oldest waves
        // first, so without this the youngest wave on each SIMD finished
        // its range last (on 156,251-word lines the waves took 61k to 99k
        // clocks for equal work, and the line ended with the slowest); a
        // wave behind the others now takes the issue slots first.
        set_priority(3);
        const uint32_t quarter = (n_clean - it + 3) / 4;
        uint32_t mark = it + quarter;
        uint32_t stage = 0;
        for (; it < n_clean; ++it) {
            if (it == mark) {
                ++stage;
                mark += quarter;
                if (stage == 1) {
                    set_priority(2);
                } else if (stage == 2) {
                    set_priority(1);
                } else {
                    set_priority(0);
                }
            }
            if (!step<false>(pool, 0, rel0 + it * kStride, we, last_word, 0u, o, q, ln, wave,
                             ts, L, O, request, lane)) {
                set_priority(0);
                fail = oldest(pool, wave, q, ts);
                return false;
            }
        }
        set_priority(0);
        for (; it < n_iter; ++it) {
            if (!step<true>(pool, ws + static_cast<uint64_t>(it) * kStride, rel0 + it * kStride,
                            we, last_word, 0u, o, q, ln, wave, ts, L, O, request, lane)) {
                fail = oldest(pool, wave, q, ts);
                return false;
            }
        }
        return true;
    }

    struct Line {
        uint32_t index;  // run line
        uint32_t tag;    // ring slot, in entry-tag position
    };

    // One iteration at W: dense stage, (EDGE: masks for the range end, the
    // line's last word and the first `skip` sub-words), queue pushes and,
    // when 64 or more entries wait, a drain.  False when a drain's survivors
    // did not fit the buffer.
    // W is read only by the EDGE body; rel_w = W - ts.
    template <bool EDGE>
    __device__ __forceinline__ static bool step(uint64_t *pool, uint64_t W, uint32_t rel_w,
                                                uint64_t we, uint64_t last_word, uint32_t skip,
                                                typename Base::Offsets &o, Carry &q, const Line &ln,
                                                int wave, uint64_t ts, const LineDesc &L,
                                                const OutDesc &O, int request, int lane) {
        uint64_t m[U];
        Base::dense(pool, o, lane, m);
        if (EDGE) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t w = W + lane + 64 * u;
                if (w >= we || static_cast<uint32_t>(u) < skip) {
                    m[u] = 0;
                } else if (w == last_word) {
                    m[u] &= L.last_mask;
                }
            }
        }
        uint64_t any = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            any |= m[u];
        }
        const uint64_t any_bal = __ballot(any != 0);
        if (!any_bal) {
            return true;
        }
        uint64_t *qm = queue_masks(pool, wave);
        uint32_t *qw = reinterpret_cast<uint32_t *>(qm + kQueue);
        const uint32_t rel = (rel_w + lane) | ln.tag;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // With one word per lane that ballot is the word's own.
            const uint64_t bal = U == 1 ? any_bal : __ballot(m[u] != 0);
            if (!bal) {
                continue;
            }
            if (q.queued == 0) {
                q.first_line = ln.index;
            }
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                static_cast<uint32_t>(bal >> 32),
                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u));
            if (m[u]) {
                qm[q.queued + rank] = m[u];
                qw[q.queued + rank] = rel + 64u * u;
            }
            q.queued += static_cast<uint32_t>(__popcll(bal));
            if (q.queued >= 64) {
                if (!drain(pool, wave, q.queued, ts, L, O, request, lane)) {
                    return false;
                }
                q.queued = 0;
            }
        }
        return true;
    }
};

// ---------------------------------------------------------------------------
// Candidate constants of a curve in LDS (u32 arrays p, magic, (-A) mod p of
// 36 entries each) and the per-denominator prime selection from them.
// ---------------------------------------------------------------------------

// Threads 0..35 write the arrays; the caller syncs.
__device__ __forceinline__ void load_candidates(uint32_t *cp, const LineDesc &L,
                                                const int *primes, int tid) {
    if (tid < kCandidates) {
        const uint32_t p = static_cast<uint32_t>(primes[tid]);
        cp[tid] = p;
        cp[kCandidates + tid] = 0xffffffffu / p;  // floor(2^32 / p), p odd
        cp[2 * kCandidates + tid] = static_cast<uint32_t>(L.neg_bound_mod[tid]);
    }
}

// src/sieve_plan.cpp:75-86: the 32 largest candidates not dividing b.  Lane
// i tests candidate i; its slot is the number of qualifying candidates above
// it.  Returns the slot (or -1) with p, its magic and u = b mod p.
__device__ __forceinline__ int select_rank(const uint32_t *cp, uint32_t b, int lane,
                                           uint32_t &p, uint32_t &magic, uint32_t &u) {
    const bool cand_lane = lane < kCandidates;
    p = cand_lane ? cp[lane] : 1u;
    magic = cand_lane ? cp[kCandidates + lane] : 0u;
    u = cand_lane ? mod_small(b, p, magic) : 0u;
    const bool ok = cand_lane && u != 0;
    const unsigned long long mask = __ballot(ok);
    const int rank = ok ? __popcll(mask >> (lane + 1)) : kSelected;
    return rank < kSelected ? rank : -1;
}

// ---------------------------------------------------------------------------
// Runs of lines.  A row tile covers nr consecutive denominators of one curve
// (nr > 1 only when the tile is the whole line).  The curve's candidate
// constants stay in LDS for the tile.  A line takes two phases, each ended
// by a barrier:
//   A(r)  every wave builds its share of line r's dense rows; wave 0
//         selects the primes of line r + 1 and the last wave the candidate
//         indices of line r + 2
//   B(r)  every thread starts loading its string word of line r + 2 and
//         stores that of line r + 1 (loaded during B(r - 1)) into place;
//         every wave sweeps its share of line r
// The per-line arrays are kept for two lines (by parity) so that a
// selection can be written while the other line's arrays are read, and
// nothing waits on global memory: a string word is loaded a whole line
// before it is stored.
// ---------------------------------------------------------------------------
template <class Row>
struct RowRun {
    using Base = typename Row::Base;
    static constexpr int kThreads = Row::kWaves * 64;
    // u32 arrays after the row tile's LDS: candidate p, magic and (-A) mod p
    // [36 each], then for two lines (by parity) the selection's candidate
    // index and b mod p [32 each].
    static constexpr int kFastRows = Base::kFast;
    static constexpr int kCand = Row::kLdsWords;
    static constexpr int kLdsWords = kCand + (3 * kCandidates + 4 * kSelected + 1) / 2;
    static_assert(kSelected * kResidWords <= kThreads, "one string word per thread");
    static_assert(Row::kRing >= 2, "line r + 1 takes its slot while line r is swept");

    // Cursor of a wave within a run: line index above bit 40, word below.
    static constexpr int kLineShift = 40;
    static constexpr uint64_t kPosMask = (1ULL << kLineShift) - 1;

    __device__ __forceinline__ static uint32_t *cand(uint64_t *pool) {
        return reinterpret_cast<uint32_t *>(pool + kCand);
    }

    // Candidate index and b mod p of run line r's selection.
    __device__ __forceinline__ static uint32_t *next_sel(uint64_t *pool, int r) {
        return cand(pool) + 3 * kCandidates + (r & 1) * 2 * kSelected;
    }

    __device__ __forceinline__ static int slot_of(int r) {
        return r % Row::kRing;
    }

    __device__ __forceinline__ static void load_curve(uint64_t *pool, const LineDesc &L,
                                                      const int *primes, int tid) {
        load_candidates(cand(pool), L, primes, tid);
    }

    // Wave 0: the selection of run line r (denominator b) for the chunk at
    // ts into the row tile's arrays of r's parity, and the line's constants
    // into its ring slot.
    __device__ __forceinline__ static void select(uint64_t *pool, uint32_t b, uint64_t ts, int r,
                                                  int lane) {
        uint32_t *cp = cand(pool);
        uint32_t *s_p = Base::small(pool, r & 1);
        uint32_t p, magic, u;
        const int rank = select_rank(cp, b, lane, p, magic, u);
        // Default selection: rank s is candidate 35 - s for every dense s.
        const bool moved = lane >= kCandidates - kFastRows && lane < kCandidates
                           && rank != kCandidates - 1 - lane;
        const bool dflt = __ballot(moved) == 0;
        if (lane == 0) {
            s_p[Base::kDefaultFlag] = dflt ? 1u : 0u;
        }
        if (rank >= 0) {
            const uint32_t c = cp[2 * kCandidates + lane];
            const uint32_t ts_mod = static_cast<uint32_t>(ts % p);
            s_p[rank] = p;
            s_p[kSelected + rank] = c;
            s_p[2 * kSelected + rank] = u;
            s_p[3 * kSelected + rank] = static_cast<uint32_t>(lane);
            s_p[4 * kSelected + rank] = ts_mod;
            s_p[5 * kSelected + rank] = magic;
            // String offset of the tile's first word: c + 64 ts (mod p).
            Row::note_line(pool, slot_of(r), rank, p, magic,
                           mod_small(c + 64u * ts_mod, p, magic), b);
        }
    }

    // One wave: candidate index and u of run line r (denominator b).
    __device__ __forceinline__ static void select_next(uint64_t *pool, uint32_t b, int r,
                                                       int lane) {
        uint32_t *cp = cand(pool);
        uint32_t *nx = next_sel(pool, r);
        uint32_t p, magic, u;
        const int rank = select_rank(cp, b, lane, p, magic, u);
        if (rank >= 0) {
            nx[rank] = static_cast<uint32_t>(lane);
            nx[kSelected + rank] = u;
        }
    }

    // One string word of run line r per thread (threads below 288).
    __device__ __forceinline__ static uint64_t fetch(uint64_t *pool, const LineDesc &L, int r,
                                                     int tid) {
        if (tid >= kSelected * kResidWords) {
            return 0;
        }
        const uint32_t *nx = next_sel(pool, r);
        const int s = tid / kResidWords;
        const int w = tid - s * kResidWords;
        return L.resid[(static_cast<size_t>(nx[s]) * kMaxPrime + nx[kSelected + s])
                       * kResidWords + w];
    }

    __device__ __forceinline__ static void store(uint64_t *pool, uint64_t v, int r, int tid) {
        if (tid < kSelected * kResidWords) {
            const int s = tid / kResidWords;
            const int w = tid - s * kResidWords;
            Row::string(pool, s, slot_of(r))[w] = v;
        }
    }

    // Run start at line r0 of nr (first denominator t0): line r0's strings
    // and selection in place, the candidate indices of r0 + 1 chosen; returns
    // this thread's string word of line r0 + 1 (loading).  Ends with a
    // barrier.
    __device__ __forceinline__ static uint64_t begin(uint64_t *pool, const LineDesc &L, int t0,
                                                     int r0, int nr, uint64_t ts, int tid, int lane,
                                                     int wave) {
        const uint32_t b0 = static_cast<uint32_t>(denominator(L, static_cast<long long>(t0) + r0));
        if (wave == 0) {
            select_next(pool, b0, r0, lane);
        }
        __syncthreads();
        store(pool, fetch(pool, L, r0, tid), r0, tid);
        if (wave == 0) {
            select(pool, b0, ts, r0, lane);
        }
        if (wave == Row::kWaves - 1 && r0 + 1 < nr) {
            select_next(pool,
                        static_cast<uint32_t>(denominator(L, static_cast<long long>(t0) + r0 + 1)),
                        r0 + 1, lane);
        }
        __syncthreads();
        return r0 + 1 < nr ? fetch(pool, L, r0 + 1, tid) : 0;
    }

    // Phase A of run line r.  `flag`, when given, is read first (no sweep
    // runs in this phase, so every wave sees the same value) into
    // `flag_value`.  Ends with a barrier.
    __device__ __forceinline__ static void build_line(uint64_t *pool, const LineDesc &L, int t0,
                                                      int r, int nr, uint64_t ts,
                                                      const typename Base::Plan &plan, int tid,
                                                      int lane, int wave, const int *flag,
                                                      int &flag_value) {
        flag_value = flag ? static_cast<int>(uniform(static_cast<uint32_t>(*flag))) : 0;
        if (Base::kPlanned && uniform(Base::small(pool, r & 1)[Base::kDefaultFlag])) {
            Base::build_planned(pool, plan, tid, lane, wave);
        } else {
            if constexpr (Base::kWide) {
                Base::build_rows_dealt(pool, lane, wave, r & 1);
            } else {
                Base::build_rows(pool, tid, r & 1);
            }
        }
        if (wave == 0 && r + 1 < nr) {
            select(pool,
                   static_cast<uint32_t>(denominator(L, static_cast<long long>(t0) + r + 1)),
                   ts, r + 1, lane);
        }
        if (wave == Row::kWaves - 1 && r + 2 < nr) {
            select_next(pool,
                        static_cast<uint32_t>(denominator(L, static_cast<long long>(t0) + r + 2)),
                        r + 2, lane);
        }
        __syncthreads();
    }

    // Start of phase B of run line r: v (this thread's word of line r + 1)
    // into place; returns its word of line r + 2 (loading).
    __device__ __forceinline__ static uint64_t advance(uint64_t *pool, const LineDesc &L, int r,
                                                       int nr, uint64_t v, int tid) {
        const uint64_t nv = r + 2 < nr ? fetch(pool, L, r + 2, tid) : 0;
        if (r + 1 < nr) {
            store(pool, v, r + 1, tid);
        }
        return nv;
    }
};

// ---------------------------------------------------------------------------
// Short-line tile: up to NB whole lines of one curve, denominators t0 .. t0 +
// nb - 1, without row expansion.  The block is swept in units (chunk of 64
// words, line), chunk-major: lane l holds word 64 c + l, and a wave walks the
// lines of the block at that chunk.  A candidate's window offset
// (-A + 64 w) mod p depends on the word and the prime, not on the line, so a
// wave reduces the 36 offsets once per chunk; per line it only ANDs the
// windows of the line's selected primes (one address add and two funnel
// shifts each), with a wave-uniform early exit once every lane's word is
// empty.  All lanes of a wave read the same string, 18 consecutive dwords at
// most, so the reads are free of bank conflicts.
// ---------------------------------------------------------------------------
template <int WAVES, int NB>
struct DirectTile {
    // LDS (u64 words): residue strings [NB][36][10] by candidate index.  A
    // candidate the line does not select (it divides b, or 32 larger ones
    // were chosen) gets an all-ones string, so every line ANDs the windows
    // of all 36 candidates with the same result as its selected 32 alone.
    static constexpr int kStringWords = 10;
    static constexpr int kLdsWords = NB * kCandidates * kStringWords;

    // Units [0, total) dealt to the waves in contiguous ranges.
    __device__ __forceinline__ static void wave_range(uint32_t total, int wave, uint32_t &qs,
                                                      uint32_t &qe) {
        const uint32_t per = (total + WAVES - 1) / WAVES;
        qs = per * wave;
        qe = qs + per < total ? qs + per : total;
        if (qs > qe) {
            qs = qe;
        }
    }

    __device__ __forceinline__ static uint32_t units(int nb, const LineDesc &L) {
        return static_cast<uint32_t>(nb) * ((static_cast<uint32_t>(L.words) + 63u) / 64u);
    }

    // A wave's windows for one chunk.  The waves' unit ranges are the same
    // in every full block of a tile, so on short lines a wave keeps its
    // chunk, and its windows, from one block to the next.
    struct Windows {
        // Window of candidate k, 12 bits: byte offset of its first dword
        // (o / 32 * 4) above bit 5, funnel shift (o mod 32) below; two
        // candidates per register (k even in the low half).
        uint32_t win[kCandidates / 2];
        uint64_t m0;      // the chunk's valid bits of this lane's word
        uint32_t w;       // this lane's word
        uint32_t chunk;   // ~0u: none yet (a new tile starts with this)
    };

    // One wave sweeps units [qs, qe) of the block; cp holds the candidates'
    // p, floor(2^32 / p) and (-A) mod p (only the last is read: the primes
    // are compile-time constants here).  With K, the windows of the chunk
    // the wave starts in are taken from K when they are that chunk's, and
    // the last chunk's are left there.  Returns qe when done, or the unit
    // whose survivors did not fit.
    template <bool KEEP>
    __device__ __forceinline__ static uint32_t sweep(const uint64_t *pool, const uint32_t *cp,
                                                     int t0, int nb, uint32_t qs, uint32_t qe,
                                                     const LineDesc &L, const OutDesc &O,
                                                     int request, int lane, Windows *K) {
        const uint32_t words = static_cast<uint32_t>(L.words);
        const uint32_t lines = static_cast<uint32_t>(nb);
        uint32_t c = qs / lines;
        uint32_t l = qs - c * lines;
        // Window of candidate k, 12 bits: byte offset of its first dword
        // (o / 32 * 4) above bit 5, funnel shift (o mod 32) below; two
        // candidates per register (k even in the low half).
        uint32_t win[kCandidates / 2];
        uint64_t m0 = 0;
        uint32_t w = 0;
        bool fresh = true;
        if constexpr (KEEP) {
            if (K->chunk == c) {
                fresh = false;
#pragma unroll
                for (int i = 0; i < kCandidates / 2; ++i) {
                    win[i] = K->win[i];
                }
                m0 = K->m0;
                w = K->w;
            }
        }
        uint32_t v = qs;
        for (; v < qe; ++v) {
            if (fresh) {
                fresh = false;
                w = 64u * c + static_cast<uint32_t>(lane);
                const uint32_t x = 64u * w;
#pragma unroll
                for (int k = 0; k < kCandidates; k += 2) {
                    // Candidate k is the (35 - k)-th largest prime below 512.
                    const uint32_t p0 = largest_prime_below_512(kCandidates - 1 - k);
                    const uint32_t p1 = largest_prime_below_512(kCandidates - 2 - k);
                    const uint32_t o0 =
                        mod_small(cp[2 * kCandidates + k] + x, p0, 0xffffffffu / p0);
                    const uint32_t o1 =
                        mod_small(cp[2 * kCandidates + k + 1] + x, p1, 0xffffffffu / p1);
                    win[k / 2] = ((o0 >> 5) << 7) | (o0 & 31u) | (((o1 >> 5) << 7) << 16)
                                 | ((o1 & 31u) << 16);
                }
                m0 = w < words ? (w == words - 1 ? L.last_mask : ~0ULL) : 0ULL;
                if constexpr (KEEP) {
                    K->chunk = c;
                }
            }
            const char *str = reinterpret_cast<const char *>(
                pool + l * (kCandidates * kStringWords));
            uint64_t m[1] = {m0};
            // Largest primes first, so that the (at least four) unselected
            // candidates come last.  After 10 primes about 1 word in 16
            // still has bits, after 12 about 1 in 60: the all-empty test
            // pays every second prime from there on.
#pragma unroll
            for (int j = 0; j < kCandidates; ++j) {
                if (j >= 10 && (j & 1) == 0 && !__ballot(m[0] != 0)) {
                    break;
                }
                const int k = kCandidates - 1 - j;
                const uint32_t x = k & 1 ? win[k / 2] >> 16 : win[k / 2] & 0xfffu;
                const uint32_t *d = reinterpret_cast<const uint32_t *>(
                    str + k * (kStringWords * 8) + (x >> 5));
                const uint32_t lo = __builtin_amdgcn_alignbit(d[1], d[0], x);
                const uint32_t hi = __builtin_amdgcn_alignbit(d[2], d[1], x);
                m[0] &= (static_cast<uint64_t>(hi) << 32) | lo;
            }
            if (__ballot(m[0] != 0)) {
                const int b = denominator(L, static_cast<long long>(t0) + l);
                if (!emit_survivors<1>(O, L.bound, m, w, b, request, lane)) {
                    break;
                }
            }
            if (++l == lines) {
                l = 0;
                ++c;
                fresh = true;
            }
        }
        if constexpr (KEEP) {
#pragma unroll
            for (int i = 0; i < kCandidates / 2; ++i) {
                K->win[i] = win[i];
            }
            K->m0 = m0;
            K->w = w;
        }
        return v;
    }
};

// Runs of short lines: a tile covers many lines of one curve, taken in blocks
// of up to NB lines.  The curve's candidate constants stay in LDS for the
// tile; per block, one wave per line selects the primes and the workgroup
// copies the block's residue strings (a block's sweep is long next to that
// one global round trip).
template <int WAVES, int NB>
struct DirectRun {
    using Tile = DirectTile<WAVES, NB>;
    static constexpr int kThreads = WAVES * 64;
    static constexpr int kCand = Tile::kLdsWords;  // u32: p, magic, (-A) mod p [36 each]
    static constexpr int kNext = kCand + (3 * kCandidates + 1) / 2;  // u32: u [NB][36]
    static constexpr int kLdsWords = kNext + (NB * kCandidates + 1) / 2;

    // Cursor of a wave: block index above bit 40, pair position below.
    static constexpr int kBlockShift = 40;
    static constexpr uint64_t kPosMask = (1ULL << kBlockShift) - 1;

    __device__ __forceinline__ static uint32_t *cand(uint64_t *pool) {
        return reinterpret_cast<uint32_t *>(pool + kCand);
    }

    __device__ __forceinline__ static void load_curve(uint64_t *pool, const LineDesc &L,
                                                      const int *primes, int tid) {
        load_candidates(cand(pool), L, primes, tid);
    }

    // Set-up of block [t0, t0 + nb); every thread calls it after the
    // previous block's sweep.  One wave per line selects its primes and
    // records u = b mod p for each selected candidate (all ones for the
    // others); then the workgroup copies the strings into place.  Ends with
    // a barrier.
    __device__ __forceinline__ static void setup(uint64_t *pool, const LineDesc &L, int t0, int nb,
                                                 int tid, int lane, int wave) {
        const uint32_t *cp = cand(pool);
        uint32_t *nx = reinterpret_cast<uint32_t *>(pool + kNext);
        for (int bl = wave; bl < nb; bl += WAVES) {
            const uint32_t b = static_cast<uint32_t>(
                denominator(L, static_cast<long long>(t0) + bl));
            uint32_t p, magic, u;
            const int rank = select_rank(cp, b, lane, p, magic, u);
            if (lane < kCandidates) {
                nx[bl * kCandidates + lane] = rank >= 0 ? u : ~0u;
            }
        }
        __syncthreads();
        for (int j = tid; j < nb * kCandidates * kResidWords; j += kThreads) {
            const int slot = j / kResidWords;  // bl * 36 + k
            const int w = j - slot * kResidWords;
            const int k = slot - (slot / kCandidates) * kCandidates;
            const uint32_t u = nx[slot];
            pool[slot * Tile::kStringWords + w] =
                u == ~0u ? ~0ULL
                         : L.resid[(static_cast<size_t>(k) * kMaxPrime + u) * kResidWords + w];
        }
        __syncthreads();
    }
};

// ---------------------------------------------------------------------------
// Per-curve tables
// ---------------------------------------------------------------------------

// One workgroup of 512 threads per (candidate prime, curve).  Thread x
// evaluates f(x) mod p by Horner's rule on the coefficient residues (padded
// with zeros above the degree) and tests it against the squares mod p (zero
// counts as a square).
__global__ __launch_bounds__(512) void chi_kernel(
    const int *__restrict__ coef_res,  // [curves][36][n_coef]
    int n_coef, const int *__restrict__ primes,
    unsigned long long *__restrict__ chi) {  // [curves][36][8]
    __shared__ unsigned char square[kMaxPrime];
    __shared__ int res[kMaxCoefficients];
    const int i = blockIdx.x;
    const size_t curve = blockIdx.y;
    const int x = threadIdx.x;
    const int p = primes[i];
    square[x] = 0;
    for (int k = x; k < n_coef; k += blockDim.x) {
        res[k] = coef_res[(curve * kCandidates + i) * n_coef + k];
    }
    __syncthreads();
    if (x < p) {
        square[(x * x) % p] = 1;
    }
    __syncthreads();
    int bit = 0;
    if (x < p) {
        int v = 0;
        for (int k = n_coef - 1; k >= 0; --k) {
            v = (v * x + res[k]) % p;
        }
        bit = square[v];
    }
    unsigned long long word = __ballot(bit);
    if ((x & 63) == 0) {
        chi[(curve * kCandidates + i) * kChiWords + (x >> 6)] = word;
    }
}

// Residue strings: one workgroup of 4 waves per (prime index i, chunk of 64
// u, curve); the curve's chi_p bitmap in LDS, each wave writing the strings
// of 16 u.  Bit z of string u is chi_p((z mod p) * u^-1 mod p), and ballot
// packs 64 consecutive z per word; lane l walks z = l + 64 k, whose product
// with u^-1 advances by 64 u^-1 mod p from one word to the next (z and z - p
// give the same product mod p, so no reduction of z is needed).  For u = 0
// (p | b) and u >= p the string is never read and is stored as zero.
constexpr int kResidChunk = 64;  // u per workgroup
__global__ __launch_bounds__(256) void resid_kernel(
    const int *__restrict__ primes,
    const int *__restrict__ inverses,        // [36][512]
    const unsigned long long *__restrict__ chi,
    const int *__restrict__ slot,            // [curves] table slot, or null
    unsigned long long *__restrict__ resid) {  // [slots][36][512][9]
    __shared__ unsigned long long ch[kChiWords];
    const int i = blockIdx.x;
    const size_t curve = blockIdx.z;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = static_cast<int>(uniform(static_cast<uint32_t>(tid >> 6)));
    if (tid < kChiWords) {
        ch[tid] = chi[(curve * kCandidates + i) * kChiWords + tid];
    }
    __syncthreads();
    const uint32_t p = static_cast<uint32_t>(primes[i]);
    const uint32_t magic = static_cast<uint32_t>(0x100000000ULL / p);
    const size_t table = slot ? static_cast<size_t>(slot[curve]) : curve;
    unsigned long long *base = resid + table * kResidTableWords
        + static_cast<size_t>(i) * kMaxPrime * kResidWords;
    for (int j = wave; j < kResidChunk; j += 4) {
        const int u = static_cast<int>(blockIdx.y) * kResidChunk + j;
        unsigned long long *dst = base + static_cast<size_t>(u) * kResidWords;
        const uint32_t inv = static_cast<uint32_t>(inverses[i * kMaxPrime + u]);
        if (inv == 0) {  // u = 0 or u >= p
            if (lane < kResidWords) {
                dst[lane] = 0;
            }
            continue;
        }
        const uint32_t step = mod_small(64u * inv, p, magic);
        uint32_t y = mod_small(static_cast<uint32_t>(lane) * inv, p, magic);
        unsigned long long mine = 0;
#pragma unroll
        for (int k = 0; k < kResidWords; ++k) {
            const int bit = static_cast<int>((ch[y >> 6] >> (y & 63)) & 1);
            const unsigned long long word = __ballot(bit);
            if (lane == k) {
                mine = word;
            }
            const uint32_t next = y + step;
            y = next >= p ? next - p : next;
        }
        if (lane < kResidWords) {
            dst[lane] = mine;
        }
    }
}

__global__ void reset_state_kernel(DeviceState *state) {
    state->count = 0;
    state->first_fail = ~0ULL;
    state->full = 0;
    state->pad = 0;
}

// The state, tile counter and failure flags of a many-curve call, cleared by
// a kernel (not a memset) so that the call's device work can be recorded in
// a graph.
__global__ void reset_many_kernel(DeviceState *state, unsigned long long *head, int *failed,
                                  int n) {
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += stride) {
        failed[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state->count = 0;
        state->first_fail = ~0ULL;
        state->full = 0;
        state->pad = 0;
        *head = 0;
    }
}

// A curve descriptor read through the constant address space (scalar loads
// into SGPRs; the descriptors are not written while a kernel runs).
__device__ __forceinline__ CurveDesc scalar_curve(const CurveDesc *at) {
    static_assert(sizeof(CurveDesc) % 4 == 0, "whole dwords");
    constexpr int kDwords = static_cast<int>(sizeof(CurveDesc) / 4);
    const auto *src = reinterpret_cast<const __attribute__((address_space(4))) uint32_t *>(
        reinterpret_cast<uintptr_t>(at));
    uint32_t w[kDwords];
#pragma unroll
    for (int i = 0; i < kDwords; ++i) {
        w[i] = src[i];
    }
    CurveDesc c;
    __builtin_memcpy(&c, w, sizeof c);
    return c;
}

// ---------------------------------------------------------------------------
// Plain launches (one curve, one denominator batch)
// ---------------------------------------------------------------------------

// Tile = (run of L.tile_t denominators, chunk of the line).  Each wave keeps
// a cursor (line of the run, position) so that a launch after a full buffer
// resumes exactly where each wave stopped.
template <class Tile, int WAVES>
__global__ __launch_bounds__(WAVES * 64, 4) void sieve_kernel(
    const SieveParams P) {
    using Run = RowRun<Tile>;
    __shared__ __attribute__((aligned(16))) uint64_t pool[Run::kLdsWords];
    __shared__ unsigned s_first;
    const LineDesc &L = P.line;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = static_cast<int>(uniform(tid >> 6));
    const int tile = blockIdx.x;
    const int run = tile / L.tiles_per_t;
    const int chunk = tile - run * L.tiles_per_t;
    const int t0 = L.first_t + run * L.tile_t;
    const int rest = L.n_t - run * L.tile_t;
    const int nr = rest < L.tile_t ? rest : L.tile_t;
    const uint64_t ts = static_cast<uint64_t>(chunk) * L.tile_words;
    const uint64_t te = ts + L.tile_words < L.words ? ts + L.tile_words : L.words;
    uint64_t ws, we;
    Tile::wave_range(ts, te, wave, ws, we);
    unsigned long long *cursor = P.cursors + static_cast<size_t>(tile) * WAVES + wave;
    uint32_t cr = 0;
    uint64_t cpos = ws;
    if (P.resume) {
        const uint64_t c = uniform64(*cursor);
        cr = static_cast<uint32_t>(c >> Run::kLineShift);
        cpos = c & Run::kPosMask;
    }
    // Skip the tile when every wave has finished it; otherwise start at the
    // earliest line any wave still needs.
    if (tid == 0) {
        s_first = static_cast<unsigned>(nr);
    }
    Run::load_curve(pool, L, P.primes, tid);
    __syncthreads();
    if (lane == 0 && cr < static_cast<uint32_t>(nr)) {
        atomicMin(&s_first, cr);
    }
    __syncthreads();
    const int r0 = static_cast<int>(uniform(s_first));
    if (r0 >= nr) {
        return;
    }
    const typename Tile::Base::Plan plan = Tile::Base::make_plan(Run::cand(pool), tid, lane, wave);
    uint64_t v = Run::begin(pool, L, t0, r0, nr, ts, tid, lane, wave);
    bool failed = false;
    typename Tile::Carry carry;
    Cursor fail{0u, 0};
    for (int r = r0; r < nr; ++r) {
        int unused;
        Run::build_line(pool, L, t0, r, nr, ts, plan, tid, lane, wave, nullptr, unused);
        v = Run::advance(pool, L, r, nr, v, tid);
        if (!failed && static_cast<uint32_t>(r) >= cr && ws < we) {
            const uint64_t from = static_cast<uint32_t>(r) == cr ? cpos : ws;
            if (!Tile::sweep_line(pool, denominator(L, static_cast<long long>(t0) + r), ts, from,
                                  we, wave, L, P.out, 0, lane, carry, static_cast<uint32_t>(r),
                                  Run::slot_of(r), fail)) {
                failed = true;
            }
        }
        // Line r + 2 takes the ring slot of line r + 2 - kRing in the next
        // line's phases: queued words of that line are drained first.
        if (!failed && !Tile::settle(pool, ts, L, P.out, 0, lane, wave, carry,
                                     static_cast<uint32_t>(r + 2), false, fail)) {
            failed = true;
        }
        __syncthreads();
    }
    if (!failed && !Tile::settle(pool, ts, L, P.out, 0, lane, wave, carry,
                                 static_cast<uint32_t>(nr), true, fail)) {
        failed = true;
    }
    if (failed) {
        cr = fail.line;
        cpos = fail.pos;
    }
    if (lane == 0) {
        *cursor = failed ? (static_cast<uint64_t>(cr) << Run::kLineShift) | cpos
                         : static_cast<uint64_t>(nr) << Run::kLineShift;
    }
}

// Tile = run of L.tile_t short lines, swept in blocks of NB lines; each wave
// keeps a (block, pair) cursor.
// KEEP: each wave keeps its windows from block to block (on lines of a few
// chunks, where the reduction of 36 offsets per block costs as much as the
// block's sweep); on longer lines the registers that holds are worth more to
// the sweep.
template <int WAVES, int NB, bool KEEP>
__global__ __launch_bounds__(WAVES * 64, 4) void direct_kernel(const SieveParams P) {
    using Run = DirectRun<WAVES, NB>;
    using Tile = DirectTile<WAVES, NB>;
    __shared__ __attribute__((aligned(16))) uint64_t pool[Run::kLdsWords];
    __shared__ unsigned s_first;
    const LineDesc &L = P.line;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = static_cast<int>(uniform(tid >> 6));
    const int tile = blockIdx.x;
    const int t0 = L.first_t + tile * L.tile_t;
    const int rest = L.n_t - tile * L.tile_t;
    const int nr = rest < L.tile_t ? rest : L.tile_t;
    const int nblocks = (nr + NB - 1) / NB;
    unsigned long long *cursor = P.cursors + static_cast<size_t>(tile) * WAVES + wave;
    uint32_t ck = 0, cq = 0;
    if (P.resume) {
        const uint64_t c = uniform64(*cursor);
        ck = static_cast<uint32_t>(c >> Run::kBlockShift);
        cq = static_cast<uint32_t>(c & Run::kPosMask);
    }
    if (tid == 0) {
        s_first = static_cast<unsigned>(nblocks);
    }
    Run::load_curve(pool, L, P.primes, tid);
    __syncthreads();
    if (lane == 0 && ck < static_cast<uint32_t>(nblocks)) {
        atomicMin(&s_first, ck);
    }
    __syncthreads();
    const int k0 = static_cast<int>(uniform(s_first));
    if (k0 >= nblocks) {
        return;
    }
    bool failed = false;
    typename Tile::Windows windows;
    windows.chunk = ~0u;
    for (int k = k0; k < nblocks; ++k) {
        const int tk = t0 + k * NB;
        const int nb = nr - k * NB < NB ? nr - k * NB : NB;
        __syncthreads();  // the previous block's sweep is over
        Run::setup(pool, L, tk, nb, tid, lane, wave);
        uint32_t qs, qe;
        Tile::wave_range(Tile::units(nb, L), wave, qs, qe);
        if (!failed && static_cast<uint32_t>(k) >= ck && qs < qe) {
            const uint32_t from = P.resume && static_cast<uint32_t>(k) == ck ? cq : qs;
            if (from < qe) {
                const uint32_t stop = Tile::template sweep<KEEP>(
                    pool, Run::cand(pool), tk, nb, from, qe, L, P.out, 0, lane, &windows);
                if (stop != qe) {
                    failed = true;
                    ck = static_cast<uint32_t>(k);
                    cq = stop;
                }
            }
        }
    }
    if (lane == 0) {
        *cursor = failed ? (static_cast<uint64_t>(ck) << Run::kBlockShift) | cq
                         : static_cast<uint64_t>(nblocks) << Run::kBlockShift;
    }
}

// ---------------------------------------------------------------------------
// Persistent multi-curve launch
// ---------------------------------------------------------------------------

// Workgroups pull tile indices from M.head until all tiles are handed out or
// the survivor buffer is full.  A tile whose survivors do not fit marks its
// curve in curve_failed; the rest of that tile and later tiles of a marked
// curve are skipped (the host searches such curves again on the plain
// path).  Tiles of all curves share one index space, curves in the order the
// host chose; bucket[g] is the curve holding tile g << bucket_shift, so a
// tile's curve is found by a binary search over a few curves at most.
template <class Row, int WAVES>
__global__ __launch_bounds__(WAVES * 64, 4) void multi_kernel(
    const MultiParams M) {
    using Run = RowRun<Row>;
    using Direct = DirectTile<WAVES, kDirectMaxDenominators>;
    using DRun = DirectRun<WAVES, kDirectMaxDenominators>;
    constexpr int kWords = Run::kLdsWords > DRun::kLdsWords ? Run::kLdsWords
                                                             : DRun::kLdsWords;
    __shared__ __attribute__((aligned(16))) uint64_t pool[kWords];
    __shared__ unsigned long long s_tile;
    __shared__ int s_curve;
    __shared__ int s_skip;
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = static_cast<int>(uniform(tid >> 6));
    for (;;) {
        if (tid == 0) {
            unsigned long long t = ~0ULL;
            if (__hip_atomic_load(&M.out.state->full, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT) == 0) {
                t = atomicAdd(M.head, 1ULL);
            }
            int c = 0;
            int skip = 1;
            if (t < M.total_tiles) {
                const unsigned long long g = t >> M.bucket_shift;
                int lo = M.bucket[g];
                int hi = M.bucket[g + 1] + 1;  // tile_start[lo] <= t < tile_start[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (M.tile_start[mid] <= t) {
                        lo = mid;
                    } else {
                        hi = mid;
                    }
                }
                c = lo;
                skip = __hip_atomic_load(&M.curve_failed[c], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
            }
            s_tile = t;
            s_curve = c;
            s_skip = skip;
            s_fail = 0;
        }
        __syncthreads();
        const unsigned long long tile = uniform64(s_tile);
        const int c = static_cast<int>(uniform(static_cast<uint32_t>(s_curve)));
        const int skip = static_cast<int>(uniform(static_cast<uint32_t>(s_skip)));
        if (tile >= M.total_tiles) {
            break;
        }
        if (!skip) {
            // The descriptor is read through the constant address space:
            // scalar loads into SGPRs, so every value derived from it is
            // known to be wave-uniform and the sweep's loops stay scalar
            // (a vector load left them as per-lane loops).
            const CurveDesc C = scalar_curve(M.curves + c);
            const LineDesc &L = C.line;
            const unsigned long long local = tile - C.tile_start;
            bool failed = false;
            if (L.direct) {
                constexpr int NB = kDirectMaxDenominators;
                const int t0 = L.first_t + static_cast<int>(local) * L.tile_t;
                const int rest = L.n_t - static_cast<int>(local) * L.tile_t;
                const int nr = rest < L.tile_t ? rest : L.tile_t;
                const int nblocks = (nr + NB - 1) / NB;
                DRun::load_curve(pool, L, M.primes, tid);
                for (int k = 0; k < nblocks; ++k) {
                    const int tk = t0 + k * NB;
                    const int nb = nr - k * NB < NB ? nr - k * NB : NB;
                    __syncthreads();  // the previous block's sweep is over
                    // s_fail is only written during sweeps, so every wave
                    // reads the same value here.
                    if (uniform(static_cast<uint32_t>(s_fail))) {
                        break;
                    }
                    DRun::setup(pool, L, tk, nb, tid, lane, wave);
                    uint32_t qs, qe;
                    Direct::wave_range(Direct::units(nb, L), wave, qs, qe);
                    if (qs < qe
                        && Direct::template sweep<false>(pool, DRun::cand(pool), tk, nb, qs,
                                                         qe, L, M.out, C.request, lane,
                                                         nullptr) != qe) {
                        failed = true;
                        s_fail = 1;  // every lane: a lane-0 branch here
                                     // made the sweep loops per-lane
                    }
                }
            } else {
                const int run = static_cast<int>(local / static_cast<unsigned>(L.tiles_per_t));
                const int chunk = static_cast<int>(local) - run * L.tiles_per_t;
                const int t0 = L.first_t + run * L.tile_t;
                const int rest = L.n_t - run * L.tile_t;
                const int nr = rest < L.tile_t ? rest : L.tile_t;
                const uint64_t ts = static_cast<uint64_t>(chunk) * L.tile_words;
                const uint64_t te = ts + L.tile_words < L.words ? ts + L.tile_words : L.words;
                uint64_t ws, we;
                Row::wave_range(ts, te, wave, ws, we);
                Run::load_curve(pool, L, M.primes, tid);
                __syncthreads();
                const typename Row::Base::Plan plan =
                    Row::Base::make_plan(Run::cand(pool), tid, lane, wave);
                uint64_t v = Run::begin(pool, L, t0, 0, nr, ts, tid, lane, wave);
                typename Row::Carry carry;
                Cursor unused_cursor{0u, 0};
                bool stopped = false;
                for (int r = 0; r < nr; ++r) {
                    int stop_run;
                    Run::build_line(pool, L, t0, r, nr, ts, plan, tid, lane, wave, &s_fail,
                                    stop_run);
                    if (stop_run) {
                        stopped = true;
                        break;
                    }
                    v = Run::advance(pool, L, r, nr, v, tid);
                    if (!failed && ws < we
                        && !Row::sweep_line(pool, denominator(L, static_cast<long long>(t0) + r),
                                            ts, ws, we, wave, L, M.out, C.request, lane, carry,
                                            static_cast<uint32_t>(r), Run::slot_of(r),
                                            unused_cursor)) {
                        failed = true;
                        s_fail = 1;  // every lane: a lane-0 branch here
                                     // made the sweep loops per-lane
                    }
                    // Queued words of the line whose ring slot line r + 2
                    // takes are drained first; s_fail, set before the
                    // barrier, is read by every wave at the next phase A.
                    if (!failed && !Row::settle(pool, ts, L, M.out, C.request, lane, wave, carry,
                                                static_cast<uint32_t>(r + 2), false,
                                                unused_cursor)) {
                        failed = true;
                        s_fail = 1;  // every lane: a lane-0 branch here
                                     // made the sweep loops per-lane
                    }
                    __syncthreads();
                }
                // A stopped run's curve is marked failed and searched again
                // on the plain path, so its queued words are dropped.
                if (!failed && !stopped
                    && !Row::settle(pool, ts, L, M.out, C.request, lane, wave, carry,
                                    static_cast<uint32_t>(nr), true, unused_cursor)) {
                    failed = true;
                }
            }
            if (failed && lane == 0) {
                atomicOr(&M.curve_failed[c], 1);
            }
        }
        __syncthreads();
    }
}

namespace {

template <class Tile, int WAVES>
hipError_t launch_variant(const SieveParams &params, unsigned grid,
                          hipStream_t stream) {
    hipLaunchKernelGGL((sieve_kernel<Tile, WAVES>), dim3(grid), dim3(WAVES * 64), 0,
                       stream, params);
    return hipGetLastError();
}

template <class Tile, int WAVES>
hipError_t launch_multi_variant(const MultiParams &params, unsigned grid,
                                hipStream_t stream) {
    hipLaunchKernelGGL((multi_kernel<Tile, WAVES>), dim3(grid), dim3(WAVES * 64), 0,
                       stream, params);
    return hipGetLastError();
}

template <class Tile, int WAVES>
hipError_t multi_occupancy(int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, multi_kernel<Tile, WAVES>,
                                                        WAVES * 64, 0);
}

// "q_u2_f12_w8": RowTileQ with U = 2, FAST = 12, 8 waves; "w_u2_f14_w16":
// the wide-row tile with FAST = 14, 16 waves.
#define RPS_VARIANT_Q(U, F, W)                                             \
    {"q_u" #U "_f" #F "_w" #W, U, F, W, 17, 0,                            \
     launch_variant<RowTileQ<U, F, W>, W>,                                 \
     launch_multi_variant<RowTileQ<U, F, W>, W>,                           \
     multi_occupancy<RowTileQ<U, F, W>, W>}

#define RPS_VARIANT_W(F, W)                                                \
    {"w_u2_f" #F "_w" #W, 2, F, W, 20, 1,                                  \
     launch_variant<RowTileQ<2, F, W, true>, W>,                           \
     launch_multi_variant<RowTileQ<2, F, W, true>, W>,                     \
     multi_occupancy<RowTileQ<2, F, W, true>, W>}
const VariantInfo kVariants[] = {
    RPS_VARIANT_Q(2, 12, 8),
    RPS_VARIANT_Q(2, 11, 8),
    RPS_VARIANT_Q(2, 10, 8),
    RPS_VARIANT_Q(1, 12, 8),
    RPS_VARIANT_Q(2, 12, 16),
    RPS_VARIANT_W(14, 16),
    RPS_VARIANT_W(13, 16),
    RPS_VARIANT_Q(1, 11, 8),
    RPS_VARIANT_Q(1, 13, 8),
    RPS_VARIANT_Q(1, 14, 8),
};

#undef RPS_VARIANT_Q
#undef RPS_VARIANT_W

}  // namespace

int variant_count() {
    return static_cast<int>(sizeof(kVariants) / sizeof(kVariants[0]));
}

const VariantInfo *variant(int index) {
    if (index < 0 || index >= variant_count()) {
        return nullptr;
    }
    return &kVariants[index];
}

hipError_t launch_chi(const int *coef_res, int n_coef, int curves,
                      const int *primes, unsigned long long *chi,
                      hipStream_t stream) {
    hipLaunchKernelGGL(chi_kernel, dim3(kCandidates, curves), dim3(512), 0, stream,
                       coef_res, n_coef, primes, chi);
    return hipGetLastError();
}

hipError_t launch_resid(const int *primes, const int *inverses,
                        const unsigned long long *chi, int curves, const int *slot,
                        unsigned long long *resid, hipStream_t stream) {
    static_assert(kMaxPrime % kResidChunk == 0, "whole chunks of u");
    hipLaunchKernelGGL(resid_kernel, dim3(kCandidates, kMaxPrime / kResidChunk, curves),
                       dim3(256), 0, stream, primes, inverses, chi, slot, resid);
    return hipGetLastError();
}

hipError_t launch_direct(const SieveParams &params, unsigned grid,
                         hipStream_t stream) {
    // Measured on lines of 128 and 256 words: 7.5 and 10.8 Tsite/s keeping
    // the windows, 6.6 and 10.3 without; from 512 words the kernel without
    // is faster (13.7 against 13.0, and 15.3 against 13.2 at 1,024).
    if (params.line.words <= 256) {
        hipLaunchKernelGGL((direct_kernel<kDirectWaves, kDirectMaxDenominators, true>),
                           dim3(grid), dim3(kDirectWaves * 64), 0, stream, params);
    } else {
        hipLaunchKernelGGL((direct_kernel<kDirectWaves, kDirectMaxDenominators, false>),
                           dim3(grid), dim3(kDirectWaves * 64), 0, stream, params);
    }
    return hipGetLastError();
}

// The allocation-free entry point's partial results (residue table, chi
// bitmaps, coefficient residues, flush state, cursors) live here, one copy
// per device, for the life of the process.  Host code never passes the
// array's name to a kernel: it holds no device address on the host side.
// scratch_init stores the address the runtime reports for the current
// device, and launches use that.
static __device__ __align__(256) unsigned char scratch[scratch_layout::kBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_lock;
void *g_scratch[kMaxScratchDevices] = {};
}  // namespace

hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) {
        return hipErrorInvalidDevice;
    }
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (g_scratch[device]) {
        return hipSuccess;
    }
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        return e;
    }
    g_scratch[device] = ptr;
    return hipSuccess;
}

void *scratch_address(int device) {
    if (device < 0 || device >= kMaxScratchDevices) {
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    return g_scratch[device];
}

hipError_t launch_reset(DeviceState *state, hipStream_t stream) {
    hipLaunchKernelGGL(reset_state_kernel, dim3(1), dim3(1), 0, stream, state);
    return hipGetLastError();
}

hipError_t launch_reset_many(DeviceState *state, unsigned long long *head, int *failed, int n,
                             hipStream_t stream) {
    const unsigned blocks = static_cast<unsigned>(
        n <= 0 ? 1 : (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(reset_many_kernel, dim3(blocks), dim3(256), 0, stream, state, head,
                       failed, n);
    return hipGetLastError();
}

}  // namespace rps
