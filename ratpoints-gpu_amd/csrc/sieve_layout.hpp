// Data layout shared by the gfx950 sieve kernels and their host launcher.
//
// The survivor set is the one defined by the reference's host sieve plan and
// CUDA sieve (paths relative to the reference repository):
//   candidate primes: the 36 largest primes below 512 (src/sieve_plan.cpp:60-61)
//   per-denominator selection of the 32 largest candidates not dividing b
//   (src/sieve_plan.cpp:75-86); chi_p(x) = [f(x) mod p is a square or zero]
//   (src/sieve_plan.cpp:115-138); (a, b) survives iff chi_p(a * b^-1 mod p)
//   for every selected p (src/sieve.cu:156-170, 302-331).
//
// The layout is this build's own:
//   * numerators are packed 64 per 64-bit word: word w holds
//     a = -A + 64 w + j, j = 0..63; a line has floor((2A + 64) / 64) words;
//   * per curve, a residue table holds, for every candidate prime p and every
//     u = b mod p, the 576-bit string E(z) = chi_p((z mod p) * u^-1 mod p),
//     z = 0..575 (nine 64-bit words).  A 64-bit window of E starting at bit
//     (-A + 64 w) mod p is exactly word w of the line for denominator b;
//   * the sieve kernel copies the strings of the 32 selected primes of its
//     denominator into LDS and expands the dense primes into rows of
//     p + 64 U - 1 words, so a wave reads 64 U consecutive row words with a
//     wave-uniform start and never wraps.
#ifndef RATPOINTS_GPU_AMD_SIEVE_LAYOUT_HPP
#define RATPOINTS_GPU_AMD_SIEVE_LAYOUT_HPP

#include <cstdint>

namespace rps {

constexpr int kCandidates = 36;     // candidate primes: 293 .. 509
constexpr int kSelected = 32;       // primes applied to each denominator
constexpr int kMaxPrime = 512;      // every candidate is below this
constexpr int kChiWords = 8;        // 512-bit chi_p bitmap per prime
constexpr int kResidWords = 9;      // 576-bit residue string per (p, u)
constexpr int kMaxCoefficients = 101;  // degree <= 100
constexpr int kMaxParamT = 46340;   // largest t with t^2 < 2^31

// Flush state of one survivor buffer.  Reservations are contiguous ranges
// handed out by an atomic add on `count`; a wave whose range does not fit
// records its start in `first_fail` (atomic min) and stops, so the written
// survivors are exactly [0, min(count, first_fail)).
struct DeviceState {
    unsigned long long count;
    unsigned long long first_fail;
    unsigned int full;
    unsigned int pad;
};

struct SieveParams {
    long long bound;                 // A
    unsigned long long words;        // 64-bit words per denominator line
    unsigned long long last_mask;    // valid bits of the line's last word
    unsigned long long tile_words;   // words per tile (a chunk of one line)
    unsigned long long capacity;     // survivor slots in out_num / out_den
    int first_t;                     // first denominator parameter
    int n_t;                         // parameters in this launch
    int tiles_per_t;                 // chunks per denominator line
    int square;                      // b = t^2
    int resume;                      // start every wave at its cursor
    int tile_t;                      // denominators per short-line tile
    unsigned magic_words;            // floor(2^32 / words), short lines
    int pad;
    int primes[kCandidates];         // ascending
    int neg_bound_mod[kCandidates];  // (-A) mod p
    const unsigned long long *resid; // [36][512][9] residue strings
    long long *out_num;
    int *out_den;
    DeviceState *state;
    unsigned long long *cursors;     // [tiles][waves] next word per wave
};

}  // namespace rps

#endif
