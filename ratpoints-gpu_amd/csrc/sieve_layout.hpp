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
//   * a row tile copies the strings of the 32 selected primes of its
//     denominator into LDS and expands the dense primes into rows of
//     p + 64 U - 1 words, so a wave reads 64 U consecutive row words with a
//     wave-uniform start and never wraps; a short-line tile keeps only the
//     strings and evaluates windows directly.
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
// Residue-table words per curve: [36][512][9].
constexpr unsigned long long kResidTableWords =
    static_cast<unsigned long long>(kCandidates) * kMaxPrime * kResidWords;

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

// One curve's search: its lines, tables and tiling.
struct LineDesc {
    long long bound;                     // A
    unsigned long long words;            // 64-bit words per line
    unsigned long long last_mask;        // valid bits of the line's last word
    unsigned long long tile_words;       // words per row tile (line chunk)
    const unsigned long long *resid;     // [36][512][9] residue strings
    const int *neg_bound_mod;            // [36] (-A) mod p
    int first_t;                         // first denominator parameter
    int n_t;                             // denominator parameters
    int square;                          // b = t^2
    int direct;                          // short lines: direct tiles
    int tiles_per_t;                     // row tiles per denominator
    int tile_t;                          // denominators per tile (a run)
};

// Where survivors go.
struct OutDesc {
    long long *num;
    int *den;
    int *req;                            // request index per survivor, or null
    DeviceState *state;
    unsigned long long capacity;
    int primitive;                       // drop survivors with gcd(a, b) > 1
    int pad;
};

// Plain launch: one curve, one denominator batch, one workgroup per tile.
struct SieveParams {
    LineDesc line;
    OutDesc out;
    const int *primes;                   // [36] ascending
    unsigned long long *cursors;         // [tiles][waves] next word per wave
    int resume;                          // start every wave at its cursor
    int pad;
};

// Persistent multi-curve launch: workgroups pull tiles from `head`.
struct CurveDesc {
    LineDesc line;
    unsigned long long tile_start;       // first tile of this curve
    int request;                         // index reported with survivors
    int pad;
};

struct MultiParams {
    const CurveDesc *curves;             // [n_curves], in tile order
    const unsigned long long *tile_start;  // [n_curves + 1] tile prefix
    const int *bucket;                   // [buckets + 1] curve of tile g << bucket_shift
    int *curve_failed;                   // [n_curves] survivor reservation failed
    unsigned long long *head;            // next tile to hand out
    unsigned long long total_tiles;
    OutDesc out;
    const int *primes;
    int n_curves;
    int bucket_shift;
};

// Layout of the allocation-free entry point's device scratch (byte offsets,
// 256-aligned): one curve's residue table, chi bitmaps and coefficient
// residues, (-A) mod p, the flush state, and the per-wave cursors of at most
// kMaxTiles tiles of at most kMaxWaves waves.
namespace scratch_layout {
constexpr unsigned long long align256(unsigned long long n) { return (n + 255) / 256 * 256; }
constexpr unsigned long long kMaxTiles = 16384;
constexpr unsigned long long kMaxWaves = 16;
constexpr unsigned long long kResid = 0;
constexpr unsigned long long kChi = align256(kResid + kResidTableWords * 8);
constexpr unsigned long long kCoef =
    align256(kChi + static_cast<unsigned long long>(kCandidates) * kChiWords * 8);
constexpr unsigned long long kNegMod =
    align256(kCoef + static_cast<unsigned long long>(kCandidates) * kMaxCoefficients * 4);
constexpr unsigned long long kState = align256(kNegMod + kCandidates * 4);
constexpr unsigned long long kCursors = align256(kState + sizeof(DeviceState));
constexpr unsigned long long kBytes = align256(kCursors + kMaxTiles * kMaxWaves * 8);
}  // namespace scratch_layout

}  // namespace rps

#endif
