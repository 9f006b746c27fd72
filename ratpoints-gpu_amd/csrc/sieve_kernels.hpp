// Host-side declarations of the gfx950 sieve kernels' launchers.
#ifndef RATPOINTS_GPU_AMD_SIEVE_KERNELS_HPP
#define RATPOINTS_GPU_AMD_SIEVE_KERNELS_HPP

#include <hip/hip_runtime_api.h>

#include "sieve_layout.hpp"

namespace rps {

// A compiled instantiation of the row-tile sieve (plain and multi-curve).
struct VariantInfo {
    const char *name;
    int words_per_lane;  // U: a wave covers 64 U words per iteration
    int fast_primes;     // primes ANDed for every word
    int waves;           // waves per workgroup
    int run_log2;        // row tiles take whole lines up to 2^run_log2 words
    int wide;            // 16-byte row elements, one workgroup per CU
    hipError_t (*launch)(const SieveParams &params, unsigned grid,
                         hipStream_t stream);
    hipError_t (*launch_multi)(const MultiParams &params, unsigned grid,
                               hipStream_t stream);
    // Workgroups of the persistent kernel resident per CU.
    hipError_t (*multi_blocks_per_cu)(int *blocks);
};

int variant_count();
const VariantInfo *variant(int index);

// chi_p bitmaps and residue strings of `curves` curves; coefficient
// residues are [curves][36][n_coef], zero-padded above each degree.  Curve
// i's residue table goes to table slot[i] of `resid` (table i when slot is
// null).
hipError_t launch_chi(const int *coef_res, int n_coef, int curves,
                      const int *primes, unsigned long long *chi,
                      hipStream_t stream);
// `inverses` is [36][512]: u^-1 mod p for each candidate p and 0 < u < p,
// 0 elsewhere (the same for every curve, computed once per context).
hipError_t launch_resid(const int *primes, const int *inverses,
                        const unsigned long long *chi, int curves, const int *slot,
                        unsigned long long *resid, hipStream_t stream);
hipError_t launch_reset(DeviceState *state, hipStream_t stream);
// Clears a many-curve call's state, tile counter and n failure flags.
hipError_t launch_reset_many(DeviceState *state, unsigned long long *head, int *failed, int n,
                             hipStream_t stream);

// Short-line kernel: tiles of up to kDirectMaxDenominators whole lines of at
// most kDirectMaxWords words.
constexpr int kDirectWaves = 8;
constexpr int kDirectMaxDenominators = 16;
constexpr unsigned kDirectMaxWords = 2048;
hipError_t launch_direct(const SieveParams &params, unsigned grid,
                         hipStream_t stream);

// Device scratch of the allocation-free entry point: one file-scope array of
// scratch_layout::kBytes per device (layout in sieve_layout.hpp).  scratch_init
// resolves the array's address on `device`, which must be the current
// device; call it once per device at set-up.  scratch_address returns that
// address, or nullptr when scratch_init has not run for the device.
hipError_t scratch_init(int device);
void *scratch_address(int device);

}  // namespace rps

#endif
