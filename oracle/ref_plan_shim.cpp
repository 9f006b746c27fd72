// Test infrastructure only: a C entry point over the reference's own host
// sieve plan (src/sieve_plan.cpp, compiled unchanged from the reference
// checkout), so tests can compare the oracle's prime selection, affine masks
// and modular inverses with the reference's code.  Nothing here is shipped.
#include <cstdint>
#include <exception>
#include <string>
#include <vector>

#include "sieve_plan.hpp"

using ratpoints_gpu::DenominatorRange;
using ratpoints_gpu::sieve_detail::SievePlan;

extern "C" {

// Selected candidate-prime values for every parameter in [first, last]:
// out[(t - first) * 32 + k].  Returns 0, or -1 if the plan threw.
int ref_plan_selected_primes(int first, int last, int square, int *out) {
    try {
        SievePlan plan(100, DenominatorRange{first, last}, square != 0);
        for (size_t i = 0; i < plan.selected_primes.size(); ++i) {
            out[i] = plan.primes[plan.selected_primes[i]];
        }
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// Word count and LDS byte count the plan derives from the bound.
int ref_plan_sizes(long long bound, long long *word_count,
                   long long *shared_mask_bytes) {
    try {
        SievePlan plan(bound, DenominatorRange{1, 1}, false);
        *word_count = plan.word_count;
        *shared_mask_bytes = static_cast<long long>(plan.shared_mask_bytes);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

long long ref_mod_inverse(long long value, long long modulus) {
    return ratpoints_gpu::sieve_detail::mod_inverse(value, modulus);
}

int ref_affine_mask(int prime, const char *const *coefficients, int count,
                    unsigned char *out) {
    ratpoints_gpu::Coefficients coeffs(coefficients, coefficients + count);
    std::vector<uint8_t> mask =
        ratpoints_gpu::sieve_detail::build_affine_mask(prime, coeffs);
    for (int x = 0; x < prime; ++x) {
        out[x] = mask[static_cast<size_t>(x)];
    }
    return 0;
}

}  // extern "C"
