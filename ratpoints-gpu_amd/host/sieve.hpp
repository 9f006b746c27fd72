// Search-backend interface: the reference's include/sieve.hpp:13-47 with the
// same types and signature, implemented by sieve_shim.cpp over the gfx950 C
// ABI (include/ratpoints_sieve.h).
#ifndef RATPOINTS_GPU_SIEVE_HPP
#define RATPOINTS_GPU_SIEVE_HPP

#include <cstddef>
#include <functional>
#include <vector>

#include "search_types.hpp"

namespace ratpoints_gpu {

struct SieveMetrics {
    float basis_ms = 0.0f;
    float sieve_ms = 0.0f;
    double plan_ms = 0.0;
    double workspace_ms = 0.0;
    double basis_wall_ms = 0.0;
    double survivor_setup_ms = 0.0;
    double verification_ms = 0.0;
    double sieve_wall_ms = 0.0;
    unsigned long long word_count = 0;
    long double initial_mask_bytes = 0.0L;
};

struct SieveResult {
    unsigned long long survivor_count = 0;
    SieveMetrics metrics;
};

// Raw modular survivors (a, b) of one flush; b is k^2 in square mode.
struct CandidateBatch {
    std::vector<long long> numerators;
    std::vector<int> denominators;

    std::size_t size() const { return numerators.size(); }
};

using CandidateBatchCallback = std::function<void(const CandidateBatch &)>;

// Streams every modular survivor of [-A, A] x denominators to `callback`.
// The batch is valid only during the call.  While the callback runs, the GPU
// already sieves the next denominator batch.  Errors are exceptions:
// std::invalid_argument for invalid input, std::runtime_error ("<op>:
// <message>") for device failures, std::overflow_error and std::logic_error
// for counter overflow and broken invariants.
SieveResult run_modular_sieve(const Coefficients &coefficients,
                              long long numerator_bound,
                              DenominatorRange denominators,
                              const CandidateBatchCallback &callback,
                              bool square_denominators = false);

}  // namespace ratpoints_gpu

#endif
