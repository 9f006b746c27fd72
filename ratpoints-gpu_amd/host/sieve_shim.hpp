// Streaming form of the sieve used by PointSearch: one call sieves the whole
// denominator range in batches and reports each completed batch, so the GPU
// sieves batch i+1 while the host verifies batch i.  No counterpart in the
// reference, whose PointSearch calls run_modular_sieve once per batch.
#ifndef RATPOINTS_GPU_SIEVE_SHIM_HPP
#define RATPOINTS_GPU_SIEVE_SHIM_HPP

#include <cstdint>
#include <functional>

#include "ratpoints_sieve.h"
#include "sieve.hpp"

namespace ratpoints_gpu {

// Called after every survivor of [range.first, range.last] was delivered;
// return false to stop the search.
using BatchDoneCallback =
    std::function<bool(DenominatorRange range, std::uint64_t survivors)>;

// The calling thread's context on its active device (opened on first use).
rps_ctx *thread_sieve_context();

SieveResult run_sieve_stream(const Coefficients &coefficients,
                             long long numerator_bound,
                             DenominatorRange denominators,
                             const CandidateBatchCallback &callback,
                             const BatchDoneCallback *done,
                             bool square_denominators, int batch_size,
                             bool *cancelled);

}  // namespace ratpoints_gpu

#endif
