// ratpoints_gpu::run_modular_sieve over the gfx950 C ABI.
//
// Each thread keeps one rps_ctx for its active device, so the per-curve
// residue table and the survivor buffers stay resident across calls (the
// reference rebuilt both on every call, src/sieve.cu:503-580).  Status codes
// map back to the reference's exception types (include/cuda_support.hpp:14-19,
// src/sieve_plan.cpp:55-58, src/sieve.cu:545-551).
#include "sieve_shim.hpp"

#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "gpu_devices.hpp"

namespace ratpoints_gpu {
namespace {

struct ThreadContext {
    rps_ctx *ctx = nullptr;
    int device = -1;

    ~ThreadContext() { rps_close(ctx); }
};

thread_local ThreadContext thread_context;

[[noreturn]] void raise_status(int code, const std::string &message) {
    switch (code) {
    case RPS_EINVAL:
        throw std::invalid_argument(message);
    case RPS_EOVERFLOW:
        throw std::overflow_error(message);
    case RPS_EINTERNAL:
        throw std::logic_error(message);
    case RPS_ENOMEM:
        throw std::runtime_error(message.empty() ? "out of memory" : message);
    default:
        throw std::runtime_error(message);
    }
}

struct StreamState {
    const CandidateBatchCallback *callback = nullptr;
    const BatchDoneCallback *done = nullptr;
    CandidateBatch batch;
    std::exception_ptr error;
};

int deliver(void *user, const int64_t *num, const int32_t *den, size_t n) {
    StreamState &state = *static_cast<StreamState *>(user);
    try {
        state.batch.numerators.assign(num, num + n);
        state.batch.denominators.assign(den, den + n);
        (*state.callback)(state.batch);
        return 0;
    } catch (...) {
        state.error = std::current_exception();
        return 1;
    }
}

int batch_done(void *user, int32_t first, int32_t last, uint64_t survivors) {
    StreamState &state = *static_cast<StreamState *>(user);
    try {
        return (*state.done)(DenominatorRange{first, last}, survivors) ? 0 : 1;
    } catch (...) {
        state.error = std::current_exception();
        return 1;
    }
}

}  // namespace

rps_ctx *thread_sieve_context() {
    const int device = active_gpu();
    if (thread_context.ctx && thread_context.device == device) {
        return thread_context.ctx;
    }
    rps_close(thread_context.ctx);
    thread_context.ctx = nullptr;
    rps_ctx *ctx = nullptr;
    const int rc = rps_open(device, &ctx);
    if (rc != RPS_OK) {
        raise_status(rc, rps_last_error(nullptr));
    }
    thread_context.ctx = ctx;
    thread_context.device = device;
    return ctx;
}

SieveResult run_sieve_stream(const Coefficients &coefficients,
                             long long numerator_bound,
                             DenominatorRange denominators,
                             const CandidateBatchCallback &callback,
                             const BatchDoneCallback *done,
                             bool square_denominators, int batch_size,
                             bool *cancelled) {
    rps_ctx *ctx = thread_sieve_context();
    std::vector<const char *> text;
    text.reserve(coefficients.size());
    for (const auto &c : coefficients) {
        text.push_back(c.c_str());
    }
    if (rps_set_option(ctx, "batch_size", batch_size) != RPS_OK) {
        raise_status(RPS_EINVAL, rps_last_error(ctx));
    }
    rps_request request{text.data(), static_cast<int32_t>(text.size()),
                        numerator_bound, denominators.first,
                        denominators.last, square_denominators ? 1 : 0};
    StreamState state;
    state.callback = &callback;
    state.done = done;
    rps_stats stats{};
    const int rc = rps_sieve_batched(ctx, &request, deliver,
                                     done ? batch_done : nullptr, &state,
                                     &stats);
    if (state.error) {
        std::rethrow_exception(state.error);
    }
    if (rc != RPS_OK && rc != RPS_ECANCELLED) {
        raise_status(rc, rps_last_error(ctx));
    }
    if (cancelled) {
        *cancelled = rc == RPS_ECANCELLED;
    }
    SieveResult result;
    result.survivor_count = stats.survivor_count;
    SieveMetrics &m = result.metrics;
    m.basis_ms = static_cast<float>(stats.basis_ms);
    m.sieve_ms = static_cast<float>(stats.sieve_ms);
    m.plan_ms = stats.plan_ms;
    m.workspace_ms = stats.workspace_ms;
    m.basis_wall_ms = stats.basis_wall_ms;
    m.survivor_setup_ms = stats.survivor_setup_ms;
    m.verification_ms = stats.verification_ms;
    m.sieve_wall_ms = stats.sieve_wall_ms;
    m.word_count = stats.word_count;
    m.initial_mask_bytes = stats.initial_mask_bytes;
    return result;
}

SieveResult run_modular_sieve(const Coefficients &coefficients,
                              long long numerator_bound,
                              DenominatorRange denominators,
                              const CandidateBatchCallback &callback,
                              bool square_denominators) {
    return run_sieve_stream(coefficients, numerator_bound, denominators,
                            callback, nullptr, square_denominators, 1 << 16,
                            nullptr);
}

}  // namespace ratpoints_gpu
