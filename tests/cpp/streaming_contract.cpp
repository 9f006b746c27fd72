// The streaming contract of run_modular_sieve, called through its C++
// signature (the reference's include/sieve.hpp:43-47 and its contract test
// tests/test_sieve_streaming.cpp:45-80): for f = 1 + P x with P the product
// of the 32 largest primes below 512 every site of the box survives, so the
// callbacks must deliver each (a, b) of [-A, A] x [1, B] exactly once and
// survivor_count must equal the number of sites.  The calling thread's
// context gets a survivor buffer of 16,384 entries, so the stream arrives in
// several flushes.
//
// Usage: streaming_contract A B   (prints "ok <sites> <flushes>")
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include <gmpxx.h>

#include "gpu_devices.hpp"
#include "ratpoints_sieve.h"
#include "sieve.hpp"
#include "sieve_shim.hpp"

namespace {

// 1 + P x: P = product of the 32 largest primes below 512.
ratpoints_gpu::Coefficients every_site_survives() {
    std::vector<int> primes;
    for (int n = 2; n < 512; ++n) {
        bool prime = true;
        for (int d = 2; d * d <= n && prime; ++d) {
            prime = n % d != 0;
        }
        if (prime) {
            primes.push_back(n);
        }
    }
    mpz_class product = 1;
    for (size_t i = primes.size() - 32; i < primes.size(); ++i) {
        product *= primes[i];
    }
    return {"1", product.get_str()};
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: streaming_contract A B\n");
        return 2;
    }
    const long long bound = std::atoll(argv[1]);
    const int last = std::atoi(argv[2]);
    try {
        ratpoints_gpu::activate_gpu(0);
        rps_ctx *ctx = ratpoints_gpu::thread_sieve_context();
        if (rps_set_option(ctx, "survivor_capacity", 16384) != RPS_OK) {
            throw std::runtime_error(rps_last_error(ctx));
        }
        const size_t width = static_cast<size_t>(2 * bound + 1);
        const size_t sites = width * static_cast<size_t>(last);
        std::vector<unsigned char> seen(sites, 0);
        size_t delivered = 0, flushes = 0;
        const ratpoints_gpu::SieveResult result = ratpoints_gpu::run_modular_sieve(
            every_site_survives(), bound, {1, last},
            [&](const ratpoints_gpu::CandidateBatch &batch) {
                if (batch.denominators.size() != batch.size()) {
                    throw std::runtime_error("numerator and denominator counts differ");
                }
                ++flushes;
                for (size_t i = 0; i < batch.size(); ++i) {
                    const long long a = batch.numerators[i];
                    const int b = batch.denominators[i];
                    if (a < -bound || a > bound || b < 1 || b > last) {
                        throw std::runtime_error("survivor outside the box");
                    }
                    unsigned char &mark = seen[static_cast<size_t>(b - 1) * width
                                               + static_cast<size_t>(a + bound)];
                    if (mark) {
                        throw std::runtime_error("survivor delivered twice");
                    }
                    mark = 1;
                    ++delivered;
                }
            });
        if (delivered != sites || result.survivor_count != sites) {
            throw std::runtime_error("incomplete survivor stream: " + std::to_string(delivered)
                                     + " of " + std::to_string(sites));
        }
        std::printf("ok %zu %zu\n", sites, flushes);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "streaming_contract: %s\n", e.what());
        return 1;
    }
}
