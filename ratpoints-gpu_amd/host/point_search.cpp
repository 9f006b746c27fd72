// Search orchestration on one GPU.
//
// Semantics follow src/point_search.cpp of the reference: options are
// validated (batch size, squarefree f), the point at infinity comes first,
// then points in increasing (b, a) order, filtered by the closed intervals,
// until the callback asks to stop.  The schedule is this build's own: the
// whole denominator range is sieved in one streaming call, so the GPU sieves
// batch i+1 while this thread (plus a pool of helpers for large flushes)
// verifies batch i exactly with GMP; a batch's points are emitted once the
// sieve reports the batch complete.
#include "point_search.hpp"

#include <algorithm>
#include <chrono>
#include <future>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

#include "exact_polynomial.hpp"
#include "sieve.hpp"
#include "sieve_shim.hpp"

namespace ratpoints_gpu {
namespace {

constexpr size_t kParallelVerifyThreshold = 1 << 15;

SearchOptions checked(SearchOptions options) {
    if (options.denominator_batch_size < 1
        || options.denominator_batch_size > (1 << 16)) {
        throw std::invalid_argument(
            "denominator batch size must be in 1..65536");
    }
    if (!ExactPolynomial::is_squarefree(options.coefficients)) {
        throw std::invalid_argument("polynomial is not squarefree");
    }
    return options;
}

long long gcd_with(long long a, long long b) {
    unsigned long long x = a < 0 ? 0ULL - static_cast<unsigned long long>(a)
                                 : static_cast<unsigned long long>(a);
    unsigned long long y = static_cast<unsigned long long>(b);
    while (y) {
        unsigned long long r = x % y;
        x = y;
        y = r;
    }
    return static_cast<long long>(x);
}

// gcd(a, b) = 1 and F(a, b) a perfect square (src/point_search.cpp:48-60).
void verify_range(const CandidateBatch &batch, size_t begin, size_t end,
                  const ExactPolynomial &polynomial,
                  std::vector<PointPair> &out) {
    mpz_class root;
    for (size_t i = begin; i < end; ++i) {
        const long long a = batch.numerators[i];
        const int b = batch.denominators[i];
        if (gcd_with(a, b) == 1 && polynomial.square_root(a, b, root)) {
            out.push_back({a, root.get_str(), b});
        }
    }
}

}  // namespace

void verify_candidates(const CandidateBatch &batch, const ExactPolynomial &polynomial,
                       std::vector<PointPair> &out) {
    const size_t n = batch.size();
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t workers = std::min<size_t>(std::min(hw, 16u), n / 4096);
    if (n < kParallelVerifyThreshold || workers < 2) {
        verify_range(batch, 0, n, polynomial, out);
        return;
    }
    std::vector<std::vector<PointPair>> parts(workers);
    std::vector<std::future<void>> jobs;
    const size_t step = (n + workers - 1) / workers;
    for (size_t w = 1; w < workers; ++w) {
        const size_t begin = std::min(n, w * step);
        const size_t end = std::min(n, begin + step);
        jobs.push_back(std::async(std::launch::async, [&, w, begin, end] {
            verify_range(batch, begin, end, polynomial, parts[w]);
        }));
    }
    verify_range(batch, 0, std::min(n, step), polynomial, parts[0]);
    for (auto &job : jobs) {
        job.get();
    }
    for (auto &part : parts) {
        out.insert(out.end(), part.begin(), part.end());
    }
}

void order_points(std::vector<PointPair> &points) {
    std::sort(points.begin(), points.end(),
              [](const PointPair &x, const PointPair &y) {
                  return x.denominator != y.denominator
                      ? x.denominator < y.denominator
                      : x.numerator < y.numerator;
              });
}

namespace {

void add_count(unsigned long long &total, unsigned long long value,
               const char *what) {
    if (value > std::numeric_limits<unsigned long long>::max() - total) {
        throw std::overflow_error(what);
    }
    total += value;
}

}  // namespace

bool SearchOptions::accepts(long long numerator, int denominator) const {
    if (intervals.empty()) {
        return true;
    }
    mpq_class x(mpz_class(std::to_string(numerator), 10), mpz_class(denominator));
    x.canonicalize();
    for (const SearchInterval &interval : intervals) {
        if ((interval.lower_infinite || x >= interval.lower)
            && (interval.upper_infinite || x <= interval.upper)) {
            return true;
        }
    }
    return false;
}

double SearchMetrics::initial_mask_gbs() const {
    return sieve_ms == 0.0
        ? 0.0
        : static_cast<double>(initial_mask_bytes / (sieve_ms * 1.0e6L));
}

struct PointSearch::Impl {
    explicit Impl(SearchOptions search_options)
        : options(checked(std::move(search_options))),
          polynomial(options.coefficients),
          device(select_gpu_devices(options.devices).front()) {}

    SearchOptions options;
    ExactPolynomial polynomial;
    GpuDevice device;
};

PointSearch::PointSearch(SearchOptions options)
    : impl_(new Impl(std::move(options))) {}

PointSearch::~PointSearch() = default;

SearchMetrics PointSearch::run(const PointCallback &point_callback) {
    const auto started = std::chrono::steady_clock::now();
    const SearchOptions &opt = impl_->options;
    SearchMetrics metrics;
    metrics.devices.push_back(DeviceSearchMetrics{});
    metrics.devices.front().device = impl_->device;
    auto finish = [&] {
        metrics.wall_ms = std::chrono::duration<double, std::milli>(
            std::chrono::steady_clock::now() - started).count();
        return metrics;
    };

    mpz_class root;
    if (opt.include_infinity && impl_->polynomial.square_root(1, 0, root)
        && !point_callback({1, root.get_str(), 0})) {
        return finish();
    }
    if (opt.denominators.first > opt.denominators.last) {
        return finish();
    }

    activate_gpu(impl_->device.id);
    const unsigned long long words =
        (2ULL * static_cast<unsigned long long>(opt.numerator_bound) + 32) / 32;
    std::vector<PointPair> pending;
    DeviceSearchMetrics &dev = metrics.devices.front();
    bool stopped = false;

    CandidateBatchCallback on_survivors = [&](const CandidateBatch &batch) {
        verify_candidates(batch, impl_->polynomial, pending);
    };
    BatchDoneCallback on_batch = [&](DenominatorRange range,
                                     std::uint64_t survivors) {
        const unsigned long long count =
            static_cast<unsigned long long>(range.count());
        dev.denominator_count += count;
        ++dev.batches;
        metrics.denominator_count += count;
        metrics.initial_mask_bytes += static_cast<long double>(count) * words
            * 14 * sizeof(uint32_t);
        add_count(metrics.modular_survivors, survivors,
                  "total survivor count overflow");
        order_points(pending);
        metrics.exact_survivors += pending.size();
        for (const PointPair &point : pending) {
            if (opt.accepts(point.numerator, point.denominator)
                && !point_callback(point)) {
                stopped = true;
                break;
            }
        }
        pending.clear();
        return !stopped;
    };

    SieveResult sieve;
    try {
        sieve = run_sieve_stream(opt.coefficients, opt.numerator_bound,
                                 opt.denominators, on_survivors, &on_batch,
                                 opt.square_denominators,
                                 opt.denominator_batch_size, nullptr);
    } catch (const std::exception &error) {
        throw std::runtime_error("GPU " + std::to_string(impl_->device.id)
            + ", denominators " + std::to_string(opt.denominators.first)
            + ".." + std::to_string(opt.denominators.last) + ": "
            + error.what());
    }
    const SieveMetrics &m = sieve.metrics;
    dev.basis_ms += m.basis_ms;
    dev.sieve_ms += m.sieve_ms;
    metrics.basis_ms = m.basis_ms;
    metrics.sieve_ms = m.sieve_ms;
    metrics.plan_ms = m.plan_ms;
    metrics.workspace_ms = m.workspace_ms;
    metrics.basis_wall_ms = m.basis_wall_ms;
    metrics.survivor_setup_ms = m.survivor_setup_ms;
    metrics.verification_ms = m.verification_ms;
    metrics.sieve_wall_ms = m.sieve_wall_ms;
    metrics.word_count = m.word_count;
    return finish();
}

}  // namespace ratpoints_gpu
