// Search orchestration (the reference's include/point_search.hpp:15-78, same
// interface): denominators are sieved in batches of at most 65,536 on one
// GPU, survivors are verified exactly with GMP on a host thread pool while
// the GPU sieves the next batch, and points are reported in (b, a) order.
#ifndef RATPOINTS_GPU_POINT_SEARCH_HPP
#define RATPOINTS_GPU_POINT_SEARCH_HPP

#include <gmpxx.h>

#include <cstddef>
#include <functional>
#include <memory>
#include <vector>

#include "gpu_devices.hpp"
#include "search_types.hpp"

namespace ratpoints_gpu {

// Closed interval of x; an infinite end is unbounded.
struct SearchInterval {
    mpq_class lower;
    mpq_class upper;
    bool lower_infinite = false;
    bool upper_infinite = false;
};

struct SearchOptions {
    Coefficients coefficients;
    long long numerator_bound = 0;
    DenominatorRange denominators{1, 0};
    std::vector<SearchInterval> intervals;
    bool include_infinity = true;
    std::vector<int> devices;  // empty: the first visible GPU
    int denominator_batch_size = 1 << 16;
    bool square_denominators = false;  // -dl/-du bound k; denominator k^2

    // x = numerator / denominator lies in one of the intervals (or there
    // are none).
    bool accepts(long long numerator, int denominator) const;
};

struct DeviceSearchMetrics {
    GpuDevice device;
    unsigned long long denominator_count = 0;
    unsigned long long batches = 0;
    double basis_ms = 0.0;
    double sieve_ms = 0.0;
};

struct SearchMetrics {
    double wall_ms = 0.0;
    std::vector<DeviceSearchMetrics> devices;
    double basis_ms = 0.0;
    double sieve_ms = 0.0;
    double plan_ms = 0.0;
    double workspace_ms = 0.0;
    double basis_wall_ms = 0.0;
    double survivor_setup_ms = 0.0;
    double verification_ms = 0.0;
    double sieve_wall_ms = 0.0;
    unsigned long long word_count = 0;
    unsigned long long denominator_count = 0;
    unsigned long long modular_survivors = 0;
    long double initial_mask_bytes = 0.0L;
    size_t exact_survivors = 0;

    double initial_mask_gbs() const;
};

// Receives points in output order; return false to stop the search.
using PointCallback = std::function<bool(const PointPair &)>;

class ExactPolynomial;
struct CandidateBatch;

// Exact check of modular survivors (src/point_search.cpp:48-60): appends
// (a, sqrt F(a, b), b) for every survivor with gcd(a, b) = 1 and F(a, b) a
// perfect square; large batches are split over host threads.
void verify_candidates(const CandidateBatch &batch, const ExactPolynomial &polynomial,
                       std::vector<PointPair> &out);

// Output order: increasing (b, a).
void order_points(std::vector<PointPair> &points);

class PointSearch {
public:
    explicit PointSearch(SearchOptions options);
    ~PointSearch();

    PointSearch(const PointSearch &) = delete;
    PointSearch &operator=(const PointSearch &) = delete;

    SearchMetrics run(const PointCallback &point_callback);

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace ratpoints_gpu

#endif
