// Serving loop over rps_sieve_many (see serve.hpp for the contract).
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

#include "exact_polynomial.hpp"
#include "ratpoints_sieve.h"
#include "sieve.hpp"

namespace ratpoints_gpu {
namespace {

using Clock = std::chrono::steady_clock;

constexpr double kTableBytes = 36.0 * 512 * 9 * 8;  // residue table per curve
constexpr double kSurvivorBytes = 16.0;              // num + den + request index
constexpr double kSurvivorRate = 1.0 / (1ULL << 30); // generous survivors per site

double line_sites(const ServeRequest &r) {
    return 2.0 * static_cast<double>(r.options.numerator_bound) + 1.0;
}

// One admitted request.
struct Active {
    ServeRequest request;
    ExactPolynomial polynomial;
    std::vector<const char *> coefficient_ptrs;
    long long next_t = 0;
    long long last_t = 0;
    long long chunk_last = 0;
    double footprint = 0.0;
    ServeResult result;
    std::vector<long long> num;
    std::vector<int> den;

    explicit Active(ServeRequest r)
        : request(std::move(r)), polynomial(request.options.coefficients) {
        for (const std::string &c : request.options.coefficients) {
            coefficient_ptrs.push_back(c.c_str());
        }
    }
};

rps_request to_request(const Active &a, long long first, long long last) {
    rps_request q{};
    q.coefficients = a.coefficient_ptrs.data();
    q.n_coefficients = static_cast<int32_t>(a.coefficient_ptrs.size());
    q.numerator_bound = a.request.options.numerator_bound;
    q.den_first = static_cast<int32_t>(first);
    q.den_last = static_cast<int32_t>(last);
    q.square_denominators = a.request.options.square_denominators ? 1 : 0;
    return q;
}

// Points of one verified chunk, in output order and filtered; returns
// whether the request is complete (-1 satisfied).
bool take_points(Active &a, std::vector<PointPair> &found) {
    order_points(found);
    for (PointPair &p : found) {
        if (a.request.options.accepts(p.numerator, p.denominator)) {
            a.result.points.push_back(std::move(p));
            if (a.request.first_point) {
                return true;
            }
        }
    }
    return false;
}

}  // namespace

double request_footprint(const ServeRequest &request, const ServeOptions &options) {
    const double chunk_sites = std::min(
        options.round_sites,
        line_sites(request) * std::max(1, request.options.denominators.count()));
    return kTableBytes + kSurvivorBytes * std::max(1024.0, chunk_sites * kSurvivorRate);
}

struct Server::Impl {
    ServeOptions options;
    ServeCallback on_done;
    Clock::time_point start = Clock::now();

    mutable std::mutex mutex;
    std::condition_variable wake;
    std::condition_variable idle;
    std::deque<std::pair<ServeRequest, double>> pending;
    size_t outstanding = 0;
    bool stopping = false;
    ServeStats totals;

    rps_ctx *ctx = nullptr;
    std::vector<std::unique_ptr<Active>> active;
    double used = 0.0;
    std::string open_error;
    std::thread worker;

    double now() const {
        return std::chrono::duration<double, std::milli>(Clock::now() - start).count();
    }

    void finish(std::unique_ptr<Active> &a) {
        a->result.finish_ms = now();
        used -= a->footprint;
        on_done(a->result);
        a.reset();
        std::lock_guard<std::mutex> lock(mutex);
        --outstanding;
        if (outstanding == 0) {
            idle.notify_all();
        }
    }

    void fail_now(ServeRequest &&request, double submit_ms, const std::string &error) {
        ServeResult r;
        r.id = request.id;
        r.error = error;
        r.submit_ms = submit_ms;
        r.admit_ms = r.finish_ms = now();
        on_done(r);
        std::lock_guard<std::mutex> lock(mutex);
        --outstanding;
        if (outstanding == 0) {
            idle.notify_all();
        }
    }

    // Moves pending requests onto the device while a slot is free and their
    // footprint fits (the first always fits an empty device).
    void admit() {
        for (;;) {
            std::pair<ServeRequest, double> next;
            {
                std::lock_guard<std::mutex> lock(mutex);
                if (pending.empty()
                    || static_cast<int>(active.size()) >= options.max_active) {
                    return;
                }
                const double fp = request_footprint(pending.front().first, options);
                if (!active.empty() && used + fp > options.memory_budget) {
                    return;
                }
                next = std::move(pending.front());
                pending.pop_front();
            }
            ServeRequest &request = next.first;
            const SearchOptions &opt = request.options;
            if (!open_error.empty()) {
                fail_now(std::move(request), next.second, open_error);
                continue;
            }
            std::vector<const char *> ptrs;
            for (const std::string &c : opt.coefficients) {
                ptrs.push_back(c.c_str());
            }
            rps_request q{};
            q.coefficients = ptrs.data();
            q.n_coefficients = static_cast<int32_t>(ptrs.size());
            q.numerator_bound = opt.numerator_bound;
            q.den_first = opt.denominators.first;
            q.den_last = std::max(opt.denominators.first, opt.denominators.last);
            q.square_denominators = opt.square_denominators ? 1 : 0;
            char message[256];
            if (rps_check_request(&q, message, sizeof message) != RPS_OK) {
                fail_now(std::move(request), next.second, message);
                continue;
            }
            std::unique_ptr<Active> a;
            try {
                if (!ExactPolynomial::is_squarefree(opt.coefficients)) {
                    throw std::invalid_argument("polynomial is not squarefree");
                }
                // A copy: `request` must stay whole for fail_now if this throws.
                a.reset(new Active(request));
            } catch (const std::exception &e) {
                fail_now(std::move(request), next.second, e.what());
                continue;
            }
            a->result.id = a->request.id;
            a->result.submit_ms = next.second;
            a->result.admit_ms = now();
            a->next_t = a->request.options.denominators.first;
            a->last_t = a->request.options.denominators.last;
            a->footprint = request_footprint(a->request, options);
            used += a->footprint;
            ++totals.admitted;
            // The point at infinity comes first (src/point_search.cpp:143-147).
            mpz_class root;
            bool done = a->next_t > a->last_t;
            if (a->request.options.include_infinity && a->polynomial.square_root(1, 0, root)) {
                a->result.points.push_back({1, root.get_str(), 0});
                done = done || a->request.first_point;
            }
            active.push_back(std::move(a));
            if (done) {
                finish(active.back());
                active.pop_back();
            }
        }
    }

    // One launch over a chunk of every active request's denominators.
    void round() {
        const size_t n = active.size();
        // Equal shares: a newcomer's first chunk waits at most one round,
        // and small requests finish in the round they join.
        const double share = options.round_sites / static_cast<double>(n);
        std::vector<rps_request> reqs(n);
        double sites = 0.0;
        for (size_t i = 0; i < n; ++i) {
            Active &a = *active[i];
            const double per_line = line_sites(a.request);
            const long long left = a.last_t - a.next_t + 1;
            const long long want = static_cast<long long>(std::floor(share / per_line));
            const long long k = std::max(1LL, std::min(left, want));
            a.chunk_last = a.next_t + k - 1;
            reqs[i] = to_request(a, a.next_t, a.chunk_last);
            a.num.clear();
            a.den.clear();
            sites += per_line * static_cast<double>(k);
        }
        rps_stats st{};
        auto sink = [](void *user, const int32_t *req, const int64_t *num,
                       const int32_t *den, size_t count) -> int {
            auto &act = *static_cast<std::vector<std::unique_ptr<Active>> *>(user);
            for (size_t i = 0; i < count; ++i) {
                Active &a = *act[static_cast<size_t>(req[i])];
                a.num.push_back(num[i]);
                a.den.push_back(den[i]);
            }
            return 0;
        };
        const int rc = rps_sieve_many(ctx, reqs.data(), n, sink, &active, &st);
        ++totals.rounds;
        totals.sites += sites;
        totals.kernel_ms += st.sieve_ms;
        totals.peak_active = std::max<unsigned long long>(totals.peak_active, n);
        if (rc != RPS_OK) {
            const std::string error = std::string("GPU round failed: ") + rps_last_error(ctx);
            for (auto &a : active) {
                a->result.error = error;
                finish(a);
            }
            active.clear();
            return;
        }
        for (auto &a : active) {
            CandidateBatch batch;
            batch.numerators = std::move(a->num);
            batch.denominators = std::move(a->den);
            std::vector<PointPair> found;
            verify_candidates(batch, a->polynomial, found);
            a->result.sites += line_sites(a->request)
                * static_cast<double>(a->chunk_last - a->next_t + 1);
            ++a->result.rounds;
            a->next_t = a->chunk_last + 1;
            const bool satisfied = take_points(*a, found);
            if (satisfied || a->next_t > a->last_t) {
                finish(a);
            }
        }
        active.erase(std::remove(active.begin(), active.end(), nullptr), active.end());
    }

    void run() {
        if (rps_open(options.device, &ctx) != RPS_OK) {
            open_error = std::string("cannot open GPU ") + std::to_string(options.device)
                + ": " + rps_last_error(nullptr);
            ctx = nullptr;
        } else {
            rps_set_option(ctx, "table_slots", std::max(1, options.max_active));
            // Non-primitive (a, b) are never points: drop them on the device.
            rps_set_option(ctx, "primitive_only", 1);
        }
        for (;;) {
            {
                std::unique_lock<std::mutex> lock(mutex);
                wake.wait(lock, [&] { return stopping || !pending.empty() || !active.empty(); });
                if (stopping && pending.empty() && active.empty()) {
                    break;
                }
            }
            admit();
            if (!active.empty()) {
                round();
            }
        }
        rps_close(ctx);
        ctx = nullptr;
    }
};

Server::Server(ServeOptions options, ServeCallback on_done) : impl_(new Impl) {
    impl_->options = options;
    impl_->on_done = std::move(on_done);
    impl_->worker = std::thread([this] { impl_->run(); });
}

Server::~Server() {
    drain();
    {
        std::lock_guard<std::mutex> lock(impl_->mutex);
        impl_->stopping = true;
    }
    impl_->wake.notify_all();
    impl_->worker.join();
}

void Server::submit(ServeRequest request) {
    {
        std::lock_guard<std::mutex> lock(impl_->mutex);
        impl_->pending.emplace_back(std::move(request), impl_->now());
        ++impl_->outstanding;
    }
    impl_->wake.notify_all();
}

void Server::drain() {
    std::unique_lock<std::mutex> lock(impl_->mutex);
    impl_->idle.wait(lock, [&] { return impl_->outstanding == 0; });
}

double Server::now_ms() const { return impl_->now(); }

ServeStats Server::stats() const {
    std::lock_guard<std::mutex> lock(impl_->mutex);
    return impl_->totals;
}

ServeResult search_alone(const ServeRequest &request) {
    ServeResult r;
    r.id = request.id;
    try {
        PointSearch search(request.options);
        const bool first = request.first_point;
        const SearchMetrics m = search.run([&](const PointPair &p) {
            r.points.push_back(p);
            return !first;
        });
        r.sites = line_sites(request) * static_cast<double>(m.denominator_count);
        r.rounds = static_cast<int>(m.devices.empty() ? 0 : m.devices.front().batches);
    } catch (const std::exception &e) {
        r.points.clear();
        r.error = e.what();
    }
    return r;
}

}  // namespace ratpoints_gpu
