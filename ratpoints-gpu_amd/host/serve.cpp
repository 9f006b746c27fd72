// Serving loop over rps_sieve_many (see serve.hpp for the contract).
//
// Three kinds of work run at once:
//   * the scheduling thread owns the device context: it admits checked
//     requests while slots and the memory budget allow, sieves one round (a
//     chunk of every active request's denominators in one rps_sieve_many
//     call) and hands each request's survivors to the worker pool;
//   * the worker pool verifies those survivors with GMP while the next
//     round is already on the GPU (a chunk of a few survivors is verified
//     on the scheduling thread, cheaper than the handoff), and runs the
//     admission checks (request limits, the squarefree test, the point at
//     infinity) of new requests with large coefficients, so a slow check
//     never stalls the GPU; small ones are checked on submit()'s thread;
//   * the scheduling thread applies verified chunks in denominator order per
//     request and reports finished requests, so results are exactly those of
//     the same search alone.
#include "serve.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <functional>
#include <map>
#include <memory>
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
constexpr int kMaxChunksInFlight = 2;                // sieved, not yet applied
constexpr size_t kInlineVerify = 32;                 // survivors verified on the scheduling thread
constexpr int kSpinMicros = 2000000;                 // scheduling thread's poll before it sleeps
constexpr size_t kInlineCheckChars = 4096;           // coefficient text checked on submit()'s thread

double line_sites(const ServeRequest &r) {
    return 2.0 * static_cast<double>(r.options.numerator_bound) + 1.0;
}

// A fixed set of threads running posted jobs in order.
class WorkerPool {
public:
    explicit WorkerPool(unsigned n) {
        for (unsigned i = 0; i < n; ++i) {
            threads_.emplace_back([this] { loop(); });
        }
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : threads_) {
            t.join();
        }
    }
    void post(std::function<void()> job) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            jobs_.push_back(std::move(job));
        }
        cv_.notify_one();
    }

private:
    void loop() {
        for (;;) {
            std::function<void()> job;
            {
                std::unique_lock<std::mutex> lock(mutex_);
                cv_.wait(lock, [&] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) {
                    return;
                }
                job = std::move(jobs_.front());
                jobs_.pop_front();
            }
            job();
        }
    }

    std::vector<std::thread> threads_;
    std::mutex mutex_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> jobs_;
    bool stop_ = false;
};

// One admitted request.
struct Active {
    ServeRequest request;
    ExactPolynomial polynomial;
    std::vector<const char *> coefficient_ptrs;
    double submit_ms = 0.0;
    long long next_t = 0;       // first denominator parameter not yet scheduled
    long long last_t = 0;
    double footprint = 0.0;
    ServeResult result;
    int next_seq = 0;           // sequence number of the next chunk to schedule
    int apply_seq = 0;          // sequence number of the next chunk to apply
    int in_flight = 0;          // chunks scheduled and not yet applied
    bool complete = false;      // -1 satisfied: nothing more is scheduled
    bool finished = false;      // reported
    std::map<int, std::pair<std::vector<PointPair>, double>> verified;  // by sequence

    Active(ServeRequest r, double submitted)
        : request(std::move(r)), polynomial(request.options.coefficients), submit_ms(submitted) {
        for (const std::string &c : request.options.coefficients) {
            coefficient_ptrs.push_back(c.c_str());
        }
        next_t = request.options.denominators.first;
        last_t = request.options.denominators.last;
    }

    bool schedulable() const {
        return !complete && next_t <= last_t && in_flight < kMaxChunksInFlight;
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

// The admission checks of one request: the C ABI's limits, a squarefree f
// and the point at infinity (src/point_search.cpp:185-189); throws with the
// message the search alone would give.
std::shared_ptr<Active> checked_request(ServeRequest request, double submit_ms) {
    const SearchOptions &opt = request.options;
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
        throw std::invalid_argument(message);
    }
    if (!ExactPolynomial::is_squarefree(opt.coefficients)) {
        throw std::invalid_argument("polynomial is not squarefree");
    }
    auto a = std::make_shared<Active>(std::move(request), submit_ms);
    a->result.id = a->request.id;
    a->result.submit_ms = submit_ms;
    mpz_class root;
    if (a->request.options.include_infinity && a->polynomial.square_root(1, 0, root)) {
        a->result.points.push_back({1, root.get_str(), 0});
        a->complete = a->request.first_point;
    }
    return a;
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

    // Shared with submit(), the workers and the scheduling thread.
    mutable std::mutex mutex;
    std::condition_variable wake;   // the scheduling thread has work
    std::atomic<unsigned long long> events{0};  // bumped with every wake
    std::condition_variable idle;   // every submitted request finished
    size_t outstanding = 0;
    bool stopping = false;
    ServeStats totals;
    size_t checking = 0;                                   // admission checks running
    std::deque<std::shared_ptr<Active>> ready;             // checked, waiting for a slot
    std::deque<ServeResult> refused;                       // failed their checks
    struct Done {
        std::shared_ptr<Active> active;
        int seq;
        std::vector<PointPair> points;
        double sites;
    };
    std::deque<Done> verified;

    // Scheduling thread only.
    rps_ctx *ctx = nullptr;
    double last_round_end = 0.0;
    std::string open_error;
    std::vector<std::shared_ptr<Active>> active;
    double used = 0.0;

    std::unique_ptr<WorkerPool> pool;
    std::thread scheduler;

    double now() const {
        return std::chrono::duration<double, std::milli>(Clock::now() - start).count();
    }

    void signal() {
        events.fetch_add(1, std::memory_order_release);
        wake.notify_all();
    }

    // Before the scheduling thread sleeps it polls (yielding the core) for
    // up to kSpinMicros: under a steady stream of requests it never sleeps,
    // so a request that arrives at an idle GPU starts without a thread
    // wake-up (tens of microseconds), at the cost of one busy core.
    void spin_for_work(unsigned long long seen) {
        const auto until = Clock::now() + std::chrono::microseconds(kSpinMicros);
        while (events.load(std::memory_order_acquire) == seen && Clock::now() < until) {
            std::this_thread::yield();
        }
    }

    void report(ServeResult &r) {
        r.finish_ms = now();
        on_done(r);
        std::lock_guard<std::mutex> lock(mutex);
        if (--outstanding == 0) {
            idle.notify_all();
        }
    }

    void finish(const std::shared_ptr<Active> &a) {
        if (a->finished) {
            return;
        }
        a->finished = true;
        used -= a->footprint;
        report(a->result);
    }

    // Worker side of admission.
    void check(ServeRequest request, double submit_ms) {
        std::shared_ptr<Active> a;
        ServeResult failed;
        try {
            a = checked_request(request, submit_ms);
        } catch (const std::exception &e) {
            failed.id = request.id;
            failed.error = e.what();
            failed.submit_ms = submit_ms;
        }
        {
            std::lock_guard<std::mutex> lock(mutex);
            --checking;
            if (a) {
                ready.push_back(std::move(a));
            } else {
                refused.push_back(std::move(failed));
            }
        }
        signal();
    }

    // Worker side of verification: one chunk of one request.
    void verify(std::shared_ptr<Active> a, int seq, CandidateBatch batch, double sites) {
        std::vector<PointPair> found;
        verify_candidates(batch, a->polynomial, found);
        order_points(found);
        std::vector<PointPair> kept;
        for (PointPair &p : found) {
            if (a->request.options.accepts(p.numerator, p.denominator)) {
                kept.push_back(std::move(p));
            }
        }
        {
            std::lock_guard<std::mutex> lock(mutex);
            verified.push_back({std::move(a), seq, std::move(kept), sites});
        }
        signal();
    }

    // Moves checked requests onto the device while a slot is free and their
    // footprint fits (the first always fits an empty device), in the order
    // their checks finished.
    void admit(std::deque<std::shared_ptr<Active>> &incoming) {
        while (!incoming.empty() && static_cast<int>(active.size()) < options.max_active) {
            std::shared_ptr<Active> a = incoming.front();
            const double fp = request_footprint(a->request, options);
            if (!active.empty() && used + fp > options.memory_budget) {
                break;
            }
            incoming.pop_front();
            a->result.admit_ms = now();
            a->footprint = fp;
            used += fp;
            {
                std::lock_guard<std::mutex> lock(mutex);
                ++totals.admitted;
            }
            if (!open_error.empty()) {
                a->result.points.clear();
                a->result.error = open_error;
                a->finished = true;
                used -= fp;
                report(a->result);
                continue;
            }
            active.push_back(a);
            if (a->complete || a->next_t > a->last_t) {
                finish(a);
            }
        }
    }

    // Verified chunks in denominator order per request; a -1 request ends
    // at its first point and its later chunks are dropped.
    void apply(std::deque<Done> &done) {
        for (Done &d : done) {
            Active &a = *d.active;
            a.verified.emplace(d.seq, std::make_pair(std::move(d.points), d.sites));
            for (auto it = a.verified.find(a.apply_seq); it != a.verified.end();
                 it = a.verified.find(a.apply_seq)) {
                if (!a.complete) {
                    a.result.sites += it->second.second;
                    ++a.result.rounds;
                    for (PointPair &p : it->second.first) {
                        a.result.points.push_back(std::move(p));
                        if (a.request.first_point) {
                            a.complete = true;
                            break;
                        }
                    }
                }
                a.verified.erase(it);
                ++a.apply_seq;
                --a.in_flight;
            }
            if (a.complete || (a.next_t > a.last_t && a.in_flight == 0)) {
                finish(d.active);
            }
        }
        prune();
    }

    // Frees the slots of reported requests with no chunk left in flight.
    void prune() {
        active.erase(std::remove_if(active.begin(), active.end(),
                                    [](const std::shared_ptr<Active> &a) {
                                        return a->finished && a->in_flight == 0;
                                    }),
                     active.end());
    }

    // One launch over a chunk of every schedulable request.  Requests take
    // the round's sites in order of least remaining work (shortest remaining
    // first): short and newly arrived requests finish in the round they
    // join, and of two long requests the one closer to its end keeps the
    // GPU, so neither waits longer than it would behind the other one at a
    // time.  Every scheduled request gets at least one denominator.
    void round() {
        std::vector<std::shared_ptr<Active>> take;
        for (const auto &a : active) {
            if (a->schedulable()) {
                take.push_back(a);
            }
        }
        if (take.empty()) {
            return;
        }
        auto remaining = [](const Active &a) {
            return line_sites(a.request) * static_cast<double>(a.last_t - a.next_t + 1);
        };
        std::sort(take.begin(), take.end(), [&](const auto &x, const auto &y) {
            return remaining(*x) < remaining(*y);
        });
        const size_t n = take.size();
        std::vector<rps_request> reqs(n);
        std::vector<long long> first(n), last(n);
        double budget = options.round_sites;
        double sites = 0.0;
        for (size_t i = 0; i < n; ++i) {
            Active &a = *take[i];
            const double per_line = line_sites(a.request);
            const long long left = a.last_t - a.next_t + 1;
            const long long want = static_cast<long long>(
                std::min(std::floor(std::max(budget, 0.0) / per_line), 9.0e18));
            const long long k = std::max(1LL, std::min(left, want));
            first[i] = a.next_t;
            last[i] = a.next_t + k - 1;
            reqs[i] = to_request(a, first[i], last[i]);
            budget -= per_line * static_cast<double>(k);
            sites += per_line * static_cast<double>(k);
        }
        std::vector<CandidateBatch> found(n);
        auto sink = [](void *user, const int32_t *req, const int64_t *num, const int32_t *den,
                       size_t count) -> int {
            auto &out = *static_cast<std::vector<CandidateBatch> *>(user);
            for (size_t i = 0; i < count; ++i) {
                CandidateBatch &b = out[static_cast<size_t>(req[i])];
                b.numerators.push_back(num[i]);
                b.denominators.push_back(den[i]);
            }
            return 0;
        };
        const double t0 = now();
        double first_wait = 0.0;  // longest submit-to-launch wait of a request's first chunk
        for (const auto &a : take) {
            if (a->next_seq == 0) {
                first_wait = std::max(first_wait, t0 - a->submit_ms);
            }
        }
        rps_stats st{};
        const int rc = rps_sieve_many(ctx, reqs.data(), n, sink, &found, &st);
        const double wall = now() - t0;
        unsigned long long round_no = 0;
        {
            std::lock_guard<std::mutex> lock(mutex);
            round_no = ++totals.rounds;
            totals.sites += sites;
            totals.kernel_ms += st.sieve_ms;
            totals.peak_active = std::max<unsigned long long>(totals.peak_active, active.size());
        }
        if (options.progress) {
            std::fprintf(stderr, "serve: round %llu: %zu requests, %.4g sites, kernel %.2f ms, "
                         "call %.2f ms (plan %.3f, tables %.3f, sieve %.3f), %zu active, "
                         "idle before %.3f ms, new request waited %.3f ms\n",
                         round_no, n, sites, st.sieve_ms, wall, st.plan_ms + st.workspace_ms,
                         st.basis_wall_ms, st.sieve_wall_ms, active.size(), t0 - last_round_end,
                         first_wait);
        }
        last_round_end = now();
        if (rc != RPS_OK) {
            const std::string error = std::string("GPU round failed: ") + rps_last_error(ctx);
            for (auto &a : active) {
                a->result.points.clear();
                a->result.error = error;
                a->complete = true;
                finish(a);
            }
            return;
        }
        for (size_t i = 0; i < n; ++i) {
            std::shared_ptr<Active> a = take[i];
            const int seq = a->next_seq++;
            ++a->in_flight;
            a->next_t = last[i] + 1;
            const double chunk_sites = line_sites(a->request)
                * static_cast<double>(last[i] - first[i] + 1);
            if (found[i].numerators.size() <= kInlineVerify) {
                // A handful of survivors: verified at once, which is quicker
                // than handing them to a worker.
                verify(std::move(a), seq, std::move(found[i]), chunk_sites);
                continue;
            }
            pool->post([this, a, seq, batch = std::move(found[i]), chunk_sites]() mutable {
                verify(std::move(a), seq, std::move(batch), chunk_sites);
            });
        }
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
            rps_set_option(ctx, "launch_timeout_ms", std::max(1LL, options.launch_timeout_ms));
        }
        std::deque<std::shared_ptr<Active>> incoming;
        unsigned long long seen = ~0ULL;
        for (;;) {
            std::deque<ServeResult> failed;
            std::deque<Done> done;
            const bool runnable =
                (!incoming.empty() && static_cast<int>(active.size()) < options.max_active)
                || std::any_of(active.begin(), active.end(),
                               [](const std::shared_ptr<Active> &a) { return a->schedulable(); });
            if (!runnable) {
                spin_for_work(seen);
            }
            {
                std::unique_lock<std::mutex> lock(mutex);
                seen = events.load(std::memory_order_acquire);
                auto busy = [&] {
                    for (const auto &a : active) {
                        if (a->schedulable()) {
                            return true;
                        }
                    }
                    return false;
                };
                wake.wait(lock, [&] {
                    return !ready.empty() || !refused.empty() || !verified.empty()
                        || (!incoming.empty()
                            && static_cast<int>(active.size()) < options.max_active)
                        || busy() || (stopping && outstanding == 0 && active.empty());
                });
                if (stopping && outstanding == 0 && active.empty()) {
                    break;
                }
                for (auto &a : ready) {
                    incoming.push_back(std::move(a));
                }
                ready.clear();
                failed.swap(refused);
                done.swap(verified);
            }
            for (ServeResult &r : failed) {
                r.admit_ms = now();
                report(r);
            }
            apply(done);
            // Requests that end at admission (-1 with a point at infinity,
            // an empty range) free their slot at once, so admission can
            // continue in the same pass.
            for (size_t before = incoming.size() + 1; incoming.size() < before;) {
                before = incoming.size();
                admit(incoming);
                prune();
            }
            round();
        }
        rps_close(ctx);
        ctx = nullptr;
    }
};

Server::Server(ServeOptions options, ServeCallback on_done) : impl_(new Impl) {
    impl_->options = options;
    impl_->on_done = std::move(on_done);
    const unsigned hw = std::max(2u, std::min(16u, std::thread::hardware_concurrency()));
    impl_->pool.reset(new WorkerPool(hw));
    impl_->scheduler = std::thread([this] { impl_->run(); });
}

Server::~Server() {
    drain();
    {
        std::lock_guard<std::mutex> lock(impl_->mutex);
        impl_->stopping = true;
    }
    impl_->signal();
    impl_->scheduler.join();
    impl_->pool.reset();
}

void Server::submit(ServeRequest request) {
    const double at = impl_->now();
    {
        std::lock_guard<std::mutex> lock(impl_->mutex);
        ++impl_->outstanding;
        ++impl_->checking;
    }
    // Small polynomials are checked on the caller's thread (microseconds,
    // and no handoff); large ones, whose squarefree test can take longer, on
    // the worker pool.  Neither runs on the scheduling thread.
    size_t chars = 0;
    for (const std::string &c : request.options.coefficients) {
        chars += c.size();
    }
    if (chars <= kInlineCheckChars) {
        impl_->check(std::move(request), at);
        return;
    }
    Impl *impl = impl_.get();
    impl_->pool->post([impl, request = std::move(request), at]() mutable {
        impl->check(std::move(request), at);
    });
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
