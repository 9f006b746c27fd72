// ratpoints_serve: replays a request trace through the serving loop (or, as
// the baseline, through one search at a time) and reports per-request
// results and latency percentiles.
//
// Trace: one request per line,
//   id arrival_ms flags bound den_first den_last n_intervals
//      [lower upper]... n_coefficients c0 ... cn
// flags: 1 = -1 (first point only), 2 = square denominators, 4 = no point at
// infinity; interval ends are decimal rationals p/q (or p) or -inf / inf.
//
// Output (--out): per request "id latency_ms sites rounds n_points
// a:y:b ..." or "id error <message>"; stdout: one JSON summary line.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "serve.hpp"

using namespace ratpoints_gpu;

namespace {

struct TraceEntry {
    ServeRequest request;
    double arrival_ms = 0.0;
};

mpq_class parse_rational(const std::string &text) {
    mpq_class q(text, 10);
    q.canonicalize();
    return q;
}

std::vector<TraceEntry> read_trace(const std::string &path) {
    std::ifstream in(path);
    if (!in) {
        throw std::runtime_error("cannot read trace " + path);
    }
    std::vector<TraceEntry> trace;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') {
            continue;
        }
        std::istringstream s(line);
        TraceEntry e;
        int flags = 0, n_int = 0, n_coef = 0;
        long long first = 0, last = 0;
        s >> e.request.id >> e.arrival_ms >> flags >> e.request.options.numerator_bound
          >> first >> last >> n_int;
        if (!s) {
            throw std::runtime_error("bad trace line: " + line);
        }
        for (int k = 0; k < n_int; ++k) {
            std::string lo, hi;
            s >> lo >> hi;
            SearchInterval iv;
            iv.lower_infinite = lo == "-inf";
            iv.upper_infinite = hi == "inf";
            if (!iv.lower_infinite) iv.lower = parse_rational(lo);
            if (!iv.upper_infinite) iv.upper = parse_rational(hi);
            e.request.options.intervals.push_back(iv);
        }
        s >> n_coef;
        for (int k = 0; k < n_coef; ++k) {
            std::string c;
            s >> c;
            e.request.options.coefficients.push_back(c);
        }
        if (!s) {
            throw std::runtime_error("bad trace line: " + line);
        }
        e.request.options.denominators = {static_cast<int>(first), static_cast<int>(last)};
        e.request.first_point = (flags & 1) != 0;
        e.request.options.square_denominators = (flags & 2) != 0;
        e.request.options.include_infinity = (flags & 4) == 0;
        trace.push_back(std::move(e));
    }
    std::stable_sort(trace.begin(), trace.end(), [](const TraceEntry &a, const TraceEntry &b) {
        return a.arrival_ms < b.arrival_ms;
    });
    return trace;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) {
        return 0.0;
    }
    std::sort(v.begin(), v.end());
    const size_t k = static_cast<size_t>(std::ceil(q * static_cast<double>(v.size())));
    return v[std::min(v.size() - 1, k ? k - 1 : 0)];
}

void usage() {
    std::cerr << "usage: ratpoints_serve --trace FILE [--mode serve|alone] [--out FILE]\n"
                 "       [--round-sites X] [--max-active N] [--memory-budget BYTES]\n"
                 "       [--time-scale S] [--progress] [--launch-timeout-ms MS]\n";
}

}  // namespace

int main(int argc, char **argv) {
    std::string trace_path, out_path, mode = "serve";
    ServeOptions options;
    double time_scale = 1.0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--trace") trace_path = value();
        else if (a == "--out") out_path = value();
        else if (a == "--mode") mode = value();
        else if (a == "--round-sites") options.round_sites = std::stod(value());
        else if (a == "--max-active") options.max_active = std::stoi(value());
        else if (a == "--memory-budget") options.memory_budget = std::stod(value());
        else if (a == "--time-scale") time_scale = std::stod(value());
        else if (a == "--progress") options.progress = true;
        else if (a == "--launch-timeout-ms") options.launch_timeout_ms = std::stoll(value());
        else {
            usage();
            return 2;
        }
    }
    if (trace_path.empty() || (mode != "serve" && mode != "alone")) {
        usage();
        return 2;
    }
    std::vector<TraceEntry> trace;
    try {
        trace = read_trace(trace_path);
    } catch (const std::exception &e) {
        std::cerr << "ratpoints_serve: " << e.what() << "\n";
        return 2;
    }

    std::map<int, ServeResult> results;
    std::map<int, double> arrival;
    std::mutex lock;
    const auto start = std::chrono::steady_clock::now();
    auto since_start = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start)
            .count();
    };
    auto wait_until = [&](double ms) {
        const double now = since_start();
        if (ms > now) {
            std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(ms - now));
        }
    };
    // Progress on stderr every 10 s: requests finished so far.
    std::atomic<size_t> finished{0};
    std::atomic<bool> running{true};
    std::thread progress([&] {
        double next = 10000.0;
        while (running.load()) {
            std::this_thread::sleep_for(std::chrono::milliseconds(200));
            if (since_start() >= next) {
                std::fprintf(stderr, "ratpoints_serve: %s %.0f s, %zu of %zu requests done\n",
                             mode.c_str(), since_start() / 1e3, finished.load(), trace.size());
                next += 10000.0;
            }
        }
    });
    ServeStats stats;
    if (mode == "serve") {
        Server server(options, [&](const ServeResult &r) {
            std::lock_guard<std::mutex> g(lock);
            ServeResult done = r;
            done.finish_ms = since_start();
            results[r.id] = std::move(done);
            ++finished;
        });
        for (const TraceEntry &e : trace) {
            const double at = e.arrival_ms * time_scale;
            wait_until(at);
            {
                std::lock_guard<std::mutex> g(lock);
                arrival[e.request.id] = at;
            }
            server.submit(e.request);
        }
        server.drain();
        stats = server.stats();
    } else {
        for (const TraceEntry &e : trace) {
            const double at = e.arrival_ms * time_scale;
            wait_until(at);
            arrival[e.request.id] = at;
            ServeResult r = search_alone(e.request);
            r.finish_ms = since_start();
            stats.sites += r.sites;
            stats.rounds += static_cast<unsigned long long>(r.rounds);
            results[r.id] = std::move(r);
            ++finished;
        }
    }
    const double makespan_ms = since_start();
    running = false;
    progress.join();

    std::vector<double> latency;
    unsigned long long errors = 0, points = 0;
    double sites = 0.0;
    std::ofstream out;
    if (!out_path.empty()) {
        out.open(out_path);
    }
    for (auto &kv : results) {
        const ServeResult &r = kv.second;
        const double lat = r.finish_ms - arrival[kv.first];
        latency.push_back(lat);
        sites += r.sites;
        points += r.points.size();
        if (!r.error.empty()) {
            ++errors;
        }
        if (out) {
            if (!r.error.empty()) {
                out << r.id << " error " << r.error << "\n";
                continue;
            }
            out << r.id << " " << lat << " " << r.sites << " " << r.rounds << " "
                << r.points.size();
            for (const PointPair &p : r.points) {
                out << " " << p.numerator << ":" << p.ordinate << ":" << p.denominator;
            }
            out << "\n";
        }
    }
    double mean = 0.0;
    for (double v : latency) {
        mean += v / static_cast<double>(latency.size());
    }
    char summary[1024];
    std::snprintf(summary, sizeof summary,
                  "{\"mode\": \"%s\", \"requests\": %zu, \"errors\": %llu, \"points\": %llu, "
                  "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"mean_ms\": %.3f, \"max_ms\": %.3f, "
                  "\"makespan_s\": %.3f, \"sites\": %.6g, \"gsites_per_s\": %.1f, "
                  "\"rounds\": %llu, \"peak_active\": %llu, \"kernel_ms\": %.1f}",
                  mode.c_str(), results.size(), errors, points, percentile(latency, 0.5),
                  percentile(latency, 0.99),
                  mean,
                  latency.empty() ? 0.0 : *std::max_element(latency.begin(), latency.end()),
                  makespan_ms / 1e3, sites, sites / (makespan_ms * 1e6), stats.rounds,
                  stats.peak_active, stats.kernel_ms);
    std::cout << summary << std::endl;
    return 0;
}
