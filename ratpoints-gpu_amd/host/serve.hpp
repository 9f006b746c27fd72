// Serving loop: many point searches share one GPU (continuous batching).
//
// No counterpart in the reference, which searches one curve per process
// (src/main.cpp:69-72 gives the -1 semantics kept here).  Requests are
// admitted while the device memory they need -- a residue table and a share
// of the survivor buffer -- fits a budget and a slot is free; every admitted
// request then advances by a chunk of its denominators in each round, one
// persistent many-curve launch per round, so requests that arrive while
// others run join at the next round and small searches finish in one or two
// rounds instead of waiting for large ones.  Residue tables stay on the device
// between rounds.  Results are exact and identical to a search of the same
// request alone: the point at infinity first, then points in increasing
// (b, a) order, filtered by the intervals; a -1 request stops after its first
// point and its remaining denominators are never scheduled.
#ifndef RATPOINTS_GPU_SERVE_HPP
#define RATPOINTS_GPU_SERVE_HPP

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "point_search.hpp"

namespace ratpoints_gpu {

struct ServeRequest {
    int id = 0;
    SearchOptions options;     // coefficients, bounds, intervals, infinity, square mode
    bool first_point = false;  // -1: stop after the first point
};

struct ServeResult {
    int id = 0;
    std::vector<PointPair> points;
    std::string error;         // empty on success
    double submit_ms = 0.0;    // on the server's clock
    double admit_ms = 0.0;
    double finish_ms = 0.0;
    double sites = 0.0;        // sites sieved for this request
    int rounds = 0;
};

struct ServeOptions {
    double round_sites = 4.0e11;  // sieve work of one round, shared by the active requests
    int max_active = 256;         // requests on the device at once
    double memory_budget = 8.0 * 1024 * 1024 * 1024;  // device bytes for admitted requests
    int device = 0;
    bool progress = false;           // one line per round on stderr
    long long launch_timeout_ms = 300000;  // a round's kernel wait; longer fails the round
};

struct ServeStats {
    double sites = 0.0;
    double kernel_ms = 0.0;
    unsigned long long rounds = 0;
    unsigned long long admitted = 0;
    unsigned long long peak_active = 0;
};

using ServeCallback = std::function<void(const ServeResult &)>;

// Device bytes an admitted request holds: its residue table and room for
// the survivors of its largest chunk.
double request_footprint(const ServeRequest &request, const ServeOptions &options);

class Server {
public:
    // on_done runs on the server's scheduling thread.
    Server(ServeOptions options, ServeCallback on_done);
    ~Server();  // finishes every submitted request, then stops

    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Thread-safe.  The admission checks of a request with short
    // coefficients run before submit returns, the others on a worker.
    void submit(ServeRequest request);
    void drain();                       // until every submitted request finished
    double now_ms() const;
    ServeStats stats() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// Baseline: one request at a time in submission order, each searched alone
// by PointSearch (the command line's path).  Same result fields.
ServeResult search_alone(const ServeRequest &request);

}  // namespace ratpoints_gpu

#endif
