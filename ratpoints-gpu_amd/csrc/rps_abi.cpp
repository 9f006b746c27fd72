// C ABI of the gfx950 modular square sieve: context, per-curve device
// tables, denominator batching, the double-buffered survivor stream and the
// persistent many-curve launch.
//
// Host orchestration replaces run_modular_sieve (src/sieve.cu:503-580 in the
// reference): the reference rebuilds its 24 MB basis table and reallocates
// every buffer on each call; here a context keeps them resident, caches the
// per-curve residue table across calls, and overlaps the survivor download and
// the caller's sink with the next batch's kernel (two buffer slots, a compute
// stream and a copy stream).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "ratpoints_sieve.h"
#include "sieve_kernels.hpp"
#include "sieve_layout.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

thread_local std::string g_open_error;

// The 36 largest primes below 512, ascending (src/sieve_plan.cpp:60-61).
std::vector<int> candidate_primes() {
    std::vector<int> primes;
    for (int n = 2; n < rps::kMaxPrime; ++n) {
        bool prime = true;
        for (int d = 2; d * d <= n; ++d) {
            if (n % d == 0) {
                prime = false;
                break;
            }
        }
        if (prime) {
            primes.push_back(n);
        }
    }
    return std::vector<int>(primes.end() - rps::kCandidates, primes.end());
}

bool is_decimal(const char *text) {
    if (!text) {
        return false;
    }
    if (*text == '-' || *text == '+') {
        ++text;
    }
    if (!*text) {
        return false;
    }
    for (; *text; ++text) {
        if (*text < '0' || *text > '9') {
            return false;
        }
    }
    return true;
}

// u^-1 mod p for 0 < u < p, p prime (extended Euclid).
int mod_inverse(int u, int p) {
    long long r0 = p, r1 = u, t0 = 0, t1 = 1;
    while (r1 != 0) {
        const long long q = r0 / r1;
        const long long r2 = r0 - q * r1;
        const long long t2 = t0 - q * t1;
        r0 = r1;
        r1 = r2;
        t0 = t1;
        t1 = t2;
    }
    return static_cast<int>(t0 < 0 ? t0 + p : t0);
}

struct HipError {
    hipError_t code;
    const char *op;
};

#define RPS_HIP(op)                                                        \
    do {                                                                   \
        hipError_t rps_status_ = (op);                                     \
        if (rps_status_ != hipSuccess) {                                   \
            throw HipError{rps_status_, #op};                              \
        }                                                                  \
    } while (0)

struct ApiError {
    int code;
    std::string message;
};

std::string op_name(const char *expr) {
    // "hipMemcpyAsync(dst, ...)" -> "hipMemcpyAsync"
    std::string s(expr);
    size_t paren = s.find('(');
    return paren == std::string::npos ? s : s.substr(0, paren);
}

// Polls `event` until it completes or `timeout_ms` passes; on expiry throws
// RPS_EDEVICE naming the wait.  Used for every wait on a sieve kernel, so a
// kernel that does not finish ends the call with an error instead of
// blocking the caller.
struct TimedOut {
    std::string message;
};

// The first millisecond is spent polling (a short launch is seen at once),
// later polls are 20 us apart.
void wait_event(hipEvent_t event, long long timeout_ms, const std::string &what) {
    const auto start = Clock::now();
    const auto deadline = start + std::chrono::milliseconds(timeout_ms);
    const auto spin_until = start + std::chrono::milliseconds(1);
    for (;;) {
        const hipError_t e = hipEventQuery(event);
        if (e == hipSuccess) {
            return;
        }
        if (e != hipErrorNotReady) {
            throw HipError{e, "hipEventQuery"};
        }
        const auto now = Clock::now();
        if (now > deadline) {
            throw TimedOut{what + " did not finish within " + std::to_string(timeout_ms) + " ms"};
        }
        if (now < spin_until) {
            std::this_thread::yield();
        } else {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
}

struct Slot {
    rps::DeviceState *d_state = nullptr;
    rps::DeviceState *h_state = nullptr;  // pinned
    long long *d_num = nullptr;
    int *d_den = nullptr;
    unsigned long long *d_cursors = nullptr;
    size_t cursor_capacity = 0;  // entries
    hipEvent_t k_start = nullptr, k_end = nullptr, ready = nullptr;
};

// One launch of the plain path: a batch of denominators of one curve.
struct Batch {
    rps::LineDesc line{};
    size_t tiles = 0;
};

// Device buffer that only grows.
template <typename T>
struct Grow {
    T *ptr = nullptr;
    size_t size = 0;

    void ensure(size_t n) {
        if (n <= size) {
            return;
        }
        if (ptr) {
            RPS_HIP(hipFree(ptr));
            ptr = nullptr;
            size = 0;
        }
        RPS_HIP(hipMalloc(&ptr, n * sizeof(T)));
        size = n;
    }

    // As ensure, but at least doubling and at least `floor` elements, for
    // per-call buffers whose size varies from call to call (each
    // reallocation costs a device synchronisation).
    void ensure_grow(size_t n, size_t floor) {
        if (n > size) {
            ensure(std::max({n, 2 * size, floor}));
        }
    }

    void release() {
        if (ptr) {
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        size = 0;
    }
};

// Pinned host buffer that only grows.
template <typename T>
struct PinnedGrow {
    T *ptr = nullptr;
    size_t size = 0;

    void ensure(size_t n) {
        if (n <= size) {
            return;
        }
        if (ptr) {
            RPS_HIP(hipHostFree(ptr));
            ptr = nullptr;
            size = 0;
        }
        RPS_HIP(hipHostMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T), 0));
        size = n;
    }

    // At least doubling and at least `floor` elements (pinned allocations
    // take milliseconds).
    void ensure_grow(size_t n, size_t floor) {
        if (n > size) {
            ensure(std::max({n, 2 * size, floor}));
        }
    }

    void release() {
        if (ptr) {
            (void)hipHostFree(ptr);
        }
        ptr = nullptr;
        size = 0;
    }
};

// Byte offsets of the sections of a many-curve call's plan, each aligned to
// 256 bytes, so that the whole plan goes to the device in one copy.
struct PlanLayout {
    size_t coef = 0, slot = 0, negmod = 0, curves = 0, tile_start = 0, bucket = 0, bytes = 0;

    PlanLayout(size_t coef_ints, size_t n_build, size_t n, size_t n_buckets) {
        auto take = [this](size_t b) {
            const size_t at = bytes;
            bytes += (std::max<size_t>(b, 1) + 255) / 256 * 256;
            return at;
        };
        coef = take(coef_ints * sizeof(int));
        slot = take(n_build * sizeof(int));
        negmod = take(n * rps::kCandidates * sizeof(int));
        curves = take(n * sizeof(rps::CurveDesc));
        tile_start = take((n + 1) * sizeof(unsigned long long));
        bucket = take(n_buckets * sizeof(int));
    }
};

// Control block of a many-curve call on the device: the survivor state, the
// tile counter and the per-curve failure flags, read back in one copy after
// every launch.
struct ControlLayout {
    static constexpr size_t kState = 0;
    static constexpr size_t kHead = 32;
    static constexpr size_t kFailed = 64;
    static_assert(sizeof(rps::DeviceState) <= kHead, "state fits its section");
    static size_t bytes(size_t n) { return kFailed + n * sizeof(int); }
};

// Survivors of a many-curve launch copied back with its control block
// before their count is known: a launch with at most this many needs no
// second copy.
constexpr size_t kSpeculativeSurvivors = 4096;

}  // namespace

struct rps_ctx {
    int device = 0;
    int cu_count = 256;
    hipStream_t compute = nullptr;
    hipStream_t copy = nullptr;
    std::string error;
    std::vector<int> primes;
    int *d_primes = nullptr;
    int *d_inverses = nullptr;  // [36][512] u^-1 mod p, in d_primes' allocation
    int *d_coef_res = nullptr;
    unsigned long long *d_chi = nullptr;
    unsigned long long *d_resid = nullptr;
    int *d_negmod = nullptr;
    int *h_negmod = nullptr;  // pinned
    int *h_coef = nullptr;    // pinned [36][101]: coefficient residues for rps_sieve_into
    unsigned long long *h_head = nullptr;  // pinned: tiles handed out by a many-curve launch
    std::vector<std::string> curve_key;
    bool curve_ready = false;
    Slot slots[2];
    size_t capacity = size_t(1) << 24;
    size_t allocated_capacity = 0;
    long long *h_num = nullptr;  // pinned staging
    int *h_den = nullptr;
    int *h_req = nullptr;
    size_t stage_capacity = size_t(1) << 20;
    int batch_size = 1 << 16;
    bool batch_forced = false;  // batch_size set by the caller: the table's is not used
    int variant = -1;  // -1: chosen per request by the tuning table
    int run_log2 = -1;  // -1: the tuning table's (or a forced variant's own) run length
    long long min_tiles = 2048;
    // The short-line kernel against the 11-row tile, kernel Tsite/s: 15.3
    // against 10.4 at 1,024 words, 15.7 against 14.5 at 1,536, 16.6 against
    // 16.5 at 2,048 (its limit).
    unsigned long long direct_words = 2048;
    int primitive = 0;  // emit only survivors with gcd(a, b) = 1
    // A kernel wait longer than this fails the call (RPS_EDEVICE) instead
    // of blocking; the context then refuses further work, since its streams
    // still hold the kernel.
    long long launch_timeout_ms = 600000;
    bool broken = false;
    hipEvent_t basis_start = nullptr, basis_end = nullptr;
    double last_kernel_ms = 0.0;
    uint64_t last_launches = 0;
    double last_sites = 0.0;
    // Many-curve launches.
    Grow<unsigned long long> m_chi;
    // Residue tables of recent curves stay on the device between calls of
    // rps_sieve_many; the least recently used are replaced.
    Grow<unsigned long long> t_resid;  // [t_cap][36][512][9]
    int t_cap = 0;
    long long table_slots = 256;
    std::map<std::vector<std::string>, int> t_index;
    std::vector<std::vector<std::string>> t_key;  // per slot
    std::vector<uint64_t> t_used;                 // per slot: last call
    uint64_t t_call = 0;
    Grow<unsigned char> m_plan;         // PlanLayout: coefficient residues, table
                                        // slots, (-A) mod p, curves, tiles, buckets
    PinnedGrow<unsigned char> h_plan;   // its host copy
    Grow<unsigned char> m_ctrl;         // ControlLayout
    PinnedGrow<unsigned char> h_ctrl;
    Grow<int> m_req;
};

namespace {

void free_slot(Slot &s) {
    if (s.d_state) (void)hipFree(s.d_state);
    if (s.h_state) (void)hipHostFree(s.h_state);
    if (s.d_num) (void)hipFree(s.d_num);
    if (s.d_den) (void)hipFree(s.d_den);
    if (s.d_cursors) (void)hipFree(s.d_cursors);
    if (s.k_start) (void)hipEventDestroy(s.k_start);
    if (s.k_end) (void)hipEventDestroy(s.k_end);
    if (s.ready) (void)hipEventDestroy(s.ready);
    s = Slot{};
}

void ensure_capacity(rps_ctx *ctx) {
    if (ctx->allocated_capacity == ctx->capacity) {
        return;
    }
    // Until both slots hold buffers of the new size, no size is recorded: a
    // failed allocation must not leave a slot without buffers looking ready.
    ctx->allocated_capacity = 0;
    for (Slot &s : ctx->slots) {
        if (s.d_num) {
            RPS_HIP(hipFree(s.d_num));
            s.d_num = nullptr;
        }
        if (s.d_den) {
            RPS_HIP(hipFree(s.d_den));
            s.d_den = nullptr;
        }
        RPS_HIP(hipMalloc(&s.d_num, ctx->capacity * sizeof(long long)));
        RPS_HIP(hipMalloc(&s.d_den, ctx->capacity * sizeof(int)));
    }
    ctx->allocated_capacity = ctx->capacity;
}

void ensure_cursors(Slot &s, size_t entries) {
    if (s.cursor_capacity >= entries) {
        return;
    }
    if (s.d_cursors) {
        RPS_HIP(hipFree(s.d_cursors));
        s.d_cursors = nullptr;
    }
    size_t n = std::max(entries, size_t(1) << 16);
    RPS_HIP(hipMalloc(&s.d_cursors, n * sizeof(unsigned long long)));
    s.cursor_capacity = n;
}

void validate(const rps_request *req) {
    if (!req) {
        throw ApiError{RPS_EINVAL, "null request"};
    }
    if (req->n_coefficients < 2 || req->n_coefficients > rps::kMaxCoefficients
        || !req->coefficients) {
        throw ApiError{RPS_EINVAL, "supported polynomial degrees are 1 through 100"};
    }
    for (int k = 0; k < req->n_coefficients; ++k) {
        if (!is_decimal(req->coefficients[k])) {
            throw ApiError{RPS_EINVAL, "coefficients must be decimal integers"};
        }
    }
    if (req->numerator_bound < 1 || req->numerator_bound > (LLONG_MAX - 32) / 2) {
        throw ApiError{RPS_EINVAL, "invalid numerator bound"};
    }
    // A line is at most 2^22 tiles of at most 2^25 words (2^47 words,
    // A < 4.5e15): the same limit on every path, so a request that passes
    // here never fails later for its size.
    if ((2ULL * static_cast<unsigned long long>(req->numerator_bound) + 64) / 64 > (1ULL << 47)) {
        throw ApiError{RPS_EINVAL, "maximum height exceeds GPU limit"};
    }
    // src/sieve_plan.cpp:55-58
    if (req->den_first < 1 || req->den_first > req->den_last
        || (req->square_denominators && req->den_last > rps::kMaxParamT)) {
        throw ApiError{RPS_EINVAL, "invalid denominator range"};
    }
}

// Launch shape per line length from the autotuner's table (autotune.py
// writes tuning_table.inc), for the plain path (rps_sieve and its streaming
// forms) and for many-curve launches (rps_sieve_many): the tile variant
// (dense primes, words per lane, waves per workgroup, row width), the run
// length (whole lines per row tile up to 2^run_log2 words) and, for the
// plain path, the denominators per launch.  The first entry whose bucket
// holds `words` applies; a name that is not compiled in falls back to
// variant 0 with its own run length.
struct TuneEntry {
    unsigned long long max_words;
    const char *variant;
    int run_log2;
    int batch;
    const char *many_variant;
    int many_run_log2;
};

const TuneEntry kTuning[] = {
#include "tuning_table.inc"
};

struct Tuned {
    int variant = 0;
    int run_log2 = 0;
    int batch = 0;  // 0: the context's batch_size
};

int variant_index(const char *name) {
    for (int v = 0; v < rps::variant_count(); ++v) {
        if (std::strcmp(rps::variant(v)->name, name) == 0) {
            return v;
        }
    }
    return -1;
}

Tuned tuned(unsigned long long words, bool many) {
    Tuned t;
    t.run_log2 = rps::variant(0)->run_log2;
    for (const TuneEntry &e : kTuning) {
        if (words <= e.max_words) {
            const int v = variant_index(many ? e.many_variant : e.variant);
            if (v >= 0) {
                t.variant = v;
                t.run_log2 = many ? e.many_run_log2 : e.run_log2;
                t.batch = many ? 0 : e.batch;
            }
            break;
        }
    }
    return t;
}

int tuned_variant(unsigned long long words) { return tuned(words, false).variant; }

// What a request runs with: the caller's options where set, else the tuning
// table's choice (a forced variant keeps its own run length).
struct Shape {
    const rps::VariantInfo *var;
    int run_log2;
    int batch;
};

Shape launch_shape(const rps_ctx *ctx, unsigned long long words, bool many, int forced = -1) {
    const Tuned t = tuned(words, many);
    const int chosen = forced >= 0 ? forced : ctx->variant >= 0 ? ctx->variant : t.variant;
    Shape sh{rps::variant(chosen), 0, 0};
    sh.run_log2 = ctx->run_log2 >= 0 ? ctx->run_log2
                  : chosen == t.variant ? t.run_log2 : sh.var->run_log2;
    sh.batch = ctx->batch_forced || t.batch <= 0 ? ctx->batch_size : t.batch;
    return sh;
}

// Line geometry of a request: words per line, last-word mask, and the
// fields of LineDesc that do not depend on the batch.
rps::LineDesc line_geometry(const rps_request *req) {
    rps::LineDesc L{};
    const unsigned long long A = static_cast<unsigned long long>(req->numerator_bound);
    L.bound = req->numerator_bound;
    L.words = (2ULL * A + 64) / 64;
    const unsigned long long valid_last = 2ULL * A - 64ULL * (L.words - 1) + 1;
    L.last_mask = valid_last >= 64 ? ~0ULL : ((1ULL << valid_last) - 1);
    L.square = req->square_denominators ? 1 : 0;
    L.tiles_per_t = 1;
    L.tile_t = 1;
    L.tile_words = L.words;
    return L;
}

// Residues of every coefficient (signed decimal text) for every candidate,
// in [0, p).  Each coefficient's text is cut into 9-digit chunks once, and
// the chunks are folded per prime, v = (v * (10^digits mod p) + chunk) mod p,
// so the digits are read once rather than once per prime.
void coefficient_residues(const rps_request *req, const std::vector<int> &primes,
                          int n_coef, int *out /* [36][n_coef] */) {
    std::vector<long long> chunk, scale;
    for (int k = 0; k < n_coef; ++k) {
        if (k >= req->n_coefficients) {
            for (int i = 0; i < rps::kCandidates; ++i) {
                out[static_cast<size_t>(i) * n_coef + k] = 0;
            }
            continue;
        }
        const char *text = req->coefficients[k];
        const bool negative = *text == '-';
        if (*text == '-' || *text == '+') {
            ++text;
        }
        chunk.clear();
        scale.clear();
        while (*text) {
            long long c = 0, sc = 1;
            for (int d = 0; d < 9 && *text; ++d, ++text) {
                c = 10 * c + (*text - '0');
                sc *= 10;
            }
            chunk.push_back(c);
            scale.push_back(sc);
        }
        for (int i = 0; i < rps::kCandidates; ++i) {
            const long long p = primes[i];
            long long v = 0;
            for (size_t j = 0; j < chunk.size(); ++j) {
                v = (v * (scale[j] % p) + chunk[j]) % p;
            }
            out[static_cast<size_t>(i) * n_coef + k] = static_cast<int>(negative && v ? p - v : v);
        }
    }
}

void prepare_curve(rps_ctx *ctx, const rps_request *req, rps_stats *st) {
    std::vector<std::string> key(req->coefficients,
                                 req->coefficients + req->n_coefficients);
    if (ctx->curve_ready && key == ctx->curve_key) {
        return;
    }
    const int n = req->n_coefficients;
    std::vector<int> res(static_cast<size_t>(rps::kCandidates) * n);
    coefficient_residues(req, ctx->primes, n, res.data());
    ctx->curve_ready = false;
    RPS_HIP(hipMemcpyAsync(ctx->d_coef_res, res.data(), res.size() * sizeof(int),
                           hipMemcpyHostToDevice, ctx->compute));
    RPS_HIP(hipEventRecord(ctx->basis_start, ctx->compute));
    RPS_HIP(rps::launch_chi(ctx->d_coef_res, n, 1, ctx->d_primes, ctx->d_chi, ctx->compute));
    RPS_HIP(rps::launch_resid(ctx->d_primes, ctx->d_inverses, ctx->d_chi, 1, nullptr,
                              ctx->d_resid, ctx->compute));
    RPS_HIP(hipEventRecord(ctx->basis_end, ctx->compute));
    wait_event(ctx->basis_end, ctx->launch_timeout_ms, "residue-table build");
    float ms = 0.0f;
    RPS_HIP(hipEventElapsedTime(&ms, ctx->basis_start, ctx->basis_end));
    if (st) {
        st->basis_ms += ms;
    }
    ctx->curve_key = std::move(key);
    ctx->curve_ready = true;
}

// Number of word chunks each denominator line is split into.  Lines are
// split only when a batch has too few denominators to fill the device, and
// always into chunks of at most 2^25 words.  A chunk keeps at least eight strides per wave so the per-tile
// row build stays amortised.
unsigned long long line_chunks(long long min_tiles, long long n_t,
                               unsigned long long words, int stride, int waves) {
    unsigned long long chunks = 1;
    if (n_t < min_tiles) {
        chunks = (static_cast<unsigned long long>(min_tiles) + n_t - 1) / n_t;
    }
    const unsigned long long min_words = 8ULL * stride * waves;
    chunks = std::min(chunks, std::max(1ULL, words / min_words));
    // At most 2^25 words per tile: the kernels index words relative to the
    // tile start in 32 bits and form 64 x index + offset without overflow.
    const unsigned long long limit = 1ULL << 25;
    return std::max(chunks, (words + limit - 1) / limit);
}

// Fills the tiling fields of L for row tiles of `chunks` chunks per line.
void row_tiling(rps::LineDesc &L, unsigned long long chunks, int stride) {
    unsigned long long tile = (L.words + chunks - 1) / chunks;
    tile = (tile + stride - 1) / stride * stride;
    L.direct = 0;
    L.tile_words = tile;
    L.tiles_per_t = static_cast<int>((L.words + tile - 1) / tile);
    L.tile_t = 1;
}

// Whole-line row tiles of lines shorter than run_words take several
// consecutive denominators (a run) so that per-tile work stays near
// run_words words (the variant's choice: one workgroup per CU wants longer
// runs, whose next-line prefetch hides the string loads); at most max_run
// lines per tile.
void row_runs(rps::LineDesc &L, long long max_run, int run_log2) {
    const unsigned long long run_words = 1ULL << run_log2;
    if (L.tiles_per_t != 1 || L.words >= run_words) {
        return;
    }
    const long long want = static_cast<long long>(run_words / L.words);
    L.tile_t = static_cast<int>(std::max(1LL, std::min(want, max_run)));
}

// Short lines: a tile is a run of lines of about kDirectRunWords words in
// all (swept in blocks of kDirectMaxDenominators lines), at most max_run lines.
constexpr unsigned long long kDirectRunWords = 1ULL << 16;

void direct_tiling(rps::LineDesc &L, long long max_run) {
    const long long want = static_cast<long long>((kDirectRunWords + L.words - 1) / L.words);
    L.direct = 1;
    L.tile_t = static_cast<int>(std::max(1LL, std::min(want, max_run)));
}

size_t tile_count(const rps::LineDesc &L) {
    const size_t runs = (static_cast<size_t>(L.n_t) + L.tile_t - 1) / L.tile_t;
    return L.direct ? runs : runs * static_cast<size_t>(L.tiles_per_t);
}

constexpr unsigned long long kMaxTilesPerLaunch = 1ULL << 22;

int fail(rps_ctx *ctx, const HipError &e) {
    ctx->error = op_name(e.op) + ": " + hipGetErrorString(e.code);
    (void)hipStreamSynchronize(ctx->compute);
    (void)hipStreamSynchronize(ctx->copy);
    ctx->curve_ready = false;
    return RPS_EDEVICE;
}

int fail(rps_ctx *ctx, const ApiError &e) {
    ctx->error = e.message;
    (void)hipStreamSynchronize(ctx->compute);
    (void)hipStreamSynchronize(ctx->copy);
    return e.code;
}

// A kernel that did not finish: no stream is synchronised (that would block
// behind it), and the context takes no more work.
int fail(rps_ctx *ctx, const TimedOut &e) {
    ctx->error = e.message + "; the context is unusable, close it";
    ctx->broken = true;
    ctx->curve_ready = false;
    return RPS_EDEVICE;
}

void check_usable(const rps_ctx *ctx) {
    if (ctx->broken) {
        throw ApiError{RPS_EDEVICE, "context unusable after a kernel timeout; close it"};
    }
}

}  // namespace

extern "C" {

int rps_open(int device, rps_ctx **out) {
    if (!out) {
        g_open_error = "rps_open: null output pointer";
        return RPS_EINVAL;
    }
    *out = nullptr;
    rps_ctx *ctx = new (std::nothrow) rps_ctx;
    if (!ctx) {
        g_open_error = "rps_open: out of host memory";
        return RPS_ENOMEM;
    }
    try {
        int count = 0;
        RPS_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) {
            throw ApiError{RPS_EINVAL, "invalid GPU device " + std::to_string(device)};
        }
        ctx->device = device;
        RPS_HIP(hipSetDevice(device));
        // The allocation-free entry point's scratch address on this device.
        RPS_HIP(rps::scratch_init(device));
        hipDeviceProp_t prop;
        RPS_HIP(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        RPS_HIP(hipStreamCreateWithFlags(&ctx->compute, hipStreamNonBlocking));
        RPS_HIP(hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
        ctx->primes = candidate_primes();
        // The candidates, then u^-1 mod p for each candidate and u (0 for u
        // = 0 and u >= p), which every residue-table build reads.
        std::vector<int> constants(static_cast<size_t>(rps::kCandidates) * (1 + rps::kMaxPrime), 0);
        for (int i = 0; i < rps::kCandidates; ++i) {
            const int p = ctx->primes[i];
            constants[i] = p;
            int *inv = constants.data() + rps::kCandidates + static_cast<size_t>(i) * rps::kMaxPrime;
            for (int u = 1; u < p; ++u) {
                inv[u] = mod_inverse(u, p);
            }
        }
        RPS_HIP(hipMalloc(&ctx->d_primes, constants.size() * sizeof(int)));
        RPS_HIP(hipMemcpy(ctx->d_primes, constants.data(), constants.size() * sizeof(int),
                          hipMemcpyHostToDevice));
        ctx->d_inverses = ctx->d_primes + rps::kCandidates;
        RPS_HIP(hipMalloc(&ctx->d_coef_res, static_cast<size_t>(rps::kCandidates)
                              * rps::kMaxCoefficients * sizeof(int)));
        RPS_HIP(hipMalloc(&ctx->d_chi, static_cast<size_t>(rps::kCandidates)
                              * rps::kChiWords * sizeof(unsigned long long)));
        RPS_HIP(hipMalloc(&ctx->d_resid, rps::kResidTableWords * sizeof(unsigned long long)));
        RPS_HIP(hipMalloc(&ctx->d_negmod, rps::kCandidates * sizeof(int)));
        RPS_HIP(hipHostMalloc(&ctx->h_negmod, rps::kCandidates * sizeof(int)));
        RPS_HIP(hipHostMalloc(&ctx->h_coef, static_cast<size_t>(rps::kCandidates)
                                  * rps::kMaxCoefficients * sizeof(int)));
        RPS_HIP(hipHostMalloc(&ctx->h_head, sizeof(unsigned long long)));
        RPS_HIP(hipEventCreate(&ctx->basis_start));
        RPS_HIP(hipEventCreate(&ctx->basis_end));
        for (Slot &s : ctx->slots) {
            RPS_HIP(hipMalloc(&s.d_state, sizeof(rps::DeviceState)));
            RPS_HIP(hipHostMalloc(&s.h_state, sizeof(rps::DeviceState)));
            RPS_HIP(hipEventCreate(&s.k_start));
            RPS_HIP(hipEventCreate(&s.k_end));
            RPS_HIP(hipEventCreateWithFlags(&s.ready, hipEventDisableTiming));
        }
        RPS_HIP(hipHostMalloc(&ctx->h_num, ctx->stage_capacity * sizeof(long long)));
        RPS_HIP(hipHostMalloc(&ctx->h_den, ctx->stage_capacity * sizeof(int)));
        RPS_HIP(hipHostMalloc(&ctx->h_req, ctx->stage_capacity * sizeof(int)));
    } catch (const HipError &e) {
        g_open_error = op_name(e.op) + ": " + hipGetErrorString(e.code);
        rps_close(ctx);
        return RPS_EDEVICE;
    } catch (const ApiError &e) {
        g_open_error = e.message;
        rps_close(ctx);
        return e.code;
    }
    *out = ctx;
    return RPS_OK;
}

void rps_close(rps_ctx *ctx) {
    if (!ctx) {
        return;
    }
    if (ctx->broken) {
        // A kernel that timed out may still run: its streams and buffers
        // are left to the process's end rather than waited for.
        delete ctx;
        return;
    }
    (void)hipSetDevice(ctx->device);
    if (ctx->compute) (void)hipStreamSynchronize(ctx->compute);
    if (ctx->copy) (void)hipStreamSynchronize(ctx->copy);
    for (Slot &s : ctx->slots) {
        free_slot(s);
    }
    if (ctx->d_primes) (void)hipFree(ctx->d_primes);
    if (ctx->d_coef_res) (void)hipFree(ctx->d_coef_res);
    if (ctx->d_chi) (void)hipFree(ctx->d_chi);
    if (ctx->d_resid) (void)hipFree(ctx->d_resid);
    if (ctx->d_negmod) (void)hipFree(ctx->d_negmod);
    if (ctx->h_negmod) (void)hipHostFree(ctx->h_negmod);
    if (ctx->h_coef) (void)hipHostFree(ctx->h_coef);
    if (ctx->h_head) (void)hipHostFree(ctx->h_head);
    if (ctx->h_num) (void)hipHostFree(ctx->h_num);
    if (ctx->h_den) (void)hipHostFree(ctx->h_den);
    if (ctx->h_req) (void)hipHostFree(ctx->h_req);
    ctx->m_chi.release();
    ctx->t_resid.release();
    ctx->m_plan.release();
    ctx->h_plan.release();
    ctx->m_ctrl.release();
    ctx->h_ctrl.release();
    ctx->m_req.release();
    if (ctx->basis_start) (void)hipEventDestroy(ctx->basis_start);
    if (ctx->basis_end) (void)hipEventDestroy(ctx->basis_end);
    if (ctx->compute) (void)hipStreamDestroy(ctx->compute);
    if (ctx->copy) (void)hipStreamDestroy(ctx->copy);
    delete ctx;
}

const char *rps_last_error(const rps_ctx *ctx) {
    return ctx ? ctx->error.c_str() : g_open_error.c_str();
}

int rps_set_option(rps_ctx *ctx, const char *name, int64_t value) {
    if (!ctx || !name) {
        return RPS_EINVAL;
    }
    std::string key(name);
    if (key == "survivor_capacity") {
        if (value < 16384 || value > (int64_t(1) << 30)) {
            ctx->error = "survivor_capacity must be in 16384..2^30";
            return RPS_EINVAL;
        }
        ctx->capacity = static_cast<size_t>(value);
    } else if (key == "batch_size") {
        if (value < 1 || value > (1 << 20)) {
            ctx->error = "batch_size must be in 1..1048576";
            return RPS_EINVAL;
        }
        ctx->batch_size = static_cast<int>(value);
        ctx->batch_forced = true;
    } else if (key == "run_log2") {
        if (value != -1 && (value < 10 || value > 24)) {
            ctx->error = "run_log2 must be -1 or in 10..24";
            return RPS_EINVAL;
        }
        ctx->run_log2 = static_cast<int>(value);
    } else if (key == "variant") {
        if (value < -1 || value >= rps::variant_count()) {
            ctx->error = "variant out of range";
            return RPS_EINVAL;
        }
        ctx->variant = static_cast<int>(value);
    } else if (key == "direct_words") {
        if (value < 0 || value > static_cast<int64_t>(rps::kDirectMaxWords)) {
            ctx->error = "direct_words must be in 0.." + std::to_string(rps::kDirectMaxWords);
            return RPS_EINVAL;
        }
        ctx->direct_words = static_cast<unsigned long long>(value);
    } else if (key == "primitive_only") {
        if (value != 0 && value != 1) {
            ctx->error = "primitive_only must be 0 or 1";
            return RPS_EINVAL;
        }
        ctx->primitive = static_cast<int>(value);
    } else if (key == "table_slots") {
        if (value < 1 || value > (1 << 16)) {
            ctx->error = "table_slots must be in 1..65536";
            return RPS_EINVAL;
        }
        ctx->table_slots = value;
    } else if (key == "launch_timeout_ms") {
        if (value < 1 || value > (int64_t(1) << 40)) {
            ctx->error = "launch_timeout_ms must be in 1..2^40";
            return RPS_EINVAL;
        }
        ctx->launch_timeout_ms = value;
    } else if (key == "min_tiles") {
        if (value < 1 || value > (1 << 22)) {
            ctx->error = "min_tiles must be in 1..4194304";
            return RPS_EINVAL;
        }
        ctx->min_tiles = value;
    } else {
        ctx->error = "unknown option: " + key;
        return RPS_EINVAL;
    }
    return RPS_OK;
}

int rps_variant_count(void) { return rps::variant_count(); }

int rps_tuned_variant(uint64_t words) { return tuned_variant(words); }

int rps_tuned(uint64_t words, int many, int *variant, int *run_log2, int *batch) {
    const Tuned t = tuned(words, many != 0);
    if (variant) {
        *variant = t.variant;
    }
    if (run_log2) {
        *run_log2 = t.run_log2;
    }
    if (batch) {
        *batch = t.batch;
    }
    return RPS_OK;
}

int rps_variant_info(int index, int *words_per_lane, int *dense_primes, int *waves,
                     int *run_log2, int *wide) {
    if (index < 0 || index >= rps::variant_count()) {
        return RPS_EINVAL;
    }
    const rps::VariantInfo *v = rps::variant(index);
    if (words_per_lane) {
        *words_per_lane = v->words_per_lane;
    }
    if (dense_primes) {
        *dense_primes = v->fast_primes;
    }
    if (waves) {
        *waves = v->waves;
    }
    if (run_log2) {
        *run_log2 = v->run_log2;
    }
    if (wide) {
        *wide = v->wide;
    }
    return RPS_OK;
}

const char *rps_variant_name(int index) {
    const rps::VariantInfo *v = rps::variant(index);
    return v ? v->name : nullptr;
}

int rps_device_count(int *count) {
    if (!count) {
        return RPS_EINVAL;
    }
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        g_open_error = std::string("hipGetDeviceCount: ") + hipGetErrorString(e);
        *count = 0;
        return RPS_EDEVICE;
    }
    return RPS_OK;
}

int rps_device_info(int device, char *name, size_t name_len, char *pci,
                    size_t pci_len, int *compute_units, uint64_t *memory_bytes,
                    int *major, int *minor) {
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        g_open_error = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
        return RPS_EDEVICE;
    }
    if (name && name_len) {
        std::snprintf(name, name_len, "%s", prop.name[0] ? prop.name : prop.gcnArchName);
    }
    if (pci && pci_len) {
        char bus[64] = {0};
        if (hipDeviceGetPCIBusId(bus, sizeof(bus), device) != hipSuccess) {
            std::snprintf(bus, sizeof(bus), "unknown");
        }
        std::snprintf(pci, pci_len, "%s", bus);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (memory_bytes) *memory_bytes = prop.totalGlobalMem;
    if (major) *major = prop.major;
    if (minor) *minor = prop.minor;
    return RPS_OK;
}

int rps_last_kernel_timing(const rps_ctx *ctx, double *kernel_ms,
                           uint64_t *launches, double *sites) {
    if (!ctx) {
        return RPS_EINVAL;
    }
    if (kernel_ms) *kernel_ms = ctx->last_kernel_ms;
    if (launches) *launches = ctx->last_launches;
    if (sites) *sites = ctx->last_sites;
    return RPS_OK;
}

int rps_sieve(rps_ctx *ctx, const rps_request *req, rps_sink sink, void *user,
              rps_stats *stats) {
    return rps_sieve_batched(ctx, req, sink, nullptr, user, stats);
}

int rps_sieve_batched(rps_ctx *ctx, const rps_request *req, rps_sink sink,
                      rps_batch_done done, void *user, rps_stats *stats) {
    if (!ctx) {
        return RPS_EINVAL;
    }
    rps_stats local{};
    rps_stats &st = stats ? *stats : local;
    st = rps_stats{};
    ctx->error.clear();
    ctx->last_kernel_ms = 0.0;
    ctx->last_launches = 0;
    ctx->last_sites = 0.0;
    const auto t_plan = Clock::now();
    try {
        check_usable(ctx);
        if (!sink) {
            throw ApiError{RPS_EINVAL, "null sink"};
        }
        validate(req);
        RPS_HIP(hipSetDevice(ctx->device));
        const rps::LineDesc geometry = line_geometry(req);
        const Shape shape = launch_shape(ctx, geometry.words, false);
        const rps::VariantInfo *var = shape.var;
        const int stride = 64 * var->words_per_lane;
        const unsigned long long words = geometry.words;
        const long long A = req->numerator_bound;
        st.word_count = (2ULL * static_cast<unsigned long long>(A) + 32) / 32;
        for (int i = 0; i < rps::kCandidates; ++i) {
            const long long p = ctx->primes[i];
            ctx->h_negmod[i] = static_cast<int>(((-A) % p + p) % p);
        }
        const auto t_workspace = Clock::now();
        st.plan_ms = ms_between(t_plan, t_workspace);
        ensure_capacity(ctx);
        RPS_HIP(hipMemcpyAsync(ctx->d_negmod, ctx->h_negmod, rps::kCandidates * sizeof(int),
                               hipMemcpyHostToDevice, ctx->compute));
        const auto t_basis = Clock::now();
        st.workspace_ms = ms_between(t_workspace, t_basis);
        prepare_curve(ctx, req, &st);
        const auto t_setup = Clock::now();
        st.basis_wall_ms = ms_between(t_basis, t_setup);

        rps::SieveParams base{};
        base.line = geometry;
        base.line.resid = ctx->d_resid;
        base.line.neg_bound_mod = ctx->d_negmod;
        base.primes = ctx->d_primes;
        base.out.capacity = ctx->capacity;
        base.out.primitive = ctx->primitive;

        const long long last_t = req->den_last;
        long long next_t = req->den_first;
        Batch inflight[2];
        bool busy[2] = {false, false};
        const double line_sites = 2.0 * static_cast<double>(A) + 1.0;

        auto launch = [&](int si, const Batch &bt, bool resume) {
            Slot &s = ctx->slots[si];
            ensure_cursors(s, bt.tiles * std::max(var->waves, rps::kDirectWaves));
            rps::SieveParams prm = base;
            prm.line = bt.line;
            prm.resume = resume ? 1 : 0;
            prm.out.num = s.d_num;
            prm.out.den = s.d_den;
            prm.out.req = nullptr;
            prm.out.state = s.d_state;
            prm.cursors = s.d_cursors;
            RPS_HIP(rps::launch_reset(s.d_state, ctx->compute));
            RPS_HIP(hipEventRecord(s.k_start, ctx->compute));
            if (bt.line.direct) {
                RPS_HIP(rps::launch_direct(prm, static_cast<unsigned>(bt.tiles), ctx->compute));
            } else {
                RPS_HIP(var->launch(prm, static_cast<unsigned>(bt.tiles), ctx->compute));
            }
            RPS_HIP(hipEventRecord(s.k_end, ctx->compute));
            RPS_HIP(hipMemcpyAsync(s.h_state, s.d_state, sizeof(rps::DeviceState),
                                   hipMemcpyDeviceToHost, ctx->compute));
            RPS_HIP(hipEventRecord(s.ready, ctx->compute));
            ++st.launches;
            if (!resume) {
                ++st.batches;
                ctx->last_sites += line_sites * bt.line.n_t;
            }
        };
        auto next_batch = [&](Batch &bt) -> bool {
            if (next_t > last_t) {
                return false;
            }
            long long n = std::min<long long>(shape.batch, last_t - next_t + 1);
            bt.line = base.line;
            if (words <= ctx->direct_words) {
                direct_tiling(bt.line, std::max(1LL, n / std::max(1LL, ctx->min_tiles)));
            } else {
                unsigned long long chunks = line_chunks(ctx->min_tiles, n, words, stride,
                                                        var->waves);
                if (chunks > kMaxTilesPerLaunch) {
                    throw ApiError{RPS_EINVAL, "maximum height exceeds GPU limit"};
                }
                n = std::min<long long>(n, static_cast<long long>(kMaxTilesPerLaunch / chunks));
                chunks = line_chunks(ctx->min_tiles, n, words, stride, var->waves);
                row_tiling(bt.line, chunks, stride);
                row_runs(bt.line, std::max(1LL, n / std::max(1LL, ctx->min_tiles)),
                         shape.run_log2);
            }
            bt.line.first_t = static_cast<int>(next_t);
            bt.line.n_t = static_cast<int>(n);
            bt.tiles = tile_count(bt.line);
            next_t += n;
            return true;
        };
        bool cancelled = false;
        // Drains one finished launch of slot si into the sink; returns
        // whether the batch still needs a resume launch.
        auto drain = [&](int si) -> bool {
            Slot &s = ctx->slots[si];
            wait_event(s.ready, ctx->launch_timeout_ms,
                       "sieve launch for denominators " + std::to_string(inflight[si].line.first_t)
                           + ".." + std::to_string(inflight[si].line.first_t
                                                   + inflight[si].line.n_t - 1));
            float kms = 0.0f;
            RPS_HIP(hipEventElapsedTime(&kms, s.k_start, s.k_end));
            st.sieve_ms += kms;
            const rps::DeviceState state = *s.h_state;
            const unsigned long long valid = std::min(state.count, state.first_fail);
            if (valid > ctx->capacity) {
                throw ApiError{RPS_EINTERNAL, "streaming survivor count exceeds capacity"};
            }
            if (valid > ~0ULL - st.survivor_count) {
                throw ApiError{RPS_EOVERFLOW, "total survivor count overflow"};
            }
            if (state.full && valid == 0 && state.first_fail == 0) {
                throw ApiError{RPS_EINTERNAL, "survivor buffer cannot hold one wave iteration"};
            }
            for (unsigned long long at = 0; at < valid && !cancelled;) {
                const size_t n = static_cast<size_t>(std::min<unsigned long long>(
                    valid - at, ctx->stage_capacity));
                RPS_HIP(hipMemcpyAsync(ctx->h_num, s.d_num + at, n * sizeof(long long),
                                       hipMemcpyDeviceToHost, ctx->copy));
                RPS_HIP(hipMemcpyAsync(ctx->h_den, s.d_den + at, n * sizeof(int),
                                       hipMemcpyDeviceToHost, ctx->copy));
                RPS_HIP(hipStreamSynchronize(ctx->copy));
                const auto v0 = Clock::now();
                const int rc = sink(user, reinterpret_cast<const int64_t *>(ctx->h_num),
                                    reinterpret_cast<const int32_t *>(ctx->h_den), n);
                st.verification_ms += ms_between(v0, Clock::now());
                st.survivor_count += n;
                at += n;
                if (rc != 0) {
                    cancelled = true;
                }
            }
            return state.full != 0;
        };

        const auto t_sieve = Clock::now();
        st.survivor_setup_ms = ms_between(t_setup, t_sieve);
        for (int si = 0; si < 2; ++si) {
            if (next_batch(inflight[si])) {
                launch(si, inflight[si], false);
                busy[si] = true;
            }
        }
        int cur = 0;
        while (busy[cur] && !cancelled) {
            const uint64_t before = st.survivor_count;
            bool more = drain(cur);
            while (more && !cancelled) {
                launch(cur, inflight[cur], true);
                more = drain(cur);
            }
            if (cancelled) {
                break;
            }
            if (done) {
                const rps::LineDesc &bl = inflight[cur].line;
                if (done(user, bl.first_t, bl.first_t + bl.n_t - 1,
                         st.survivor_count - before) != 0) {
                    cancelled = true;
                    break;
                }
            }
            busy[cur] = next_batch(inflight[cur]);
            if (busy[cur]) {
                launch(cur, inflight[cur], false);
            }
            cur ^= 1;
        }
        RPS_HIP(hipStreamSynchronize(ctx->compute));
        RPS_HIP(hipStreamSynchronize(ctx->copy));
        st.sieve_wall_ms = ms_between(t_sieve, Clock::now()) - st.verification_ms;
        st.initial_mask_bytes = static_cast<double>(
            static_cast<long long>(req->den_last) - req->den_first + 1)
            * static_cast<double>(st.word_count) * 14.0 * 4.0;
        ctx->last_kernel_ms = st.sieve_ms;
        ctx->last_launches = st.launches;
        return cancelled ? RPS_ECANCELLED : RPS_OK;
    } catch (const HipError &e) {
        return fail(ctx, e);
    } catch (const TimedOut &e) {
        return fail(ctx, e);
    } catch (const ApiError &e) {
        return fail(ctx, e);
    } catch (const std::bad_alloc &) {
        ctx->error = "out of host memory";
        return RPS_ENOMEM;
    } catch (const std::exception &e) {
        // Nothing may unwind through the C ABI.
        ctx->error = std::string("internal error: ") + e.what();
        (void)hipStreamSynchronize(ctx->compute);
        (void)hipStreamSynchronize(ctx->copy);
        return RPS_EINTERNAL;
    }
}

int rps_sieve_into(rps_ctx *ctx, const rps_request *req, int64_t *numerators,
                   int32_t *denominators, uint64_t capacity, uint64_t *survivor_count) {
    if (!ctx) {
        return RPS_EINVAL;
    }
    ctx->error.clear();
    ctx->last_kernel_ms = 0.0;
    ctx->last_launches = 0;
    ctx->last_sites = 0.0;
    if (survivor_count) {
        *survivor_count = 0;
    }
    try {
        check_usable(ctx);
        if (!numerators || !denominators || !survivor_count) {
            throw ApiError{RPS_EINVAL, "null output pointer"};
        }
        validate(req);
        RPS_HIP(hipSetDevice(ctx->device));
        unsigned char *scratch = static_cast<unsigned char *>(rps::scratch_address(ctx->device));
        if (!scratch) {
            throw ApiError{RPS_EINTERNAL, "device scratch not initialised"};
        }
        namespace S = rps::scratch_layout;
        auto *resid = reinterpret_cast<unsigned long long *>(scratch + S::kResid);
        auto *chi = reinterpret_cast<unsigned long long *>(scratch + S::kChi);
        auto *coef = reinterpret_cast<int *>(scratch + S::kCoef);
        auto *negmod = reinterpret_cast<int *>(scratch + S::kNegMod);
        auto *state = reinterpret_cast<rps::DeviceState *>(scratch + S::kState);
        auto *cursors = reinterpret_cast<unsigned long long *>(scratch + S::kCursors);

        const rps::LineDesc geometry = line_geometry(req);
        const Shape shape = launch_shape(ctx, geometry.words, false);
        const rps::VariantInfo *var = shape.var;
        const int stride = 64 * var->words_per_lane;
        const int waves = std::max(var->waves, rps::kDirectWaves);
        if (static_cast<unsigned long long>(waves) > S::kMaxWaves) {
            throw ApiError{RPS_EINTERNAL, "variant has more waves than the scratch cursors hold"};
        }
        const unsigned long long words = geometry.words;
        // Every launch must fit the scratch's cursors: a line alone is cut
        // into at least ceil(words / 2^25) tiles.
        if ((words + (1ULL << 25) - 1) >> 25 > S::kMaxTiles) {
            throw ApiError{RPS_EINVAL, "maximum height exceeds the allocation-free limit "
                                       "(lines of at most 2^39 words)"};
        }

        // One curve's tables into the scratch: coefficient residues and
        // (-A) mod p staged in the context's pinned buffers.
        const long long A = req->numerator_bound;
        coefficient_residues(req, ctx->primes, req->n_coefficients, ctx->h_coef);
        for (int i = 0; i < rps::kCandidates; ++i) {
            const long long p = ctx->primes[i];
            ctx->h_negmod[i] = static_cast<int>(((-A) % p + p) % p);
        }
        RPS_HIP(hipMemcpyAsync(coef, ctx->h_coef,
                               static_cast<size_t>(rps::kCandidates) * req->n_coefficients
                                   * sizeof(int),
                               hipMemcpyHostToDevice, ctx->compute));
        RPS_HIP(hipMemcpyAsync(negmod, ctx->h_negmod, rps::kCandidates * sizeof(int),
                               hipMemcpyHostToDevice, ctx->compute));
        RPS_HIP(rps::launch_chi(coef, req->n_coefficients, 1, ctx->d_primes, chi, ctx->compute));
        RPS_HIP(rps::launch_resid(ctx->d_primes, ctx->d_inverses, chi, 1, nullptr, resid,
                                  ctx->compute));

        // All batches reserve from one flush state, reset once: survivors go
        // to [0, count) of the caller's buffers; a reservation that does not
        // fit is recorded and every later one fails too, so the written
        // survivors stay exactly [0, min(count, first_fail)).
        rps::SieveParams base{};
        base.line = geometry;
        base.line.resid = resid;
        base.line.neg_bound_mod = negmod;
        base.primes = ctx->d_primes;
        base.out.num = reinterpret_cast<long long *>(numerators);
        base.out.den = denominators;
        base.out.req = nullptr;
        base.out.state = state;
        base.out.capacity = capacity;
        base.out.primitive = ctx->primitive;
        base.cursors = cursors;
        base.resume = 0;
        RPS_HIP(rps::launch_reset(state, ctx->compute));
        const double line_sites = 2.0 * static_cast<double>(A) + 1.0;
        for (long long next_t = req->den_first; next_t <= req->den_last;) {
            long long n = std::min<long long>(shape.batch, req->den_last - next_t + 1);
            n = std::min<long long>(n, static_cast<long long>(S::kMaxTiles));
            rps::SieveParams prm = base;
            if (words <= ctx->direct_words) {
                direct_tiling(prm.line, std::max(1LL, n / std::max(1LL, ctx->min_tiles)));
            } else {
                unsigned long long chunks = line_chunks(ctx->min_tiles, n, words, stride,
                                                        var->waves);
                n = std::min<long long>(n, static_cast<long long>(S::kMaxTiles / chunks));
                chunks = line_chunks(ctx->min_tiles, n, words, stride, var->waves);
                row_tiling(prm.line, chunks, stride);
                row_runs(prm.line, std::max(1LL, n / std::max(1LL, ctx->min_tiles)),
                         shape.run_log2);
            }
            prm.line.first_t = static_cast<int>(next_t);
            prm.line.n_t = static_cast<int>(n);
            const size_t tiles = tile_count(prm.line);
            if (tiles > S::kMaxTiles) {
                throw ApiError{RPS_EINTERNAL, "batch exceeds the scratch's tile limit"};
            }
            if (prm.line.direct) {
                RPS_HIP(rps::launch_direct(prm, static_cast<unsigned>(tiles), ctx->compute));
            } else {
                RPS_HIP(var->launch(prm, static_cast<unsigned>(tiles), ctx->compute));
            }
            ++ctx->last_launches;
            ctx->last_sites += line_sites * static_cast<double>(n);
            next_t += n;
        }
        rps::DeviceState *h_state = ctx->slots[0].h_state;
        RPS_HIP(hipMemcpyAsync(h_state, state, sizeof(rps::DeviceState), hipMemcpyDeviceToHost,
                               ctx->compute));
        RPS_HIP(hipEventRecord(ctx->slots[0].ready, ctx->compute));
        wait_event(ctx->slots[0].ready, ctx->launch_timeout_ms,
                   "rps_sieve_into's " + std::to_string(ctx->last_launches) + " launches");
        const unsigned long long valid = std::min(h_state->count, h_state->first_fail);
        *survivor_count = std::min<unsigned long long>(valid, capacity);
        if (h_state->full) {
            throw ApiError{RPS_EOVERFLOW, "survivors exceed the output capacity of "
                                          + std::to_string(capacity) + " entries"};
        }
        return RPS_OK;
    } catch (const HipError &e) {
        return fail(ctx, e);
    } catch (const TimedOut &e) {
        return fail(ctx, e);
    } catch (const ApiError &e) {
        return fail(ctx, e);
    } catch (const std::exception &e) {
        ctx->error = std::string("internal error: ") + e.what();
        (void)hipStreamSynchronize(ctx->compute);
        return RPS_EINTERNAL;
    }
}

int rps_check_request(const rps_request *request, char *message, size_t message_len) {
    std::string reason;
    int rc = RPS_OK;
    try {
        validate(request);
    } catch (const ApiError &e) {
        rc = e.code;
        reason = e.message;
    }
    if (message && message_len) {
        std::snprintf(message, message_len, "%s", reason.c_str());
    }
    return rc;
}

}  // extern "C"

namespace {

// Survivors of the plain path handed to a many-curve sink as request 0.
struct SingleSink {
    rps_multi_sink sink;
    void *user;
    std::vector<int32_t> zeros;

    static int call(void *u, const int64_t *num, const int32_t *den, size_t count) {
        SingleSink &s = *static_cast<SingleSink *>(u);
        if (s.zeros.size() < count) {
            s.zeros.assign(count, 0);
        }
        return s.sink(s.user, s.zeros.data(), num, den, count);
    }
};

// One persistent launch over all the requests, with `forced` (>= 0) or the
// tuning table's variant for the request with the most sites.  A single
// request takes the plain path instead: the same survivors, from a grid
// and tiles sized for that one curve.
// The many-curve launch's short-line crossover.  There the row tiles stay
// ahead of the short-line tiles above about 1,152 words: on the 4,096
// batched curves, short-line tiles up to 2,048 words made the launches of
// the 11- and 12-row groups 850 and 488 ms against 812 and 449.
constexpr unsigned long long kManyDirectWords = 1152;

int sieve_many_launch(rps_ctx *ctx, const rps_request *requests, size_t n_requests,
                      rps_multi_sink sink, void *user, rps_stats *stats, int forced) {
    if (n_requests == 1 && requests && sink && forced < 0) {
        SingleSink single{sink, user, {}};
        const int rc = rps_sieve_batched(ctx, requests, SingleSink::call, nullptr, &single,
                                         stats);
        if (stats) {
            stats->batches = 1;
        }
        return rc;
    }
    rps_stats local{};
    rps_stats &st = stats ? *stats : local;
    st = rps_stats{};
    ctx->error.clear();
    ctx->last_kernel_ms = 0.0;
    ctx->last_launches = 0;
    ctx->last_sites = 0.0;
    const auto t_plan = Clock::now();
    try {
        check_usable(ctx);
        if (!sink || (!requests && n_requests)) {
            throw ApiError{RPS_EINVAL, "null requests or sink"};
        }
        if (n_requests > (size_t(1) << 20)) {
            throw ApiError{RPS_EINVAL, "too many requests in one call (at most 2^20)"};
        }
        if (n_requests == 0) {
            return RPS_OK;
        }
        const int n = static_cast<int>(n_requests);
        int n_coef = 2;
        for (int r = 0; r < n; ++r) {
            validate(&requests[r]);
            n_coef = std::max(n_coef, requests[r].n_coefficients);
        }
        RPS_HIP(hipSetDevice(ctx->device));
        // One variant and run length per launch: the options' or the forced
        // variant, else the tuning table's many-curve choice for the request
        // with the most sites.
        int largest = 0;
        double most = -1.0;
        for (int r = 0; r < n; ++r) {
            const rps::LineDesc g = line_geometry(&requests[r]);
            const double work = (2.0 * static_cast<double>(g.bound) + 1.0)
                * (static_cast<double>(requests[r].den_last) - requests[r].den_first + 1.0);
            if (work > most) {
                most = work;
                largest = r;
            }
        }
        const Shape shape = launch_shape(ctx, line_geometry(&requests[largest]).words, true,
                                         forced);
        const rps::VariantInfo *var = shape.var;
        const int stride = 64 * var->words_per_lane;

        // Geometry and tiling per curve.  Long lines are cut into chunks of
        // at most 2^18 words so one curve's tiles stay small next to the rest
        // of the launch; lines of more than 2^40 words take longer chunks
        // (at most 2^22 per line, of at most 2^25 words, as on the plain
        // path).
        std::vector<rps::LineDesc> lines(n);
        std::vector<double> cost(n);
        double sites = 0.0;
        for (int r = 0; r < n; ++r) {
            rps::LineDesc L = line_geometry(&requests[r]);
            L.first_t = requests[r].den_first;
            L.n_t = requests[r].den_last - requests[r].den_first + 1;
            if (L.words <= std::min(ctx->direct_words, kManyDirectWords)) {
                direct_tiling(L, L.n_t);
            } else {
                const unsigned long long chunks = std::min(
                    std::max((L.words + (1ULL << 18) - 1) >> 18, 1ULL), kMaxTilesPerLaunch);
                if ((L.words + chunks - 1) / chunks > (1ULL << 25)) {
                    throw ApiError{RPS_EINVAL, "maximum height exceeds GPU limit"};
                }
                row_tiling(L, chunks, stride);
                row_runs(L, L.n_t, shape.run_log2);
            }
            lines[r] = L;
            const double line_sites = 2.0 * static_cast<double>(L.bound) + 1.0;
            sites += line_sites * L.n_t;
            cost[r] = (line_sites + 150000.0) * L.n_t;  // + per-line setup
        }
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return cost[a] > cost[b]; });

        // Residue tables: one per distinct curve, kept between calls.  Hits
        // are marked first; misses then take the least recently used slots
        // not needed by this call.
        ++ctx->t_call;
        std::vector<std::vector<std::string>> keys(n);
        std::map<std::vector<std::string>, int> distinct;
        for (int k = 0; k < n; ++k) {
            const rps_request &q = requests[order[k]];
            keys[k].assign(q.coefficients, q.coefficients + q.n_coefficients);
            distinct.emplace(keys[k], 0);
        }
        const int cap = static_cast<int>(std::max<long long>(
            ctx->table_slots, static_cast<long long>(distinct.size())));
        if (cap > ctx->t_cap) {
            // Empty the cache before the old tables go: if the larger
            // allocation fails, no slot may point into freed memory.
            ctx->t_cap = 0;
            ctx->t_index.clear();
            ctx->t_key.clear();
            ctx->t_used.clear();
            ctx->t_resid.release();
            ctx->t_resid.ensure(static_cast<size_t>(cap) * rps::kResidTableWords);
            ctx->t_cap = cap;
            ctx->t_key.assign(cap, {});
            ctx->t_used.assign(cap, 0);
        }
        for (auto &d : distinct) {
            auto it = ctx->t_index.find(d.first);
            d.second = -1;
            if (it != ctx->t_index.end()) {
                d.second = it->second;
                ctx->t_used[it->second] = ctx->t_call;
            }
        }
        std::vector<int> spare;
        for (int slot = 0; slot < ctx->t_cap; ++slot) {
            if (ctx->t_used[slot] != ctx->t_call) {
                spare.push_back(slot);
            }
        }
        std::stable_sort(spare.begin(), spare.end(), [&](int a, int b) {
            return ctx->t_used[a] < ctx->t_used[b];
        });
        std::vector<int> slot_of(n), build_k, build_slot;
        size_t next_spare = 0;
        for (int k = 0; k < n; ++k) {
            int &slot = distinct[keys[k]];
            if (slot < 0) {
                slot = spare.at(next_spare++);
                if (!ctx->t_key[slot].empty()) {
                    ctx->t_index.erase(ctx->t_key[slot]);
                }
                ctx->t_key[slot] = keys[k];
                ctx->t_index[keys[k]] = slot;
                ctx->t_used[slot] = ctx->t_call;
                build_k.push_back(k);
                build_slot.push_back(slot);
            }
            slot_of[k] = slot;
        }
        const int n_build = static_cast<int>(build_k.size());
        // Until the new tables are built, a failure must not leave them in
        // the cache.
        struct TableGuard {
            rps_ctx *ctx;
            bool built = false;
            ~TableGuard() {
                if (!built) {
                    ctx->t_index.clear();
                    ctx->t_key.assign(ctx->t_cap, {});
                    ctx->t_used.assign(ctx->t_cap, 0);
                }
            }
        } guard{ctx};

        ctx->m_chi.ensure_grow(static_cast<size_t>(n_build) * rps::kCandidates * rps::kChiWords,
                               size_t(64) * rps::kCandidates * rps::kChiWords);
        ctx->m_req.ensure(ctx->capacity);
        ensure_capacity(ctx);
        std::vector<rps::CurveDesc> curves(n);
        std::vector<unsigned long long> tile_start(n + 1, 0);
        for (int k = 0; k < n; ++k) {
            rps::CurveDesc &C = curves[k];
            C.line = lines[order[k]];
            C.request = order[k];
            C.tile_start = tile_start[k];
            if (__builtin_add_overflow(tile_start[k], tile_count(C.line), &tile_start[k + 1])
                || tile_start[k + 1] > (1ULL << 62)) {
                throw ApiError{RPS_EINVAL, "too much work in one call"};
            }
        }
        const unsigned long long total_tiles = tile_start[n];
        // Curve of every 2^shift-th tile (at most 2^20 buckets), so the
        // kernel's search for a tile's curve spans only the curves of one
        // bucket.
        int bucket_shift = 0;
        while ((total_tiles >> bucket_shift) > (1ULL << 20)) {
            ++bucket_shift;
        }
        const size_t n_buckets = static_cast<size_t>(
            ((total_tiles + (1ULL << bucket_shift) - 1) >> bucket_shift) + 1);

        // The plan goes to the device in one copy from pinned memory.
        const size_t coef_ints = static_cast<size_t>(n_build) * rps::kCandidates * n_coef;
        const PlanLayout plan(coef_ints, static_cast<size_t>(n_build), static_cast<size_t>(n),
                              n_buckets);
        ctx->m_plan.ensure_grow(plan.bytes, size_t(1) << 20);
        ctx->h_plan.ensure_grow(plan.bytes, size_t(1) << 20);
        ctx->m_ctrl.ensure_grow(ControlLayout::bytes(static_cast<size_t>(n)), size_t(1) << 16);
        ctx->h_ctrl.ensure_grow(ControlLayout::bytes(static_cast<size_t>(n)), size_t(1) << 16);
        unsigned char *hp = ctx->h_plan.ptr;
        unsigned char *dp = ctx->m_plan.ptr;
        int *h_coef = reinterpret_cast<int *>(hp + plan.coef);
        for (int j = 0; j < n_build; ++j) {
            coefficient_residues(&requests[order[build_k[j]]], ctx->primes, n_coef,
                                 h_coef + static_cast<size_t>(j) * rps::kCandidates * n_coef);
        }
        std::copy(build_slot.begin(), build_slot.end(), reinterpret_cast<int *>(hp + plan.slot));
        int *h_negmod = reinterpret_cast<int *>(hp + plan.negmod);
        int *d_negmod = reinterpret_cast<int *>(dp + plan.negmod);
        for (int k = 0; k < n; ++k) {
            const int r = order[k];
            for (int i = 0; i < rps::kCandidates; ++i) {
                const long long p = ctx->primes[i];
                h_negmod[static_cast<size_t>(k) * rps::kCandidates + i] =
                    static_cast<int>(((-requests[r].numerator_bound) % p + p) % p);
            }
            curves[k].line.resid = ctx->t_resid.ptr
                + static_cast<size_t>(slot_of[k]) * rps::kResidTableWords;
            curves[k].line.neg_bound_mod = d_negmod + static_cast<size_t>(k) * rps::kCandidates;
        }
        std::memcpy(hp + plan.curves, curves.data(), n * sizeof(rps::CurveDesc));
        std::memcpy(hp + plan.tile_start, tile_start.data(),
                    (n + 1) * sizeof(unsigned long long));
        int *h_bucket = reinterpret_cast<int *>(hp + plan.bucket);
        for (size_t g = 0, k = 0; g < n_buckets; ++g) {
            const unsigned long long first = std::min<unsigned long long>(
                static_cast<unsigned long long>(g) << bucket_shift,
                total_tiles ? total_tiles - 1 : 0);
            while (k + 1 < static_cast<size_t>(n) && tile_start[k + 1] <= first) {
                ++k;
            }
            h_bucket[g] = static_cast<int>(k);
        }
        const auto t_workspace = Clock::now();
        st.plan_ms = ms_between(t_plan, t_workspace);
        RPS_HIP(hipMemcpyAsync(dp, hp, plan.bytes, hipMemcpyHostToDevice, ctx->compute));
        unsigned char *ctrl = ctx->m_ctrl.ptr;
        rps::DeviceState *d_state = reinterpret_cast<rps::DeviceState *>(ctrl + ControlLayout::kState);
        auto *d_head = reinterpret_cast<unsigned long long *>(ctrl + ControlLayout::kHead);
        int *d_failed = reinterpret_cast<int *>(ctrl + ControlLayout::kFailed);
        RPS_HIP(rps::launch_reset_many(d_state, d_head, d_failed, n, ctx->compute));
        const auto t_basis = Clock::now();
        st.workspace_ms = ms_between(t_workspace, t_basis);
        RPS_HIP(hipEventRecord(ctx->basis_start, ctx->compute));
        if (n_build) {
            RPS_HIP(rps::launch_chi(reinterpret_cast<const int *>(dp + plan.coef), n_coef, n_build,
                                    ctx->d_primes, ctx->m_chi.ptr, ctx->compute));
            RPS_HIP(rps::launch_resid(ctx->d_primes, ctx->d_inverses, ctx->m_chi.ptr, n_build,
                                      reinterpret_cast<const int *>(dp + plan.slot),
                                      ctx->t_resid.ptr, ctx->compute));
        }
        RPS_HIP(hipEventRecord(ctx->basis_end, ctx->compute));
        const auto t_sieve = Clock::now();
        st.basis_wall_ms = ms_between(t_basis, t_sieve);

        // Persistent launches until every tile was handed out.  Each launch's
        // control block and first survivors come back in the same stream,
        // so a launch costs one wait.  Survivors are kept per request on the
        // host until the end, because a request whose reservation fails is
        // searched again on the plain path.
        Slot &s = ctx->slots[0];
        rps::MultiParams M{};
        M.curves = reinterpret_cast<const rps::CurveDesc *>(dp + plan.curves);
        M.tile_start = reinterpret_cast<const unsigned long long *>(dp + plan.tile_start);
        M.bucket = reinterpret_cast<const int *>(dp + plan.bucket);
        M.bucket_shift = bucket_shift;
        M.curve_failed = d_failed;
        M.head = d_head;
        M.total_tiles = total_tiles;
        M.out.num = s.d_num;
        M.out.den = s.d_den;
        M.out.req = ctx->m_req.ptr;
        M.out.state = d_state;
        M.out.capacity = ctx->capacity;
        M.out.primitive = ctx->primitive;
        M.primes = ctx->d_primes;
        M.n_curves = n;
        // One workgroup per resident slot: the variant's occupancy per CU.
        int per_cu = 0;
        RPS_HIP(var->multi_blocks_per_cu(&per_cu));
        // No more workgroups than tiles: a small call would otherwise start
        // every resident slot only to find the tiles handed out.
        const unsigned grid = static_cast<unsigned>(std::min<unsigned long long>(
            static_cast<unsigned long long>(ctx->cu_count) * std::max(1, per_cu),
            std::max(1ULL, total_tiles)));
        const size_t spec = std::min({kSpeculativeSurvivors, ctx->stage_capacity, ctx->capacity});
        const size_t ctrl_bytes = ControlLayout::bytes(static_cast<size_t>(n));
        std::vector<std::vector<long long>> kept_num(n);
        std::vector<std::vector<int>> kept_den(n);
        auto keep = [&](size_t k) {
            for (size_t i = 0; i < k; ++i) {
                const int r = ctx->h_req[i];
                if (r < 0 || r >= n) {
                    throw ApiError{RPS_EINTERNAL, "survivor with an invalid request index"};
                }
                kept_num[r].push_back(ctx->h_num[i]);
                kept_den[r].push_back(ctx->h_den[i]);
            }
        };
        unsigned long long head = 0;
        while (head < total_tiles) {
            if (st.launches) {
                RPS_HIP(rps::launch_reset(d_state, ctx->compute));
            }
            RPS_HIP(hipEventRecord(s.k_start, ctx->compute));
            RPS_HIP(var->launch_multi(M, grid, ctx->compute));
            RPS_HIP(hipEventRecord(s.k_end, ctx->compute));
            RPS_HIP(hipMemcpyAsync(ctx->h_ctrl.ptr, ctrl, ctrl_bytes, hipMemcpyDeviceToHost,
                                   ctx->compute));
            RPS_HIP(hipMemcpyAsync(ctx->h_num, s.d_num, spec * sizeof(long long),
                                   hipMemcpyDeviceToHost, ctx->compute));
            RPS_HIP(hipMemcpyAsync(ctx->h_den, s.d_den, spec * sizeof(int),
                                   hipMemcpyDeviceToHost, ctx->compute));
            RPS_HIP(hipMemcpyAsync(ctx->h_req, ctx->m_req.ptr, spec * sizeof(int),
                                   hipMemcpyDeviceToHost, ctx->compute));
            RPS_HIP(hipEventRecord(s.ready, ctx->compute));
            wait_event(s.ready, ctx->launch_timeout_ms,
                       "many-curve launch " + std::to_string(st.launches + 1) + " (tiles "
                           + std::to_string(head) + ".." + std::to_string(total_tiles)
                           + " of " + std::to_string(n) + " requests)");
            if (!st.launches) {
                guard.built = true;
                float bms = 0.0f;
                RPS_HIP(hipEventElapsedTime(&bms, ctx->basis_start, ctx->basis_end));
                st.basis_ms = bms;
            }
            std::memcpy(&head, ctx->h_ctrl.ptr + ControlLayout::kHead, sizeof head);
            ++st.launches;
            float kms = 0.0f;
            RPS_HIP(hipEventElapsedTime(&kms, s.k_start, s.k_end));
            st.sieve_ms += kms;
            rps::DeviceState state;
            std::memcpy(&state, ctx->h_ctrl.ptr + ControlLayout::kState, sizeof state);
            const unsigned long long valid = std::min(state.count, state.first_fail);
            if (valid > ctx->capacity) {
                throw ApiError{RPS_EINTERNAL, "streaming survivor count exceeds capacity"};
            }
            keep(static_cast<size_t>(std::min<unsigned long long>(valid, spec)));
            for (unsigned long long at = spec; at < valid;) {
                const size_t k = static_cast<size_t>(std::min<unsigned long long>(
                    valid - at, ctx->stage_capacity));
                RPS_HIP(hipMemcpyAsync(ctx->h_num, s.d_num + at, k * sizeof(long long),
                                       hipMemcpyDeviceToHost, ctx->copy));
                RPS_HIP(hipMemcpyAsync(ctx->h_den, s.d_den + at, k * sizeof(int),
                                       hipMemcpyDeviceToHost, ctx->copy));
                RPS_HIP(hipMemcpyAsync(ctx->h_req, ctx->m_req.ptr + at, k * sizeof(int),
                                       hipMemcpyDeviceToHost, ctx->copy));
                RPS_HIP(hipStreamSynchronize(ctx->copy));
                keep(k);
                at += k;
            }
            if (head > total_tiles) {
                head = total_tiles;
            }
        }
        std::vector<int> failed(n);
        std::memcpy(failed.data(), ctx->h_ctrl.ptr + ControlLayout::kFailed, n * sizeof(int));
        const double multi_kernel_ms = st.sieve_ms;
        const uint64_t multi_launches = st.launches;

        // Deliver in request order; a request that failed is searched again
        // alone on the plain path, whose per-wave cursors resume exactly.
        std::vector<char> is_failed(n, 0);
        for (int k = 0; k < n; ++k) {
            if (failed[k]) {
                is_failed[curves[k].request] = 1;
            }
        }
        bool cancelled = false;
        for (int r = 0; r < n && !cancelled; ++r) {
            if (is_failed[r]) {
                struct Fwd {
                    rps_multi_sink sink;
                    void *user;
                    int request;
                    std::vector<int> req;
                    double ms;
                } fwd{sink, user, r, {}, 0.0};
                auto forward = [](void *u, const int64_t *num, const int32_t *den,
                                  size_t count) -> int {
                    Fwd &f = *static_cast<Fwd *>(u);
                    f.req.assign(count, f.request);
                    const auto v0 = Clock::now();
                    const int rc = f.sink(f.user, f.req.data(), num, den, count);
                    f.ms += ms_between(v0, Clock::now());
                    return rc;
                };
                rps_stats one{};
                const int rc = rps_sieve(ctx, &requests[r], forward, &fwd, &one);
                st.survivor_count += one.survivor_count;
                st.launches += one.launches;
                st.sieve_ms += one.sieve_ms;
                st.verification_ms += fwd.ms;
                ++st.requeued;
                if (rc == RPS_ECANCELLED) {
                    cancelled = true;
                } else if (rc != RPS_OK) {
                    return rc;
                }
                continue;
            }
            const size_t count = kept_num[r].size();
            for (size_t at = 0; at < count && !cancelled;) {
                const size_t k = std::min(count - at, ctx->stage_capacity);
                for (size_t i = 0; i < k; ++i) {
                    ctx->h_req[i] = r;
                }
                const auto v0 = Clock::now();
                const int rc = sink(user, ctx->h_req,
                                    reinterpret_cast<const int64_t *>(kept_num[r].data() + at),
                                    reinterpret_cast<const int32_t *>(kept_den[r].data() + at), k);
                st.verification_ms += ms_between(v0, Clock::now());
                st.survivor_count += k;
                at += k;
                if (rc != 0) {
                    cancelled = true;
                }
            }
        }
        st.batches = static_cast<uint64_t>(n);
        st.sieve_wall_ms = ms_between(t_sieve, Clock::now()) - st.verification_ms;
        ctx->last_kernel_ms = multi_kernel_ms;
        ctx->last_launches = multi_launches;
        ctx->last_sites = sites;
        return cancelled ? RPS_ECANCELLED : RPS_OK;
    } catch (const HipError &e) {
        return fail(ctx, e);
    } catch (const TimedOut &e) {
        return fail(ctx, e);
    } catch (const ApiError &e) {
        return fail(ctx, e);
    } catch (const std::bad_alloc &) {
        ctx->error = "out of host memory";
        return RPS_ENOMEM;
    } catch (const std::exception &e) {
        // Nothing may unwind through the C ABI.
        ctx->error = std::string("internal error: ") + e.what();
        (void)hipStreamSynchronize(ctx->compute);
        (void)hipStreamSynchronize(ctx->copy);
        return RPS_EINTERNAL;
    }
}

// Survivor sink of one group of requests: maps the group's request indices
// back to the caller's.
struct GroupSink {
    rps_multi_sink sink;
    void *user;
    const std::vector<int32_t> *index;
    std::vector<int32_t> mapped;

    static int call(void *u, const int32_t *req, const int64_t *num, const int32_t *den,
                    size_t count) {
        GroupSink &g = *static_cast<GroupSink *>(u);
        g.mapped.resize(count);
        for (size_t i = 0; i < count; ++i) {
            g.mapped[i] = (*g.index)[static_cast<size_t>(req[i])];
        }
        return g.sink(g.user, g.mapped.data(), num, den, count);
    }
};

void add_stats(rps_stats &into, const rps_stats &s) {
    into.survivor_count += s.survivor_count;
    into.basis_ms += s.basis_ms;
    into.sieve_ms += s.sieve_ms;
    into.plan_ms += s.plan_ms;
    into.workspace_ms += s.workspace_ms;
    into.basis_wall_ms += s.basis_wall_ms;
    into.survivor_setup_ms += s.survivor_setup_ms;
    into.verification_ms += s.verification_ms;
    into.sieve_wall_ms += s.sieve_wall_ms;
    into.launches += s.launches;
    into.batches += s.batches;
    into.requeued += s.requeued;
}

// Requests per launch at most: the launch keeps one residue table per
// distinct curve on the device (1.33 MB each, 10.9 GB for this many).
constexpr size_t kMaxCurvesPerLaunch = 8192;

// Sites below which a group of requests wanting one variant does not get a
// launch of its own (about 5 ms of kernel time on mid-length lines).
constexpr double kOwnLaunchSites = 1.5e11;

// With the variant left to the tuning table, requests are grouped by the
// many-curve column's choice for their line length, and each group runs in
// launches of its own (a launch has one LDS layout and one tile shape: the
// wide tile would cost short lines half their occupancy, and a narrow tile
// with 14 dense rows builds more per line than short lines repay).  A group
// with fewer than kOwnLaunchSites sites joins the group of its tile family
// (wide or not) with the most sites, so that a small call (a serving round)
// is not cut into many launches.  Each group is cut into launches of at
// most kMaxCurvesPerLaunch requests.
int sieve_many_split(rps_ctx *ctx, const rps_request *requests, size_t n_requests,
                     rps_multi_sink sink, void *user, rps_stats *stats) {
    if (n_requests < 2 || !requests || !sink || n_requests > (size_t(1) << 20)) {
        return sieve_many_launch(ctx, requests, n_requests, sink, user, stats, -1);
    }
    const int n_variants = rps::variant_count();
    std::vector<std::vector<int32_t>> by_variant(static_cast<size_t>(n_variants));
    std::vector<double> group_sites(static_cast<size_t>(n_variants), 0.0);
    try {
        for (size_t r = 0; r < n_requests; ++r) {
            validate(&requests[r]);
            int v = 0;
            if (ctx->variant < 0) {
                const rps::LineDesc L = line_geometry(&requests[r]);
                v = tuned(L.words, true).variant;
                group_sites[static_cast<size_t>(v)] +=
                    (2.0 * static_cast<double>(L.bound) + 1.0)
                    * (static_cast<double>(requests[r].den_last) - requests[r].den_first + 1.0);
            }
            by_variant[static_cast<size_t>(v)].push_back(static_cast<int32_t>(r));
        }
    } catch (const ApiError &e) {
        ctx->last_kernel_ms = 0.0;
        ctx->last_launches = 0;
        ctx->last_sites = 0.0;
        if (stats) {
            *stats = rps_stats{};
        }
        return fail(ctx, e);
    }
    // Small groups join the largest group of their tile family.
    int lead[2] = {-1, -1};
    for (int v = 0; v < n_variants; ++v) {
        const int f = rps::variant(v)->wide ? 1 : 0;
        if (!by_variant[static_cast<size_t>(v)].empty()
            && (lead[f] < 0 || group_sites[static_cast<size_t>(v)] > group_sites[static_cast<size_t>(lead[f])])) {
            lead[f] = v;
        }
    }
    std::vector<std::vector<int32_t>> family;
    for (int v = 0; v < n_variants; ++v) {
        const int f = rps::variant(v)->wide ? 1 : 0;
        if (!by_variant[static_cast<size_t>(v)].empty() && v != lead[f]
            && group_sites[static_cast<size_t>(v)] < kOwnLaunchSites) {
            std::vector<int32_t> &into = by_variant[static_cast<size_t>(lead[f])];
            into.insert(into.end(), by_variant[static_cast<size_t>(v)].begin(),
                        by_variant[static_cast<size_t>(v)].end());
            by_variant[static_cast<size_t>(v)].clear();
        }
    }
    for (std::vector<int32_t> &g : by_variant) {
        if (!g.empty()) {
            std::sort(g.begin(), g.end());
            family.push_back(std::move(g));
        }
    }
    if (n_requests <= kMaxCurvesPerLaunch && family.size() == 1) {
        return sieve_many_launch(ctx, requests, n_requests, sink, user, stats, -1);
    }
    std::vector<std::vector<int32_t>> groups;
    for (const std::vector<int32_t> &f : family) {
        for (size_t at = 0; at < f.size(); at += kMaxCurvesPerLaunch) {
            const size_t end = std::min(f.size(), at + kMaxCurvesPerLaunch);
            groups.emplace_back(f.begin() + static_cast<std::ptrdiff_t>(at),
                                f.begin() + static_cast<std::ptrdiff_t>(end));
        }
    }
    rps_stats total{};
    double kernel_ms = 0.0, sites = 0.0;
    uint64_t launches = 0;
    int rc = RPS_OK;
    for (const std::vector<int32_t> &g : groups) {
        std::vector<rps_request> sub(g.size());
        for (size_t k = 0; k < g.size(); ++k) {
            sub[k] = requests[g[k]];
        }
        GroupSink gs{sink, user, &g, {}};
        rps_stats one{};
        rc = sieve_many_launch(ctx, sub.data(), sub.size(), GroupSink::call, &gs, &one, -1);
        add_stats(total, one);
        kernel_ms += ctx->last_kernel_ms;
        launches += ctx->last_launches;
        sites += ctx->last_sites;
        if (rc != RPS_OK) {
            break;
        }
    }
    ctx->last_kernel_ms = kernel_ms;
    ctx->last_launches = launches;
    ctx->last_sites = sites;
    if (stats) {
        *stats = total;
    }
    return rc;
}

}  // namespace

extern "C" {

int rps_sieve_many(rps_ctx *ctx, const rps_request *requests, size_t n_requests,
                   rps_multi_sink sink, void *user, rps_stats *stats) {
    if (!ctx) {
        return RPS_EINVAL;
    }
    try {
        return sieve_many_split(ctx, requests, n_requests, sink, user, stats);
    } catch (const std::bad_alloc &) {
        ctx->error = "out of host memory";
        return RPS_ENOMEM;
    } catch (const std::exception &e) {
        ctx->error = std::string("internal error: ") + e.what();
        return RPS_EINTERNAL;
    }
}

}  // extern "C"
