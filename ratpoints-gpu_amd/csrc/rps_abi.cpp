// C ABI of the gfx950 modular square sieve: context, per-curve device
// tables, denominator batching and the double-buffered survivor stream.
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
#include <string>
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

// Signed decimal text mod p, in [0, p).
int decimal_mod(const char *text, int p) {
    bool negative = *text == '-';
    if (*text == '-' || *text == '+') {
        ++text;
    }
    long long v = 0;
    for (; *text; ++text) {
        v = (10 * v + (*text - '0')) % p;
    }
    return static_cast<int>(negative && v ? p - v : v);
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

struct Slot {
    rps::DeviceState *d_state = nullptr;
    rps::DeviceState *h_state = nullptr;  // pinned
    long long *d_num = nullptr;
    int *d_den = nullptr;
    unsigned long long *d_cursors = nullptr;
    size_t cursor_capacity = 0;  // entries
    hipEvent_t k_start = nullptr, k_end = nullptr, ready = nullptr;
};

struct Batch {
    int first_t = 0;
    int n_t = 0;
    int tiles_per_t = 1;
    unsigned long long tile_words = 0;
    size_t tiles = 0;
    bool direct = false;  // short lines: tiles of tile_t whole lines
    int tile_t = 1;
};

}  // namespace

struct rps_ctx {
    int device = 0;
    hipStream_t compute = nullptr;
    hipStream_t copy = nullptr;
    std::string error;
    std::vector<int> primes;
    int *d_primes = nullptr;
    int *d_coef_res = nullptr;
    unsigned long long *d_chi = nullptr;
    unsigned long long *d_resid = nullptr;
    std::vector<std::string> curve_key;
    bool curve_ready = false;
    Slot slots[2];
    size_t capacity = size_t(1) << 24;
    size_t allocated_capacity = 0;
    long long *h_num = nullptr;  // pinned staging
    int *h_den = nullptr;
    size_t stage_capacity = size_t(1) << 20;
    int batch_size = 1 << 16;
    int variant = 0;
    long long min_tiles = 2048;
    unsigned long long direct_words = rps::kDirectMaxWords;
    hipEvent_t basis_start = nullptr, basis_end = nullptr;
    double last_kernel_ms = 0.0;
    uint64_t last_launches = 0;
    double last_sites = 0.0;
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

void prepare_curve(rps_ctx *ctx, const rps_request *req, rps_stats *st) {
    std::vector<std::string> key(req->coefficients,
                                 req->coefficients + req->n_coefficients);
    if (ctx->curve_ready && key == ctx->curve_key) {
        return;
    }
    const int n = req->n_coefficients;
    std::vector<int> res(static_cast<size_t>(rps::kCandidates) * n);
    for (int i = 0; i < rps::kCandidates; ++i) {
        for (int k = 0; k < n; ++k) {
            res[static_cast<size_t>(i) * n + k] =
                decimal_mod(req->coefficients[k], ctx->primes[i]);
        }
    }
    ctx->curve_ready = false;
    RPS_HIP(hipMemcpyAsync(ctx->d_coef_res, res.data(),
                           res.size() * sizeof(int), hipMemcpyHostToDevice,
                           ctx->compute));
    RPS_HIP(hipEventRecord(ctx->basis_start, ctx->compute));
    RPS_HIP(rps::launch_chi(ctx->d_coef_res, n, ctx->d_primes, ctx->d_chi,
                            ctx->compute));
    RPS_HIP(rps::launch_resid(ctx->d_primes, ctx->d_chi, ctx->d_resid,
                              ctx->compute));
    RPS_HIP(hipEventRecord(ctx->basis_end, ctx->compute));
    RPS_HIP(hipEventSynchronize(ctx->basis_end));
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
// always into chunks of at most 2^31 words (the kernel's local word offsets
// are 32-bit).  A chunk keeps at least eight strides per wave so the per-tile
// row build stays amortised.
unsigned long long line_chunks(const rps_ctx *ctx, long long n_t,
                               unsigned long long words, int stride,
                               int waves) {
    unsigned long long chunks = 1;
    if (n_t < ctx->min_tiles) {
        chunks = (static_cast<unsigned long long>(ctx->min_tiles) + n_t - 1) / n_t;
    }
    const unsigned long long min_words = 8ULL * stride * waves;
    chunks = std::min(chunks, std::max(1ULL, words / min_words));
    const unsigned long long limit = 1ULL << 31;
    return std::max(chunks, (words + limit - 1) / limit);
}

Batch make_batch(int first_t, int n_t, unsigned long long chunks,
                 unsigned long long words, int stride) {
    Batch bt;
    bt.first_t = first_t;
    bt.n_t = n_t;
    unsigned long long tile = (words + chunks - 1) / chunks;
    tile = (tile + stride - 1) / stride * stride;
    chunks = (words + tile - 1) / tile;
    bt.tile_words = tile;
    bt.tiles_per_t = static_cast<int>(chunks);
    bt.tiles = static_cast<size_t>(n_t) * chunks;
    return bt;
}

constexpr unsigned long long kMaxTilesPerLaunch = 1ULL << 22;

// Short lines go to the direct kernel: enough denominators per tile to give
// every wave of the workgroup a full 64-pair slice, at most
// kDirectMaxDenominators.
Batch make_direct_batch(int first_t, int n_t, unsigned long long words) {
    Batch bt;
    bt.first_t = first_t;
    bt.n_t = n_t;
    bt.direct = true;
    const unsigned long long want = 64ULL * rps::kDirectWaves;
    int per = static_cast<int>(std::min<unsigned long long>(
        rps::kDirectMaxDenominators, (want + words - 1) / words));
    bt.tile_t = std::max(1, per);
    bt.tiles = (static_cast<size_t>(n_t) + bt.tile_t - 1) / bt.tile_t;
    bt.tile_words = words;
    return bt;
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
        RPS_HIP(hipStreamCreateWithFlags(&ctx->compute, hipStreamNonBlocking));
        RPS_HIP(hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
        ctx->primes = candidate_primes();
        RPS_HIP(hipMalloc(&ctx->d_primes, rps::kCandidates * sizeof(int)));
        RPS_HIP(hipMemcpy(ctx->d_primes, ctx->primes.data(),
                          rps::kCandidates * sizeof(int), hipMemcpyHostToDevice));
        RPS_HIP(hipMalloc(&ctx->d_coef_res, static_cast<size_t>(rps::kCandidates)
                              * rps::kMaxCoefficients * sizeof(int)));
        RPS_HIP(hipMalloc(&ctx->d_chi, static_cast<size_t>(rps::kCandidates)
                              * rps::kChiWords * sizeof(unsigned long long)));
        RPS_HIP(hipMalloc(&ctx->d_resid, static_cast<size_t>(rps::kCandidates)
                              * rps::kMaxPrime * rps::kResidWords
                              * sizeof(unsigned long long)));
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
    if (ctx->h_num) (void)hipHostFree(ctx->h_num);
    if (ctx->h_den) (void)hipHostFree(ctx->h_den);
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
    } else if (key == "variant") {
        if (value < 0 || value >= rps::variant_count()) {
            ctx->error = "variant out of range";
            return RPS_EINVAL;
        }
        ctx->variant = static_cast<int>(value);
    } else if (key == "direct_words") {
        if (value < 0 || value > static_cast<int64_t>(rps::kDirectMaxWords)) {
            ctx->error = "direct_words must be in 0..768";
            return RPS_EINVAL;
        }
        ctx->direct_words = static_cast<unsigned long long>(value);
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
        std::snprintf(name, name_len, "%s", prop.name);
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
        if (!req || !sink) {
            throw ApiError{RPS_EINVAL, "null request or sink"};
        }
        if (req->n_coefficients < 2
            || req->n_coefficients > rps::kMaxCoefficients
            || !req->coefficients) {
            throw ApiError{RPS_EINVAL, "supported polynomial degrees are 1 through 100"};
        }
        for (int k = 0; k < req->n_coefficients; ++k) {
            if (!is_decimal(req->coefficients[k])) {
                throw ApiError{RPS_EINVAL, "coefficients must be decimal integers"};
            }
        }
        if (req->numerator_bound < 1
            || req->numerator_bound > (LLONG_MAX - 32) / 2) {
            throw ApiError{RPS_EINVAL, "invalid numerator bound"};
        }
        // src/sieve_plan.cpp:55-58
        if (req->den_first < 1 || req->den_first > req->den_last
            || (req->square_denominators && req->den_last > rps::kMaxParamT)) {
            throw ApiError{RPS_EINVAL, "invalid denominator range"};
        }
        RPS_HIP(hipSetDevice(ctx->device));
        const rps::VariantInfo *var = rps::variant(ctx->variant);
        const int stride = 64 * var->words_per_lane;
        const long long A = req->numerator_bound;
        const unsigned long long words = (2ULL * static_cast<unsigned long long>(A) + 64) / 64;
        const unsigned long long valid_last =
            2ULL * static_cast<unsigned long long>(A) - 64ULL * (words - 1) + 1;
        const unsigned long long last_mask =
            valid_last >= 64 ? ~0ULL : ((1ULL << valid_last) - 1);
        st.word_count = (2ULL * static_cast<unsigned long long>(A) + 32) / 32;

        rps::SieveParams base{};
        base.bound = A;
        base.words = words;
        base.last_mask = last_mask;
        base.square = req->square_denominators ? 1 : 0;
        for (int i = 0; i < rps::kCandidates; ++i) {
            const long long p = ctx->primes[i];
            base.primes[i] = static_cast<int>(p);
            base.neg_bound_mod[i] = static_cast<int>(((-A) % p + p) % p);
        }
        const auto t_workspace = Clock::now();
        st.plan_ms = ms_between(t_plan, t_workspace);
        ensure_capacity(ctx);
        base.capacity = ctx->capacity;
        const auto t_basis = Clock::now();
        st.workspace_ms = ms_between(t_workspace, t_basis);
        prepare_curve(ctx, req, &st);
        base.resid = ctx->d_resid;
        const auto t_setup = Clock::now();
        st.basis_wall_ms = ms_between(t_basis, t_setup);

        const long long last_t = req->den_last;
        long long next_t = req->den_first;
        Batch inflight[2];
        bool busy[2] = {false, false};
        const double line_sites = 2.0 * static_cast<double>(A) + 1.0;

        auto launch = [&](int si, const Batch &bt, bool resume) {
            Slot &s = ctx->slots[si];
            ensure_cursors(s, bt.tiles * std::max(var->waves, rps::kDirectWaves));
            rps::SieveParams prm = base;
            prm.first_t = bt.first_t;
            prm.n_t = bt.n_t;
            prm.tiles_per_t = bt.tiles_per_t;
            prm.tile_words = bt.tile_words;
            prm.resume = resume ? 1 : 0;
            prm.out_num = s.d_num;
            prm.out_den = s.d_den;
            prm.state = s.d_state;
            prm.cursors = s.d_cursors;
            RPS_HIP(rps::launch_reset(s.d_state, ctx->compute));
            RPS_HIP(hipEventRecord(s.k_start, ctx->compute));
            if (bt.direct) {
                prm.tile_t = bt.tile_t;
                prm.magic_words = words == 1 ? 0xFFFFFFFFu
                    : static_cast<unsigned>(0x100000000ULL / words);
                RPS_HIP(rps::launch_direct(prm, static_cast<unsigned>(bt.tiles),
                                           ctx->compute));
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
                ctx->last_sites += line_sites * bt.n_t;
            }
        };
        auto next_batch = [&](Batch &bt) -> bool {
            if (next_t > last_t) {
                return false;
            }
            long long n = std::min<long long>(ctx->batch_size, last_t - next_t + 1);
            if (words <= ctx->direct_words) {
                bt = make_direct_batch(static_cast<int>(next_t), static_cast<int>(n), words);
                next_t += n;
                return true;
            }
            unsigned long long chunks = line_chunks(ctx, n, words, stride, var->waves);
            if (chunks > kMaxTilesPerLaunch) {
                throw ApiError{RPS_EINVAL, "maximum height exceeds GPU limit"};
            }
            n = std::min<long long>(n, static_cast<long long>(kMaxTilesPerLaunch / chunks));
            chunks = line_chunks(ctx, n, words, stride, var->waves);
            bt = make_batch(static_cast<int>(next_t), static_cast<int>(n), chunks, words, stride);
            next_t += n;
            return true;
        };
        bool cancelled = false;
        // Drains one finished launch of slot si into the sink; returns
        // whether the batch still needs a resume launch.
        auto drain = [&](int si) -> bool {
            Slot &s = ctx->slots[si];
            RPS_HIP(hipEventSynchronize(s.ready));
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
                const Batch &bt = inflight[cur];
                if (done(user, bt.first_t, bt.first_t + bt.n_t - 1,
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
        ctx->error = op_name(e.op) + ": " + hipGetErrorString(e.code);
        (void)hipStreamSynchronize(ctx->compute);
        (void)hipStreamSynchronize(ctx->copy);
        ctx->curve_ready = false;
        return RPS_EDEVICE;
    } catch (const ApiError &e) {
        ctx->error = e.message;
        (void)hipStreamSynchronize(ctx->compute);
        (void)hipStreamSynchronize(ctx->copy);
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->error = "out of host memory";
        return RPS_ENOMEM;
    }
}

}  // extern "C"
