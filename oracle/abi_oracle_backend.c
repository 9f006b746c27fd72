/*
 * CPU test double of the C ABI in include/ratpoints_sieve.h.  TEST
 * INFRASTRUCTURE ONLY -- it is NOT a search backend.
 *
 * Linked (instead of the HIP library) into oracle/_build/ratpoints_gpu_oracle,
 * a build of the C++ host used only by tests/test_cli.py to check the host
 * logic -- option parsing, batching, ordering, interval filtering, output
 * format, -1 and error paths -- without a GPU, as the reference's
 * tests/mock_backend.cpp does for its host.  Survivors come from the oracle's
 * per-site evaluation (sieve_oracle.c), are delivered in reverse order and in
 * flushes of at most 97 so the host's sorting is exercised, and boxes larger
 * than 5e8 sites are refused so the double cannot be mistaken for a backend.
 * Its device is named "Oracle CPU test double".
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ratpoints_sieve.h"

typedef struct rpo_curve rpo_curve;
rpo_curve *rpo_curve_new(const char *const *coeffs, int n_coeffs);
void rpo_curve_free(rpo_curve *curve);
long long rpo_sieve_direct(const rpo_curve *curve, long long bound, int first,
                           int last, int square, int threads, long long *num,
                           int *den, long long cap, uint64_t digest[2]);

int rps_check_request(const rps_request *req, char *message, size_t message_len);

struct rps_ctx {
    char error[256];
    int batch_size;
    int primitive;
};

static char open_error[256];

int rps_open(int device, rps_ctx **out) {
    if (!out) {
        return RPS_EINVAL;
    }
    *out = NULL;
    if (getenv("RPS_TEST_OPEN_FAIL")) {
        snprintf(open_error, sizeof open_error, "hipSetDevice: injected failure");
        return RPS_EDEVICE;
    }
    if (device != 0) {
        snprintf(open_error, sizeof open_error, "invalid GPU device %d", device);
        return RPS_EINVAL;
    }
    rps_ctx *ctx = (rps_ctx *)calloc(1, sizeof(rps_ctx));
    if (!ctx) {
        return RPS_ENOMEM;
    }
    ctx->batch_size = 65536;
    *out = ctx;
    return RPS_OK;
}

void rps_close(rps_ctx *ctx) { free(ctx); }

const char *rps_last_error(const rps_ctx *ctx) {
    return ctx ? ctx->error : open_error;
}

int rps_set_option(rps_ctx *ctx, const char *name, int64_t value) {
    if (!ctx || !name) {
        return RPS_EINVAL;
    }
    if (strcmp(name, "batch_size") == 0 && value >= 1 && value <= (1 << 20)) {
        ctx->batch_size = (int)value;
        return RPS_OK;
    }
    if (strcmp(name, "primitive_only") == 0 && (value == 0 || value == 1)) {
        ctx->primitive = (int)value;
        return RPS_OK;
    }
    if (strcmp(name, "survivor_capacity") == 0 || strcmp(name, "variant") == 0
        || strcmp(name, "min_tiles") == 0 || strcmp(name, "table_slots") == 0) {
        return RPS_OK;
    }
    snprintf(ctx->error, sizeof ctx->error, "unknown option: %s", name);
    return RPS_EINVAL;
}

int rps_variant_count(void) { return 1; }

const char *rps_variant_name(int variant) {
    return variant == 0 ? "oracle" : NULL;
}

int rps_device_count(int *count) {
    if (!count) {
        return RPS_EINVAL;
    }
    const char *forced = getenv("RPS_TEST_DEVICE_COUNT");
    *count = forced ? atoi(forced) : 1;
    return RPS_OK;
}

int rps_device_info(int device, char *name, size_t name_len, char *pci,
                    size_t pci_len, int *compute_units, uint64_t *memory_bytes,
                    int *major, int *minor) {
    if (device != 0) {
        snprintf(open_error, sizeof open_error, "hipGetDeviceProperties: invalid device");
        return RPS_EDEVICE;
    }
    if (name && name_len) snprintf(name, name_len, "Oracle CPU test double");
    if (pci && pci_len) snprintf(pci, pci_len, "none");
    if (compute_units) *compute_units = 1;
    if (memory_bytes) *memory_bytes = 0;
    if (major) *major = 0;
    if (minor) *minor = 0;
    return RPS_OK;
}

int rps_last_kernel_timing(const rps_ctx *ctx, double *kernel_ms,
                           uint64_t *launches, double *sites) {
    (void)ctx;
    if (kernel_ms) *kernel_ms = 0.0;
    if (launches) *launches = 0;
    if (sites) *sites = 0.0;
    return RPS_OK;
}

int rps_sieve_batched(rps_ctx *ctx, const rps_request *req, rps_sink sink,
                      rps_batch_done done, void *user, rps_stats *stats) {
    rps_stats local;
    memset(&local, 0, sizeof local);
    if (!ctx || !req || !sink) {
        return RPS_EINVAL;
    }
    ctx->error[0] = '\0';
    if (rps_check_request(req, ctx->error, sizeof ctx->error) != RPS_OK) {
        return RPS_EINVAL;
    }
    double sites = (2.0 * (double)req->numerator_bound + 1.0)
        * ((double)req->den_last - req->den_first + 1.0);
    if (sites > 5e8) {
        snprintf(ctx->error, sizeof ctx->error,
                 "oracle test double refuses %.3g sites", sites);
        return RPS_EINVAL;
    }
    rpo_curve *curve = rpo_curve_new(req->coefficients, req->n_coefficients);
    if (!curve) {
        snprintf(ctx->error, sizeof ctx->error, "oracle curve setup failed");
        return RPS_ENOMEM;
    }
    local.word_count = (2ULL * (unsigned long long)req->numerator_bound + 32) / 32;
    int status = RPS_OK;
    for (long long t = req->den_first; t <= req->den_last && status == RPS_OK;) {
        long long last = t + ctx->batch_size - 1;
        if (last > req->den_last) {
            last = req->den_last;
        }
        uint64_t digest[2];
        long long n = rpo_sieve_direct(curve, req->numerator_bound, (int)t, (int)last,
                                       req->square_denominators, 1, NULL, NULL, 0, digest);
        long long *num = (long long *)malloc(sizeof(long long) * (size_t)(n + 1));
        int *den = (int *)malloc(sizeof(int) * (size_t)(n + 1));
        if (!num || !den || n < 0
            || rpo_sieve_direct(curve, req->numerator_bound, (int)t, (int)last,
                                req->square_denominators, 1, num, den, n, digest) != n) {
            free(num);
            free(den);
            snprintf(ctx->error, sizeof ctx->error, "oracle sieve failed");
            status = RPS_EINTERNAL;
            break;
        }
        if (ctx->primitive) {
            /* Keep gcd(a, b) = 1 only. */
            long long kept = 0;
            for (long long i = 0; i < n; ++i) {
                unsigned long long x = num[i] < 0 ? 0ULL - (unsigned long long)num[i]
                                                   : (unsigned long long)num[i];
                unsigned long long y = (unsigned long long)den[i];
                while (y) {
                    unsigned long long r = x % y;
                    x = y;
                    y = r;
                }
                if (x == 1) {
                    num[kept] = num[i];
                    den[kept] = den[i];
                    ++kept;
                }
            }
            n = kept;
        }
        /* Reverse order, flushes of at most 97. */
        for (long long i = 0; i < n / 2; ++i) {
            long long tn = num[i];
            num[i] = num[n - 1 - i];
            num[n - 1 - i] = tn;
            int td = den[i];
            den[i] = den[n - 1 - i];
            den[n - 1 - i] = td;
        }
        for (long long at = 0; at < n; at += 97) {
            size_t k = (size_t)(n - at < 97 ? n - at : 97);
            local.survivor_count += k;
            if (sink(user, (const int64_t *)(num + at), (const int32_t *)(den + at), k)) {
                status = RPS_ECANCELLED;
                break;
            }
        }
        free(num);
        free(den);
        ++local.batches;
        ++local.launches;
        if (status == RPS_OK && done
            && done(user, (int32_t)t, (int32_t)last, (uint64_t)n) != 0) {
            status = RPS_ECANCELLED;
        }
        t = last + 1;
    }
    rpo_curve_free(curve);
    if (stats) {
        *stats = local;
    }
    return status;
}

int rps_sieve(rps_ctx *ctx, const rps_request *req, rps_sink sink, void *user,
              rps_stats *stats) {
    return rps_sieve_batched(ctx, req, sink, NULL, user, stats);
}

int rps_check_request(const rps_request *req, char *message, size_t message_len) {
    const char *reason = "";
    int rc = RPS_OK;
    if (!req || req->n_coefficients < 2 || req->n_coefficients > 101 || !req->coefficients) {
        reason = "supported polynomial degrees are 1 through 100";
        rc = RPS_EINVAL;
    } else if (req->numerator_bound < 1 || req->numerator_bound > (INT64_MAX - 32) / 2) {
        reason = "invalid numerator bound";
        rc = RPS_EINVAL;
    } else if ((2 * (uint64_t)req->numerator_bound + 64) / 64 > ((uint64_t)1 << 47)) {
        /* the device library's line limit: 2^47 64-bit words */
        reason = "maximum height exceeds GPU limit";
        rc = RPS_EINVAL;
    } else if (req->den_first < 1 || req->den_first > req->den_last
               || (req->square_denominators && req->den_last > 46340)) {
        reason = "invalid denominator range";
        rc = RPS_EINVAL;
    }
    if (message && message_len) {
        snprintf(message, message_len, "%s", reason);
    }
    return rc;
}

/* Many requests: each searched alone, in reverse request order, survivors
 * tagged with their request index. */
typedef struct {
    rps_multi_sink sink;
    void *user;
    int32_t request;
    int32_t *tags;
    size_t tags_len;
} many_fwd;

static int many_forward(void *u, const int64_t *num, const int32_t *den, size_t n) {
    many_fwd *f = (many_fwd *)u;
    if (n > f->tags_len) {
        int32_t *grown = (int32_t *)realloc(f->tags, n * sizeof *grown);
        if (!grown) {
            return 1;
        }
        f->tags = grown;
        f->tags_len = n;
    }
    for (size_t i = 0; i < n; ++i) {
        f->tags[i] = f->request;
    }
    return f->sink(f->user, f->tags, num, den, n);
}

int rps_sieve_many(rps_ctx *ctx, const rps_request *requests, size_t n_requests,
                   rps_multi_sink sink, void *user, rps_stats *stats) {
    rps_stats total;
    memset(&total, 0, sizeof total);
    if (!ctx || !sink || (!requests && n_requests)) {
        return RPS_EINVAL;
    }
    for (size_t r = 0; r < n_requests; ++r) {
        if (rps_check_request(&requests[r], ctx->error, sizeof ctx->error) != RPS_OK) {
            return RPS_EINVAL;
        }
    }
    many_fwd fwd = {sink, user, 0, NULL, 0};
    int status = RPS_OK;
    for (size_t k = n_requests; k-- > 0 && status == RPS_OK;) {
        rps_stats one;
        fwd.request = (int32_t)k;
        status = rps_sieve(ctx, &requests[k], many_forward, &fwd, &one);
        total.survivor_count += one.survivor_count;
        total.launches += one.launches;
    }
    free(fwd.tags);
    total.batches = n_requests;
    if (stats) {
        *stats = total;
    }
    return status;
}
