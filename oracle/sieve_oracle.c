/*
 * CPU oracle for the modular square sieve.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker (or as the timed CPU baseline).  The
 * product path (ratpoints-gpu_amd/) never links, loads or calls it.
 *
 * It restates, in plain C, the survivor set defined by the reference's host
 * sieve plan and CUDA sieve (paths relative to the reference repository):
 *
 *   candidate primes   the 36 largest primes below 512
 *                      (src/sieve_plan.cpp:28-45, 60-61, sieve_plan.hpp:27-30)
 *   prime selection    for denominator b, the first 32 candidates in
 *                      descending order that do not divide b
 *                      (src/sieve_plan.cpp:71-90)
 *   decimal_mod        signed decimal string mod p (src/sieve_plan.cpp:15-26)
 *   mod_inverse        extended Euclid (src/sieve_plan.cpp:98-113)
 *   affine mask        chi_p(x) = [f(x) mod p is a square or zero]
 *                      (src/sieve_plan.cpp:115-138)
 *   survivor predicate (a, b), |a| <= A, survives iff
 *                      chi_p((a mod p) * (b mod p)^-1 mod p) = 1 for every
 *                      selected p (src/sieve.cu:141-172 builds the bits,
 *                      src/sieve.cu:296-331 ANDs them and masks the tail;
 *                      src/sieve.cu:250-252 squares b in square mode;
 *                      src/sieve.cu:179 gives a = -A + 32*word + bit).
 *
 * Two independent evaluations of the same set are provided:
 *   rpo_sieve_direct  per-site evaluation of the predicate, prime by prime;
 *   rpo_sieve_words   64-bit packed rows, the CPU counterpart of the GPU
 *                     kernel's layout (used as the timed CPU baseline).
 *
 * Pinning (see tests/test_oracle.py): the direct method is checked against
 * an independent Python Legendre-symbol brute force of F(a, b) mod p on tiny
 * boxes, and both methods against every survivor count the reference holds
 * (tests/test_sieve_streaming.cpp dense case, docs/RTX4090-*.json and
 * experiments/elliptic-curves/results/rank31-gpu-fixed-*.json) that a CPU can
 * reach in seconds; the word method is checked against the direct method on
 * seeded random curves.  The selection/mask/inverse helpers are also checked
 * against the reference's own src/sieve_plan.cpp compiled into oracle/_ref.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define RPO_MAX_PRIME 512
#define RPO_CANDIDATES 36
#define RPO_SELECTED 32
#define RPO_FAST 14

/* ---------------------------------------------------------------------- */
/* Plan helpers                                                           */
/* ---------------------------------------------------------------------- */

/* Sieve of Eratosthenes below `limit` (src/sieve_plan.cpp:28-45).  Writes the
 * 36 largest primes below 512 in ascending order, the order of
 * SievePlan::primes (src/sieve_plan.cpp:60-61). */
int rpo_candidate_primes(int *out) {
    unsigned char composite[RPO_MAX_PRIME] = {0};
    int primes[RPO_MAX_PRIME];
    int count = 0;
    for (int n = 2; n < RPO_MAX_PRIME; ++n) {
        if (composite[n]) {
            continue;
        }
        primes[count++] = n;
        for (int m = n * n; m < RPO_MAX_PRIME; m += n) {
            composite[m] = 1;
        }
    }
    for (int i = 0; i < RPO_CANDIDATES; ++i) {
        out[i] = primes[count - RPO_CANDIDATES + i];
    }
    return RPO_CANDIDATES;
}

/* Signed decimal string mod p in [0, p) (src/sieve_plan.cpp:15-26). */
int rpo_decimal_mod(const char *text, int modulus) {
    int negative = 0;
    if (*text == '-' || *text == '+') {
        negative = *text == '-';
        ++text;
    }
    long long value = 0;
    for (; *text; ++text) {
        value = (10 * value + (*text - '0')) % modulus;
    }
    return (int)(negative && value ? modulus - value : value);
}

/* Inverse of value modulo m by the extended Euclidean algorithm
 * (src/sieve_plan.cpp:98-113); returns a value in [0, m). */
long long rpo_mod_inverse(long long value, long long modulus) {
    long long r0 = value, r1 = modulus, s0 = 1, s1 = 0;
    while (r1) {
        long long q = r0 / r1, t = r0 - q * r1;
        r0 = r1;
        r1 = t;
        t = s0 - q * s1;
        s0 = s1;
        s1 = t;
    }
    return ((s0 % modulus) + modulus) % modulus;
}

/* chi_p over x in [0, p): 1 iff f(x) mod p is a square (zero included)
 * (src/sieve_plan.cpp:115-138).  Coefficients are c0..cn as decimal text. */
int rpo_affine_mask(int prime, const char *const *coeffs, int n_coeffs,
                    unsigned char *out) {
    int *res = (int *)malloc(sizeof(int) * (size_t)n_coeffs);
    unsigned char *square = (unsigned char *)calloc((size_t)prime, 1);
    if (!res || !square) {
        free(res);
        free(square);
        return -1;
    }
    for (int i = 0; i < n_coeffs; ++i) {
        res[i] = rpo_decimal_mod(coeffs[i], prime);
    }
    for (long long y = 0; y < prime; ++y) {
        square[y * y % prime] = 1;
    }
    for (long long x = 0; x < prime; ++x) {
        long long v = 0;
        for (int i = n_coeffs - 1; i >= 0; --i) {
            v = (v * x + res[i]) % prime;
        }
        out[x] = square[v];
    }
    free(res);
    free(square);
    return 0;
}

/* Indices (into the ascending candidate list) of the 32 largest candidates
 * that do not divide `denominator`, largest first (src/sieve_plan.cpp:75-86).
 * Returns 32, or -1 when fewer than 32 qualify. */
int rpo_select_primes(long long denominator, int *out) {
    int primes[RPO_CANDIDATES];
    rpo_candidate_primes(primes);
    int count = 0;
    for (int i = RPO_CANDIDATES - 1; i >= 0 && count < RPO_SELECTED; --i) {
        if (denominator % primes[i] != 0) {
            out[count++] = i;
        }
    }
    return count == RPO_SELECTED ? count : -1;
}

/* ---------------------------------------------------------------------- */
/* Order-independent survivor digest                                      */
/* ---------------------------------------------------------------------- */

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* Digest of one survivor; the set digest is the pair of wrapping sums of
 * h and splitmix64(h ^ C) over all survivors. */
static inline uint64_t site_hash(long long a, int b) {
    return splitmix64((uint64_t)a ^ splitmix64((uint64_t)(uint32_t)b));
}

void rpo_digest(const long long *num, const int *den, long long n,
                uint64_t out[2]) {
    uint64_t h0 = 0, h1 = 0;
    for (long long i = 0; i < n; ++i) {
        uint64_t h = site_hash(num[i], den[i]);
        h0 += h;
        h1 += splitmix64(h ^ 0xA5A5A5A5A5A5A5A5ULL);
    }
    out[0] = h0;
    out[1] = h1;
}

/* ---------------------------------------------------------------------- */
/* Curve context: chi_p tables for the 36 candidates                      */
/* ---------------------------------------------------------------------- */

typedef struct rpo_curve {
    int primes[RPO_CANDIDATES];
    unsigned char chi[RPO_CANDIDATES][RPO_MAX_PRIME];
} rpo_curve;

rpo_curve *rpo_curve_new(const char *const *coeffs, int n_coeffs) {
    rpo_curve *curve = (rpo_curve *)calloc(1, sizeof(rpo_curve));
    if (!curve) {
        return NULL;
    }
    rpo_candidate_primes(curve->primes);
    for (int i = 0; i < RPO_CANDIDATES; ++i) {
        if (rpo_affine_mask(curve->primes[i], coeffs, n_coeffs,
                            curve->chi[i]) != 0) {
            free(curve);
            return NULL;
        }
    }
    return curve;
}

void rpo_curve_free(rpo_curve *curve) { free(curve); }

/* ---------------------------------------------------------------------- */
/* Survivor enumeration                                                   */
/* ---------------------------------------------------------------------- */

typedef struct {
    long long *num;
    int *den;
    long long size, cap;
    uint64_t h0, h1;
    int keep;
    int failed;
} sink_t;

static void sink_push(sink_t *s, long long a, int b) {
    uint64_t h = site_hash(a, b);
    s->h0 += h;
    s->h1 += splitmix64(h ^ 0xA5A5A5A5A5A5A5A5ULL);
    if (s->keep && !s->failed) {
        if (s->size == s->cap) {
            long long cap = s->cap ? 2 * s->cap : 1024;
            long long *num = (long long *)realloc(s->num, sizeof(long long) * (size_t)cap);
            int *den = num ? (int *)realloc(s->den, sizeof(int) * (size_t)cap) : NULL;
            if (!num || !den) {
                if (num) {
                    s->num = num;
                }
                s->failed = 1;
                ++s->size;
                return;
            }
            s->num = num;
            s->den = den;
            s->cap = cap;
        }
        s->num[s->size] = a;
        s->den[s->size] = b;
    }
    ++s->size;
}

typedef struct {
    const rpo_curve *curve;
    long long bound;
    long long first, last; /* denominator parameters, inclusive */
    int square;
    int method;            /* 0 = direct, 1 = words */
    sink_t sink;
} job_t;

static long long pos_mod(long long a, long long p) {
    long long r = a % p;
    return r < 0 ? r + p : r;
}

/* Per-site evaluation: for each selected prime, in selection order, look up
 * chi_p(a * b^-1 mod p) and stop at the first failure. */
static void sieve_direct_one(const rpo_curve *curve, long long bound,
                             long long b, sink_t *sink) {
    int sel[RPO_SELECTED];
    if (rpo_select_primes(b, sel) < 0) {
        sink->failed = 1;
        return;
    }
    int p[RPO_SELECTED];
    long long inv[RPO_SELECTED];
    for (int k = 0; k < RPO_SELECTED; ++k) {
        p[k] = curve->primes[sel[k]];
        inv[k] = rpo_mod_inverse(b % p[k], p[k]);
    }
    for (long long a = -bound; a <= bound; ++a) {
        int alive = 1;
        for (int k = 0; k < RPO_SELECTED && alive; ++k) {
            long long x = pos_mod(a, p[k]) * inv[k] % p[k];
            alive = curve->chi[sel[k]][x];
        }
        if (alive) {
            sink_push(sink, a, (int)b);
        }
    }
}

/* Packed evaluation: word w holds a = -A + 64w + j, j = 0..63.  Row k entry t
 * holds bit j = chi_p(((-A + 64t + j) mod p) * b^-1 mod p); word w uses entry
 * w mod p.  The 14 largest selected primes are ANDed for every word, the rest
 * only for words that still have bits set; the last word is masked to the
 * 2A + 1 sites of the line. */
static void sieve_words_one(const rpo_curve *curve, long long bound,
                            long long b, sink_t *sink,
                            uint64_t *rows /* [32][512] scratch */) {
    int sel[RPO_SELECTED];
    if (rpo_select_primes(b, sel) < 0) {
        sink->failed = 1;
        return;
    }
    unsigned long long span = 2ULL * (unsigned long long)bound; /* a + A max */
    unsigned long long words = (span + 64) / 64;
    int p[RPO_SELECTED];
    for (int k = 0; k < RPO_SELECTED; ++k) {
        p[k] = curve->primes[sel[k]];
        long long r = rpo_mod_inverse(b % p[k], p[k]);
        long long c = pos_mod(-bound, p[k]);
        unsigned char perm[RPO_MAX_PRIME];
        for (int z = 0; z < p[k]; ++z) {
            perm[z] = curve->chi[sel[k]][(long long)z * r % p[k]];
        }
        unsigned long long entries = words < (unsigned long long)p[k]
            ? words : (unsigned long long)p[k];
        /* Bit z of `string` is perm[z mod p] for z < p + 64, so the 64
         * numerators from residue z0 on are bits z0 .. z0 + 63 of it. */
        uint64_t string[(RPO_MAX_PRIME + 64) / 64 + 2] = {0};
        for (int z = 0; z < p[k] + 64; ++z) {
            string[z >> 6] |= (uint64_t)perm[z % p[k]] << (z & 63);
        }
        for (unsigned long long t = 0; t < entries; ++t) {
            const long long z0 = (long long)((c + 64 * (long long)t) % p[k]);
            const int shift = (int)(z0 & 63);
            const uint64_t lo = string[z0 >> 6];
            const uint64_t hi = string[(z0 >> 6) + 1];
            rows[k * RPO_MAX_PRIME + t] = shift ? (lo >> shift) | (hi << (64 - shift)) : lo;
        }
    }
    int off[RPO_FAST];
    for (int k = 0; k < RPO_FAST; ++k) {
        off[k] = 0;
    }
    for (unsigned long long w = 0; w < words; ++w) {
        uint64_t m = ~0ULL;
        for (int k = 0; k < RPO_FAST; ++k) {
            m &= rows[k * RPO_MAX_PRIME + off[k]];
            if (++off[k] == p[k]) {
                off[k] = 0;
            }
        }
        if (!m) {
            continue;
        }
        for (int k = RPO_FAST; k < RPO_SELECTED && m; ++k) {
            m &= rows[k * RPO_MAX_PRIME + (int)(w % (unsigned long long)p[k])];
        }
        if (w == words - 1) {
            unsigned long long valid = span - 64 * w + 1;
            if (valid < 64) {
                m &= (1ULL << valid) - 1;
            }
        }
        while (m) {
            int j = __builtin_ctzll(m);
            m &= m - 1;
            long long a = (long long)((unsigned long long)(-bound) + 64 * w + (unsigned)j);
            sink_push(sink, a, (int)b);
        }
    }
}

static void *run_job(void *arg) {
    job_t *job = (job_t *)arg;
    uint64_t *rows = NULL;
    if (job->method == 1) {
        rows = (uint64_t *)malloc(sizeof(uint64_t) * RPO_SELECTED * RPO_MAX_PRIME);
        if (!rows) {
            job->sink.failed = 1;
            return NULL;
        }
    }
    for (long long t = job->first; t <= job->last && !job->sink.failed; ++t) {
        long long b = job->square ? t * t : t;
        if (job->method == 0) {
            sieve_direct_one(job->curve, job->bound, b, &job->sink);
        } else {
            sieve_words_one(job->curve, job->bound, b, &job->sink, rows);
        }
    }
    free(rows);
    return NULL;
}

/* Enumerate the survivor set of (curve, A, [first, last], square mode).
 * Returns the survivor count, or -1 on invalid input / allocation failure.
 * digest[2] receives the order-independent set digest.  When `num`/`den` are
 * non-NULL and `cap` >= count, survivors are written sorted by (b, a). */
static long long sieve_common(const rpo_curve *curve, long long bound,
                              int first, int last, int square, int threads,
                              int method, long long *num, int *den,
                              long long cap, uint64_t digest[2]) {
    if (!curve || bound < 1 || first < 1 || first > last
        || (square && last > 46340)) {
        return -1;
    }
    long long count = (long long)last - first + 1;
    if (threads < 1) {
        threads = 1;
    }
    if (threads > count) {
        threads = (int)count;
    }
    job_t *jobs = (job_t *)calloc((size_t)threads, sizeof(job_t));
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    if (!jobs || !tid) {
        free(jobs);
        free(tid);
        return -1;
    }
    /* Contiguous, balanced ranges so concatenation keeps (b, a) order.  In
     * square mode work grows with nothing but the count, so equal counts
     * balance too. */
    long long start = first;
    for (int i = 0; i < threads; ++i) {
        long long share = count / threads + (i < count % threads);
        jobs[i].curve = curve;
        jobs[i].bound = bound;
        jobs[i].first = start;
        jobs[i].last = start + share - 1;
        jobs[i].square = square;
        jobs[i].method = method;
        jobs[i].sink.keep = num != NULL;
        start += share;
    }
    int spawned = 0;
    for (int i = 1; i < threads; ++i) {
        if (pthread_create(&tid[i], NULL, run_job, &jobs[i]) != 0) {
            break;
        }
        spawned = i;
    }
    run_job(&jobs[0]);
    for (int i = 1; i <= spawned; ++i) {
        pthread_join(tid[i], NULL);
    }
    for (int i = spawned + 1; i < threads; ++i) {
        run_job(&jobs[i]);
    }
    long long total = 0;
    uint64_t h0 = 0, h1 = 0;
    int failed = 0;
    for (int i = 0; i < threads; ++i) {
        total += jobs[i].sink.size;
        h0 += jobs[i].sink.h0;
        h1 += jobs[i].sink.h1;
        failed |= jobs[i].sink.failed;
    }
    if (num && !failed && total <= cap) {
        long long at = 0;
        for (int i = 0; i < threads; ++i) {
            memcpy(num + at, jobs[i].sink.num, sizeof(long long) * (size_t)jobs[i].sink.size);
            memcpy(den + at, jobs[i].sink.den, sizeof(int) * (size_t)jobs[i].sink.size);
            at += jobs[i].sink.size;
        }
    }
    for (int i = 0; i < threads; ++i) {
        free(jobs[i].sink.num);
        free(jobs[i].sink.den);
    }
    free(jobs);
    free(tid);
    if (digest) {
        digest[0] = h0;
        digest[1] = h1;
    }
    return failed ? -1 : total;
}

long long rpo_sieve_direct(const rpo_curve *curve, long long bound, int first,
                           int last, int square, int threads, long long *num,
                           int *den, long long cap, uint64_t digest[2]) {
    return sieve_common(curve, bound, first, last, square, threads, 0, num,
                        den, cap, digest);
}

long long rpo_sieve_words(const rpo_curve *curve, long long bound, int first,
                          int last, int square, int threads, long long *num,
                          int *den, long long cap, uint64_t digest[2]) {
    return sieve_common(curve, bound, first, last, square, threads, 1, num,
                        den, cap, digest);
}
