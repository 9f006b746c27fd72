/*
 * ratpoints_sieve.h -- C ABI of the MI355X (gfx950) modular square sieve.
 *
 * This is the drop-in boundary for the search backend of ratpoints_gpu.  It
 * replaces, one for one, what the reference binds at its backend interface
 * (paths relative to the reference repository):
 *
 *   run_modular_sieve(coefficients, numerator_bound, denominators, callback,
 *                     square_denominators)             include/sieve.hpp:43-47
 *   CandidateBatch / CandidateBatchCallback            include/sieve.hpp:32-40
 *   SieveResult / SieveMetrics                         include/sieve.hpp:13-30
 *   check_cuda error convention (op: message)          include/cuda_support.hpp:14-19
 *   available_gpu_devices / activate_gpu               include/gpu_devices.hpp:20-23
 *
 * Only plain C types cross it: pointers, sizes and fixed-width integers.  The
 * C++ surface with the reference's exact signature (ratpoints_gpu::
 * run_modular_sieve) is a thin shim over these functions, in
 * ratpoints-gpu_amd/host/sieve_shim.cpp.
 *
 * Threading: one rps_ctx per thread; distinct contexts may run concurrently.
 * A context binds its device at rps_open and owns its streams, device tables
 * and pinned staging buffers until rps_close.
 */
#ifndef RATPOINTS_SIEVE_H
#define RATPOINTS_SIEVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPS_ABI_VERSION 1

/* Status codes.  Every entry point returning int returns one of these; the
 * message of the last failure on a context is rps_last_error(ctx). */
enum {
    RPS_OK = 0,
    RPS_EINVAL = 1,     /* invalid argument (std::invalid_argument)        */
    RPS_EDEVICE = 2,    /* HIP runtime or kernel failure (runtime_error)   */
    RPS_ENOMEM = 3,     /* device or pinned allocation failed              */
    RPS_ECANCELLED = 4, /* the sink asked to stop; not an error for -1     */
    RPS_EOVERFLOW = 5,  /* survivor total overflowed (overflow_error)      */
    RPS_EINTERNAL = 6   /* invariant violated (logic_error)                */
};

typedef struct rps_ctx rps_ctx;

/* One sieve request: the reference's run_modular_sieve arguments
 * (include/sieve.hpp:43-47, include/search_types.hpp:9-22). */
typedef struct rps_request {
    const char *const *coefficients; /* decimal c0 .. cn, optional sign     */
    int32_t n_coefficients;          /* n + 1, 2..101                       */
    int64_t numerator_bound;         /* A >= 1: numerators a in [-A, A];
                                        A < 2^52 (lines of at most 2^47
                                        64-bit words), else RPS_EINVAL      */
    int32_t den_first;               /* first denominator parameter, >= 1   */
    int32_t den_last;                /* last parameter, inclusive           */
    int32_t square_denominators;     /* nonzero: b = t^2, den_last <= 46340 */
} rps_request;

/* Survivor sink: the reference's CandidateBatchCallback
 * (include/sieve.hpp:38-40).  `numerators[i] / denominators[i]` are raw
 * modular survivors (a, b); b is t^2 in square mode.  Order within and across
 * calls is unspecified, each survivor is delivered exactly once.  The arrays
 * are valid only during the call.  Return 0 to continue; any nonzero value
 * cancels the request, which then returns RPS_ECANCELLED after the current
 * denominator batch. */
typedef int (*rps_sink)(void *user, const int64_t *numerators,
                        const int32_t *denominators, size_t count);

/* Metrics of one request: SieveResult + SieveMetrics
 * (include/sieve.hpp:13-30).  word_count and initial_mask_bytes keep the
 * reference's definitions (32-bit words, (2A+32)/32; den x words x 14 x 4 B)
 * so the -v metrics line stays comparable. */
typedef struct rps_stats {
    uint64_t survivor_count;  /* raw modular survivors delivered          */
    uint64_t word_count;      /* (2A + 32) / 32, as SievePlan::word_count */
    double basis_ms;          /* device time of the per-curve table build */
    double sieve_ms;          /* device time of all sieve launches        */
    double plan_ms;
    double workspace_ms;
    double basis_wall_ms;
    double survivor_setup_ms;
    double verification_ms;   /* wall time spent inside the sink          */
    double sieve_wall_ms;
    double initial_mask_bytes;
    uint64_t launches;        /* sieve kernel launches, resumes included  */
    uint64_t batches;         /* denominator batches (requests, for many) */
    uint64_t requeued;        /* requests searched again on the plain
                                 path after a survivor-buffer overflow    */
} rps_stats;

/* Opens a context on HIP device `device` (0-based among visible devices).
 * Replaces activate_gpu (include/gpu_devices.hpp:23) plus the per-call
 * workspace construction of src/sieve.cu:65-95, hoisted to open time. */
int rps_open(int device, rps_ctx **out);

/* Releases every device and host resource of the context. */
void rps_close(rps_ctx *ctx);

/* Message of the last failure on `ctx` ("" if none).  A NULL ctx returns
 * the message of the last failed rps_open on this thread. */
const char *rps_last_error(const rps_ctx *ctx);

/* Runs the modular square sieve over [-A, A] x [den_first, den_last] and
 * streams every survivor to `sink`.  Replaces run_modular_sieve
 * (include/sieve.hpp:43-47, src/sieve.cu:503-580).  Errors follow the
 * reference: invalid range -> RPS_EINVAL ("invalid denominator range",
 * src/sieve_plan.cpp:55-58), HIP failures -> RPS_EDEVICE ("<op>: <msg>").
 * `stats` may be NULL. */
int rps_sieve(rps_ctx *ctx, const rps_request *request, rps_sink sink,
              void *user, rps_stats *stats);

/* Batch-completion notice: every survivor with a denominator parameter in
 * [first_t, last_t] has been delivered to the sink, and batches complete in
 * increasing parameter order.  Return 0 to continue; nonzero cancels the
 * request like a nonzero sink return.  No counterpart in the reference, whose
 * caller sees one batch per run_modular_sieve call
 * (src/point_search.cpp:195-237); it lets a caller sieve a whole range in one
 * call -- so the GPU works on batch i+1 while the caller verifies batch i --
 * and still emit points in (b, a) order and stop early for -1. */
typedef int (*rps_batch_done)(void *user, int32_t first_t, int32_t last_t,
                              uint64_t survivors);

/* rps_sieve with a batch-completion callback (`done` may be NULL). */
int rps_sieve_batched(rps_ctx *ctx, const rps_request *request, rps_sink sink,
                      rps_batch_done done, void *user, rps_stats *stats);

/* Allocation-free form of rps_sieve.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory while they run (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The call makes no device or host allocation: the survivors go to the
 * caller's device buffers `numerators` and `denominators` (`capacity`
 * entries each, any order), and the kernels' partial results (the curve's
 * residue table, flush state, cursors) live in a fixed device scratch array
 * of 3.4 MB per device, held until the process ends, whose address rps_open
 * resolves.  Limits: lines of at most 2^39 64-bit words (A < 1.7e13), else
 * RPS_EINVAL; at most 16,384 tiles per launch, which the call's batching
 * keeps to.  Returns after the survivors are written, with
 * *survivor_count set.  When they exceed `capacity` the call returns
 * RPS_EOVERFLOW and *survivor_count survivors of the request (a subset)
 * are in the buffers.  Every call on one device shares its scratch array:
 * calls that overlap on one GPU (from two threads or contexts) race. */
int rps_sieve_into(rps_ctx *ctx, const rps_request *request, int64_t *numerators,
                   int32_t *denominators, uint64_t capacity, uint64_t *survivor_count);

/* Checks a request without touching a device: RPS_OK, or RPS_EINVAL with the
 * reason rps_sieve would give written to `message` (NUL-terminated, at most
 * message_len bytes; may be NULL).  No counterpart in the reference, whose
 * checks run inside run_modular_sieve (src/sieve_plan.cpp:55-58); a server
 * that batches many requests per launch uses it to refuse a bad request on
 * admission instead of failing the whole launch. */
int rps_check_request(const rps_request *request, char *message, size_t message_len);

/* Survivor sink of a many-curve request: request[i] is the index into the
 * request array of survivor i.  Same validity, order and cancel rules as
 * rps_sink. */
typedef int (*rps_multi_sink)(void *user, const int32_t *request,
                              const int64_t *numerators,
                              const int32_t *denominators, size_t count);

/* Sieves many requests (curves with their own coefficients, bounds, ranges
 * and modes) in one persistent launch: workgroups pull (curve, denominator
 * block) and (curve, line chunk) tiles of all requests from a device work
 * counter, so short searches share the GPU instead of each paying a launch
 * and a synchronisation.  With the variant left to the tuning table,
 * requests whose lines take wide-row tiles (one workgroup per CU) get a
 * launch of their own; more than 8,192 requests are cut into launches of at
 * most 8,192 (each keeps one 1.33 MB residue table per curve on the device).  Each request's survivor set equals what rps_sieve
 * gives for it alone.  A request whose survivors overflow the buffer during
 * the shared launch is searched again alone (stats->requeued counts them).
 * No counterpart in the reference, which searches one curve per process
 * (src/main.cpp:48-85); it is the batched form of the same
 * run_modular_sieve contract (include/sieve.hpp:43-47). */
int rps_sieve_many(rps_ctx *ctx, const rps_request *requests,
                   size_t n_requests, rps_multi_sink sink, void *user,
                   rps_stats *stats);

/* Tuning and test knobs; returns RPS_EINVAL for an unknown name or a value
 * out of range.  No counterpart in the reference, whose knobs were
 * compile-time macros (src/sieve_plan.hpp:10-22, src/sieve.cu:15-17):
 *   "survivor_capacity"  survivors per flush buffer, >= 16384 (default 2^24)
 *   "batch_size"         denominators per launch, 1..1048576 (default 65536;
 *                        until set, rps_tuned's choice where it has one)
 *   "variant"            kernel variant index, 0..rps_variant_count()-1, or
 *                        -1 (default): chosen per request by rps_tuned
 *   "run_log2"           row tiles take whole lines up to 2^run_log2 words,
 *                        10..24, or -1 (default): rps_tuned's, or a forced
 *                        variant's own
 *   "min_tiles"          split denominators into word chunks until at least
 *                        this many tiles per launch exist (default 2048)
 *   "direct_words"       lines of at most this many 64-bit words use the
 *                        short-line kernel, 0..2048 (default 2048; 0 = never;
 *                        rps_sieve_many takes at most 1152 of it)
 *   "primitive_only"     1: the device drops survivors with gcd(a, b) > 1
 *                        before they are copied (they can never be points;
 *                        the exact stage, src/point_search.cpp:48-60, skips
 *                        them).  Default 0: the raw modular survivor set.
 *   "launch_timeout_ms"  longest wait for one kernel launch, 1..2^40 ms
 *                        (default 600,000): a launch still running then
 *                        fails the call with RPS_EDEVICE naming it, and the
 *                        context refuses further work (close it; its device
 *                        memory stays allocated until the process ends)
 *   "table_slots"        residue tables (1.33 MB each) kept on the device
 *                        between rps_sieve_many calls, least recently used
 *                        replaced, 1..65536 (default 256); a call with more
 *                        distinct curves grows the cache to fit them
 */
int rps_set_option(rps_ctx *ctx, const char *name, int64_t value);

/* Number of compiled kernel variants and a short name for each, e.g.
 * "q_u2_f12_w8" (q_: late primes through a compaction queue; words per lane,
 * dense primes, waves per workgroup).  Variant 0 is the default. */
int rps_variant_count(void);
const char *rps_variant_name(int variant);

/* The variant the autotuner's table picks for lines of `words` 64-bit words
 * (used when the "variant" option is -1, the default).  No counterpart in the
 * reference, whose kernel has one configuration. */
int rps_tuned_variant(uint64_t words);

/* The autotuner's full choice for lines of `words` words on the plain path
 * (many == 0: rps_sieve and its streaming forms) or for many-curve launches
 * (many != 0): variant, run length (row tiles take whole lines up to
 * 2^run_log2 words) and denominators per launch (0: the "batch_size"
 * option's; many-curve launches have none).  Used when the "variant",
 * "run_log2" and "batch_size" options are left at their defaults.  The
 * reference's counterpart is its tuning script's BLOCK / NUM_PRIMES /
 * INITIAL_PRIMES sweep (tests/tune_cuda.py:90-158), whose result is compiled
 * in by hand. */
int rps_tuned(uint64_t words, int many, int *variant, int *run_log2, int *batch);

/* Shape of a compiled variant: 64-bit words per lane per iteration, dense
 * (row-table) primes, waves per workgroup, default run length, and whether
 * its rows hold 16-byte elements (one workgroup per CU).  RPS_EINVAL for an
 * index out of range.  No counterpart in the reference. */
int rps_variant_info(int variant, int *words_per_lane, int *dense_primes, int *waves,
                     int *run_log2, int *wide);

/* Device enumeration for --list-devices (include/gpu_devices.hpp:20-21,
 * src/gpu_devices.cpp:55-97).  rps_device_info writes a NUL-terminated name
 * and PCI bus id; any output pointer may be NULL. */
int rps_device_count(int *count);
int rps_device_info(int device, char *name, size_t name_len, char *pci,
                    size_t pci_len, int *compute_units,
                    uint64_t *memory_bytes, int *major, int *minor);

/* Timing of the dominant kernel over the last rps_sieve call on `ctx`,
 * measured with HIP events on the stream the kernel runs on: total device
 * milliseconds of all sieve launches, launch count, and the number of
 * numerator sites those launches covered ((2A+1) x denominators). */
int rps_last_kernel_timing(const rps_ctx *ctx, double *kernel_ms,
                           uint64_t *launches, double *sites);

#ifdef __cplusplus
}
#endif

#endif /* RATPOINTS_SIEVE_H */
