"""Autotuner: picks the launch shape per line-length bucket.

Runs on a GPU.  For each bucket of line lengths (64-bit words per
denominator line) it times, over one representative line length of a
planted sextic (about 3e11 sites per point):

  * on the plain path (rps_sieve): every compiled variant -- dense-prime
    count, words per lane, waves per workgroup, row width -- at three run
    lengths each (whole lines per row tile up to 2^run_log2 words: the
    variant's own and one step either side), by kernel time; then, for the
    fastest, three batch sizes (denominators per launch) by wall time;
  * on the many-curve path (rps_sieve_many, the search split into 64
    requests of equal shape in one call): every variant at the same run
    lengths, by kernel time;

and writes ``csrc/tuning_table.inc``, the table the library uses when the
caller leaves the "variant", "run_log2" and "batch_size" options at their
defaults (``rps_tuned``).  Rebuild the library afterwards.  The reference's
counterpart is tests/tune_cuda.py:90-158 (a BLOCK / NUM_PRIMES /
INITIAL_PRIMES sweep whose result is compiled in by hand).

    python ratpoints-gpu_amd/autotune.py            # measure and write
    python ratpoints-gpu_amd/autotune.py --dry-run  # measure and print
    python ratpoints-gpu_amd/autotune.py --run q_u1_f12_w8 --words 4096 [--many] [--lib SO]
                                     # one timed sieve of one variant (profiling,
                                     # or comparing another build of the library)
"""

import argparse
import importlib.util
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "csrc", "tuning_table.inc")
# Upper ends of the buckets (words per line) and the length timed for each.
BUCKETS = [(1536, 1024), (3072, 2048), (6144, 4096), (12288, 8192),
           (24576, 16384), (49152, 32768), (98304, 65536), (1 << 62, 156251)]
BATCHES = (32768, 65536, 131072)
MANY_REQUESTS = 64


def load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module
    spec.loader.exec_module(module)
    return module


def run_lengths(pkg, v):
    own = pkg.variant_info(v)["run_log2"]
    return [r for r in (own - 1, own, own + 1) if 10 <= r <= 24]


def many_requests(coeffs, bound, lines):
    per = max(1, lines // MANY_REQUESTS)
    return [{"coefficients": coeffs, "bound": bound, "first": 1 + k * per,
             "last": (k + 1) * per} for k in range(MANY_REQUESTS)]


def measure(sites_per_point=3e11):
    pkg = load("ratpoints_gpu_amd", os.path.join(HERE, "__init__.py"))
    wl = load("ratpoints_gpu_amd_workloads", os.path.join(HERE, "workloads.py"))
    coeffs = [str(c) for c in wl.sextic_a5e6_curve()[0]]
    names = pkg.variants()
    rows = []
    no_sink = lambda n, d: False  # noqa: E731
    with pkg.SieveContext(0) as ctx:
        for upper, words in BUCKETS:
            bound = 32 * words - 32
            lines = max(65536, int(sites_per_point / (2 * bound + 1)) // 65536 * 65536)
            row = {"max_words": upper, "words": words, "plain": {}, "many": {}, "batch": {}}
            ctx.set_option("batch_size", 65536)
            for v, name in enumerate(names):
                ctx.set_option("variant", v)
                for r in run_lengths(pkg, v):
                    ctx.set_option("run_log2", r)
                    ctx.run(coeffs, bound, 1, 65536, False, no_sink)
                    ctx.run(coeffs, bound, 1, lines, False, no_sink)
                    ms, _, sites = ctx.kernel_timing()
                    row["plain"]["%s/%d" % (name, r)] = round(sites / ms / 1e6, 1)
                    reqs = many_requests(coeffs, bound, lines)
                    ctx.run_many(reqs[:2], lambda q, n, d: False)
                    ctx.run_many(reqs, lambda q, n, d: False)
                    ms, _, sites = ctx.kernel_timing()
                    row["many"]["%s/%d" % (name, r)] = round(sites / ms / 1e6, 1)
            best = max(row["plain"], key=row["plain"].get)
            name, r = best.split("/")
            ctx.set_option("variant", names.index(name))
            ctx.set_option("run_log2", int(r))
            for batch in BATCHES:
                ctx.set_option("batch_size", batch)
                ctx.run(coeffs, bound, 1, lines, False, no_sink)
                t0 = time.perf_counter()
                ctx.run(coeffs, bound, 1, lines, False, no_sink)
                wall = time.perf_counter() - t0
                _, _, sites = ctx.kernel_timing()
                row["batch"][str(batch)] = round(sites / wall / 1e9, 1)
            row["best"] = best
            row["best_batch"] = int(max(row["batch"], key=row["batch"].get))
            row["best_many"] = max(row["many"], key=row["many"].get)
            print(json.dumps(row), flush=True)
            rows.append(row)
            ctx.set_option("variant", -1)
            ctx.set_option("run_log2", -1)
    return rows


def write_table(rows, path=TABLE):
    lines = [
        "// Generated by ratpoints-gpu_amd/autotune.py on one MI355X (%s): kernel" % time.strftime("%Y-%m-%d"),
        "// Gsite/s of every variant/run_log2 per line-length bucket on the plain",
        "// path and in many-curve launches, and end-to-end Gsite/s of the plain",
        "// path's fastest shape per batch size; the fastest of each chosen.",
        "// {largest words per line of the bucket, plain variant, run_log2, batch,",
        "//  many-curve variant, run_log2}",
    ]
    for row in rows:
        upper = min(row["max_words"], 10 ** 12)
        for kind in ("plain", "many"):
            rates = ", ".join("%s %.0f" % (k, v) for k, v in sorted(row[kind].items()))
            lines.append("// <= %d words (timed at %d), %s: %s" % (upper, row["words"], kind,
                                                                  rates))
        rates = ", ".join("%s %.0f" % (k, v) for k, v in row["batch"].items())
        lines.append("// <= %d words, plain %s by batch: %s" % (upper, row["best"], rates))
    for row in rows:
        upper = "~0ULL" if row["max_words"] >= 1 << 62 else "%dULL" % row["max_words"]
        name, r = row["best"].split("/")
        many, mr = row["best_many"].split("/")
        lines.append('{%s, "%s", %s, %d, "%s", %s},' % (upper, name, r, row["best_batch"],
                                                        many, mr))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def run_one(variant, words, lines, many, lib=None):
    """One timed sieve of `lines` lines of `words` words on the planted
    sextic with one variant, on the plain path or (many) as 64 requests of
    equal shape in one rps_sieve_many call: the workload a profiler wraps.
    `lib`: another build of the library (same C ABI) to time instead."""
    pkg = load("ratpoints_gpu_amd", os.path.join(HERE, "__init__.py"))
    wl = load("ratpoints_gpu_amd_workloads", os.path.join(HERE, "workloads.py"))
    coeffs = [str(c) for c in wl.sextic_a5e6_curve()[0]]
    bound = 32 * words - 32
    handle = pkg.bind_library(lib) if lib else None
    with pkg.SieveContext(0, lib=handle) as ctx:
        names = [ctx.lib.rps_variant_name(i).decode() for i in range(ctx.lib.rps_variant_count())]
        ctx.set_option("variant", names.index(variant))
        if many:
            per = max(1, lines // 64)
            reqs = [{"coefficients": coeffs, "bound": bound, "first": 1 + k * per,
                     "last": (k + 1) * per} for k in range(64)]
            ctx.run_many(reqs[:2], lambda q, n, d: False)
            t0 = time.perf_counter()
            ctx.run_many(reqs, lambda q, n, d: False)
        else:
            ctx.run(coeffs, bound, 1, 65536, False, lambda n, d: False)
            t0 = time.perf_counter()
            ctx.run(coeffs, bound, 1, lines, False, lambda n, d: False)
        wall = time.perf_counter() - t0
        ms, launches, sites = ctx.kernel_timing()
    print(json.dumps({"variant": variant, "words": words, "lines": lines, "many": many,
                      "lib": lib or "default",
                      "kernel_ms": round(ms, 3), "launches": launches, "sites": sites,
                      "kernel_gsites": round(sites / ms / 1e6, 1),
                      "wall_gsites": round(sites / wall / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--dry-run", action="store_true")
    ap.add_argument("--out", default=TABLE, help="table path (default: csrc/tuning_table.inc)")
    ap.add_argument("--run", metavar="VARIANT",
                    help="run one variant at --words (for profilers) instead of tuning")
    ap.add_argument("--words", type=int, default=4096)
    ap.add_argument("--lines", type=int, default=262144)
    ap.add_argument("--many", action="store_true", help="--run through rps_sieve_many")
    ap.add_argument("--lib", help="--run on another build of the library")
    args = ap.parse_args()
    if args.run:
        run_one(args.run, args.words, args.lines, args.many, args.lib)
        return
    rows = measure()
    if not args.dry_run:
        write_table(rows, args.out)
        print("wrote", args.out)


if __name__ == "__main__":
    main()
