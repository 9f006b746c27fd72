"""Autotuner: picks the row-tile kernel variant per line-length bucket.

Runs on a GPU.  For each bucket of line lengths (64-bit words per
denominator line) it times every compiled variant on the plain path over
one representative line length (a planted sextic, about 3e11 sites, runs of
lines on), and writes ``csrc/tuning_table.inc`` -- the table the library
uses to choose a variant for a request when the caller has not forced one
with ``rps_set_option("variant")``.  Rebuild the library afterwards.

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


def load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module
    spec.loader.exec_module(module)
    return module


def measure(sites_per_point=3e11):
    pkg = load("ratpoints_gpu_amd", os.path.join(HERE, "__init__.py"))
    wl = load("ratpoints_gpu_amd_workloads", os.path.join(HERE, "workloads.py"))
    coeffs = [str(c) for c in wl.sextic_a5e6_curve()[0]]
    names = pkg.variants()
    rows = []
    with pkg.SieveContext(0) as ctx:
        for upper, words in BUCKETS:
            bound = 32 * words - 32
            lines = max(65536, int(sites_per_point / (2 * bound + 1)) // 65536 * 65536)
            rates = {}
            for v, name in enumerate(names):
                ctx.set_option("variant", v)
                ctx.run(coeffs, bound, 1, 65536, False, lambda n, d: False)
                ctx.run(coeffs, bound, 1, lines, False, lambda n, d: False)
                ms, _, sites = ctx.kernel_timing()
                rates[name] = sites / ms / 1e6
            best = max(rates, key=rates.get)
            row = {"max_words": upper, "words": words, "best": best,
                   "gsites": {k: round(r, 1) for k, r in rates.items()}}
            print(json.dumps(row), flush=True)
            rows.append(row)
    return rows


def write_table(rows, path=TABLE):
    lines = [
        "// Generated by ratpoints-gpu_amd/autotune.py on one MI355X (%s);" % time.strftime("%Y-%m-%d"),
        "// kernel Gsite/s of every variant per line-length bucket, the fastest chosen.",
        "// {largest words per line of the bucket, variant name}",
    ]
    for row in rows:
        rates = ", ".join("%s %.0f" % (k, v) for k, v in sorted(row["gsites"].items()))
        lines.append("// <= %d words (timed at %d): %s" % (min(row["max_words"], 10 ** 12),
                                                           row["words"], rates))
    for row in rows:
        upper = "~0ULL" if row["max_words"] >= 1 << 62 else "%dULL" % row["max_words"]
        lines.append('{%s, "%s"},' % (upper, row["best"]))
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
