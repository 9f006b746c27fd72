"""Writes reference_survivor_counts.json from the result files the reference
repository holds (read as JSON data; nothing of the reference is executed).

Sources, relative to the reference repository:
  docs/RTX4090-NUM-PRIMES.json, docs/RTX4090-SCREEN.json   record curve
  experiments/elliptic-curves/results/rank31-gpu-fixed-*.json  cubics
  tests/test_sieve_streaming.cpp:74-80, tests/test_multi_gpu.py:122-130
                                                       dense 1 + P x case
Only rows produced with the default 32-prime sieve are kept, since the
survivor count depends on the number of primes.

Usage: python tests/golden/extract_reference_counts.py [/root/reference]
"""

import glob
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
RECORD = "247747600 -985905640 567207969 2396040466 52485681 -470135160 82342800"


def main(ref):
    rows = []
    primes = [n for n in range(2, 512) if all(n % d for d in range(2, math.isqrt(n) + 1))]
    rows.append({"name": "dense-1-P", "coefficients": f"1 {math.prod(primes[-32:])}",
                 "bound": 5000, "first": 1, "last": 2, "square": False,
                 "survivors": 20002, "exact": 1,
                 "source": "tests/test_sieve_streaming.cpp:74-80; tests/test_multi_gpu.py:122-130"})
    for name, path in (("num-primes", "docs/RTX4090-NUM-PRIMES.json"),
                       ("screen", "docs/RTX4090-SCREEN.json")):
        data = json.load(open(os.path.join(ref, path)))
        seen = set()
        for v in data["variants"]:
            if v.get("num_primes") != 32:
                continue
            key = (v["height"], v["denominator_max"])
            if key in seen:
                continue
            seen.add(key)
            rows.append({"name": f"record-{name}-{v['height']}-{v['denominator_max']}",
                         "coefficients": RECORD, "bound": v["height"], "first": 1,
                         "last": v["denominator_max"], "square": False,
                         "survivors": v["modular_survivors"], "exact": v["exact_survivors"],
                         "output_sha256": v.get("output_sha256"), "source": path})
    for path in sorted(glob.glob(os.path.join(
            ref, "experiments/elliptic-curves/results/rank31-gpu-fixed-*.json"))):
        data = json.load(open(path))
        rel = os.path.relpath(path, ref)
        cmd = data.get("command")
        if cmd:
            coeffs = cmd[1]
            bound = int(cmd[2])
            first = int(cmd[cmd.index("-dl") + 1])
            last = int(cmd[cmd.index("-du") + 1])
        else:
            coeffs = data["coefficients"]
            bound = int(data["height"])
            first = int(data["denominator_min"])
            last = int(data["denominator_max"])
        metrics = data.get("metrics", {})
        surv = data.get("modular_survivors", metrics.get("survivors"))
        exact = data.get("exact_survivors", metrics.get("exact_survivors"))
        rows.append({"name": os.path.basename(path)[:-5], "coefficients": coeffs,
                     "bound": bound, "first": first, "last": last, "square": False,
                     "survivors": int(surv), "exact": int(exact), "source": rel})
    out = os.path.join(HERE, "reference_survivor_counts.json")
    with open(out, "w") as fh:
        json.dump(rows, fh, indent=1)
        fh.write("\n")
    print(f"wrote {len(rows)} rows to {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
