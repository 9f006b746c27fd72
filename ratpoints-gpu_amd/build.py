"""Builds every native artefact of the package in-tree.

  lib/libratpoints_sieve.so   gfx950 HIP kernels + C ABI (hipcc)
  bin/ratpoints_gpu           C++ drop-in CLI over the C ABI (GMP from the
                              first of /usr, /opt/conda that has gmpxx.h)

Run ``python ratpoints-gpu_amd/build.py`` or ``__graft_entry__.build()``.
Every call recompiles; nothing is cached.
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions stay disabled in every HIP translation unit.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]

HOST_SOURCES = [
    "command_line.cpp", "exact_polynomial.cpp", "output_formatter.cpp",
    "point_search.cpp", "sieve_shim.cpp", "gpu_devices.cpp", "main.cpp",
]


def gmp_prefix():
    for prefix in ("/usr", "/usr/local", "/opt/conda"):
        if os.path.exists(os.path.join(prefix, "include", "gmpxx.h")):
            return prefix
    return None


def run(cmd, cwd=ROOT):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")
    return proc.stdout


def build_library():
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    out = os.path.join(HERE, "lib", "libratpoints_sieve.so")
    tmp = out + ".tmp"
    objs = []
    jobs = []
    for src in ("sieve_kernels.hip", "rps_abi.cpp"):
        obj = os.path.join(HERE, "lib", src.rsplit(".", 1)[0] + ".o")
        extra = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append([HIPCC, *HIP_FLAGS, "-I" + os.path.join(ROOT, "include"),
                     "-I" + os.path.join(HERE, "csrc"), *extra, "-c",
                     os.path.join(HERE, "csrc", src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(len(jobs)) as pool:
        list(pool.map(run, jobs))
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    for obj in objs:
        os.remove(obj)
    return out


def build_cli():
    host = os.path.join(HERE, "host")
    sources = [os.path.join(host, s) for s in HOST_SOURCES]
    if not all(os.path.exists(s) for s in sources):
        return None
    prefix = gmp_prefix()
    if prefix is None:
        raise RuntimeError("GMP (gmpxx.h) not found under /usr, /usr/local or /opt/conda")
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    out = os.path.join(HERE, "bin", "ratpoints_gpu")
    libdir = os.path.join(HERE, "lib")
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", "-pthread",
           "-I" + host, "-I" + os.path.join(ROOT, "include")]
    if prefix != "/usr":
        cmd += ["-I" + os.path.join(prefix, "include")]
    cmd += [*sources, "-o", out + ".tmp", "-L" + libdir, "-lratpoints_sieve",
            "-Wl,-rpath,$ORIGIN/../lib"]
    if prefix != "/usr":
        cmd += ["-L" + os.path.join(prefix, "lib"), "-Wl,-rpath," + os.path.join(prefix, "lib")]
    cmd += ["-lgmpxx", "-lgmp"]
    run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    """The oracle is test infrastructure; it is built beside the product so
    tests, smoke() and the bench's CPU baseline find it."""
    oracle = os.path.join(ROOT, "oracle")
    run(["make", "-B", "-C", oracle])
    if os.path.isdir("/root/reference/src"):
        run(["make", "-B", "-C", oracle, "ref"])


def build_all():
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    lib = build_library()
    cli = build_cli()
    build_oracle()
    return lib, cli


if __name__ == "__main__":
    lib, cli = build_all()
    print("built", lib)
    if cli:
        print("built", cli)
    sys.exit(0)
