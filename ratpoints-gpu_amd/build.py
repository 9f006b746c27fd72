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

HOST_COMMON = [
    "exact_polynomial.cpp", "point_search.cpp", "sieve_shim.cpp", "gpu_devices.cpp",
]
HOST_SOURCES = HOST_COMMON + ["command_line.cpp", "output_formatter.cpp", "main.cpp"]
SERVE_SOURCES = HOST_COMMON + ["serve.cpp", "serve_main.cpp"]


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


def build_cli(backend=None, out=None, sources=None, name="ratpoints_gpu", extra=()):
    """The CLI over the HIP library, or -- for tests only -- over another
    implementation of the C ABI given as a list of object files; `extra` adds
    source files by path."""
    host = os.path.join(HERE, "host")
    sources = [os.path.join(host, s) for s in (sources or HOST_SOURCES)] + list(extra)
    if not all(os.path.exists(s) for s in sources):
        return None
    prefix = gmp_prefix()
    if prefix is None:
        raise RuntimeError("GMP (gmpxx.h) not found under /usr, /usr/local or /opt/conda")
    out = out or os.path.join(HERE, "bin", name)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    libdir = os.path.join(HERE, "lib")
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-Wextra", "-pthread",
           "-I" + host, "-I" + os.path.join(ROOT, "include")]
    # The library directory relative to the binary, so the tree can move.
    origin_lib = "$ORIGIN/" + os.path.relpath(libdir, os.path.dirname(os.path.abspath(out)))
    rpath = "-Wl,-rpath," + origin_lib
    if prefix != "/usr":
        # GMP is linked by full path and found at run time through a
        # directory of symlinks to it alone, so the prefix's other libraries
        # (an older libstdc++ under /opt/conda) never shadow the system ones.
        gmp_dir = os.path.join(libdir, "gmp")
        os.makedirs(gmp_dir, exist_ok=True)
        for soname in ("libgmp.so.10", "libgmpxx.so.4"):
            link = os.path.join(gmp_dir, soname)
            if os.path.lexists(link):
                os.remove(link)
            os.symlink(os.path.join(prefix, "lib", soname), link)
        cmd += ["-I" + os.path.join(prefix, "include")]
        rpath += ",-rpath," + origin_lib + "/gmp"
        gmp_libs = [os.path.join(prefix, "lib", "libgmpxx.so"),
                    os.path.join(prefix, "lib", "libgmp.so")]
    else:
        gmp_libs = ["-lgmpxx", "-lgmp"]
    if backend is None:
        abi = ["-L" + libdir, "-lratpoints_sieve"]
    else:
        abi = list(backend)
        rpath = rpath.replace(origin_lib, libdir)
    cmd += [*sources, "-o", out + ".tmp", *abi, rpath, *gmp_libs]
    run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_serve(backend=None, out=None):
    """bin/ratpoints_serve: the serving loop's trace driver."""
    return build_cli(backend, out, SERVE_SOURCES, "ratpoints_serve")


def build_oracle():
    """The oracle is test infrastructure; it is built beside the product so
    tests, smoke() and the bench's CPU baseline find it."""
    oracle = os.path.join(ROOT, "oracle")
    run(["make", "-B", "-C", oracle])
    if os.path.isdir("/root/reference/src"):
        run(["make", "-B", "-C", oracle, "ref"])
    # The host CLI over the oracle's CPU test double of the C ABI, for the
    # host-logic tests (never shipped; see oracle/abi_oracle_backend.c).
    build_dir = os.path.join(oracle, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    for src in ("abi_oracle_backend.c", "sieve_oracle.c"):
        obj = os.path.join(build_dir, src[:-2] + ".o")
        run(["gcc", "-O2", "-std=c11", "-fPIC", "-pthread", "-I" + os.path.join(ROOT, "include"),
             "-c", os.path.join(oracle, src), "-o", obj])
        objs.append(obj)
    build_cli(backend=objs, out=os.path.join(build_dir, "ratpoints_gpu_oracle"))
    build_serve(backend=objs, out=os.path.join(build_dir, "ratpoints_serve_oracle"))


HOST_CLANG = os.environ.get("HOST_CLANG", "/opt/rocm/llvm/bin/clang++")


def build_emulator():
    """Test infrastructure: the library's own sources (kernels and C ABI)
    compiled for the host against the CPU wave emulator
    (tests/emu/_build/libratpoints_sieve_emu.so), which the CPU tests drive
    through the same C ABI.  Never loaded by the package."""
    emu = os.path.join(ROOT, "tests", "emu")
    out_dir = os.path.join(emu, "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libratpoints_sieve_emu.so")
    run([HOST_CLANG, "-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall",
         "-Wno-unknown-attributes", "-Wno-unused-function",
         "-I" + os.path.join(emu, "include"), "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(HERE, "csrc"),
         "-x", "c++", os.path.join(HERE, "csrc", "sieve_kernels.hip"),
         "-x", "none", os.path.join(HERE, "csrc", "rps_abi.cpp"),
         os.path.join(emu, "wave_emu.cpp"), "-shared", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_test_programs():
    """Test infrastructure: the C++ streaming-contract test over the HIP
    library (tests/cpp/_build/streaming_contract) and over the emulator
    (streaming_contract_emu), and the serving driver over the emulator."""
    emu_lib = build_emulator()
    cpp = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(cpp, "_build")
    contract = [os.path.join(cpp, "streaming_contract.cpp")]
    shim = ["exact_polynomial.cpp", "sieve_shim.cpp", "gpu_devices.cpp"]
    build_cli(sources=shim, extra=contract, out=os.path.join(out_dir, "streaming_contract"))
    build_cli(backend=[emu_lib], sources=shim, extra=contract,
              out=os.path.join(out_dir, "streaming_contract_emu"))
    build_serve(backend=[emu_lib],
                out=os.path.join(ROOT, "tests", "emu", "_build", "ratpoints_serve_emu"))


def build_all():
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    lib = build_library()
    cli = build_cli()
    build_serve()
    build_oracle()
    build_test_programs()
    return lib, cli


if __name__ == "__main__":
    lib, cli = build_all()
    print("built", lib)
    if cli:
        print("built", cli)
    sys.exit(0)
