"""Builds libmallocmc_hip.so in-tree with hipcc for gfx950 (and the CPU oracle).

Usage: python -m mallocmc_amd.build [--verbose]
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
LIB_NAME = "libmallocmc_hip.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

SOURCES = [
    "mallocmc_capi.hip",
    "mallocmc_harness.hip",
    "mallocmc_batched.hip",
    "mallocmc_serve.hip",
    "mallocmc_snapshot.hip",
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions are never wanted on this path
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link the shared library."""
    cc = hipcc()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with tempfile.TemporaryDirectory(prefix="mmc_build_") as tmp:
        objs = [os.path.join(tmp, s.replace(".hip", ".o")) for s in srcs]
        cmds = [[cc, *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, s), "-o", obj]
                for s, obj in zip(srcs, objs)]
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        # one hipcc per source, at most four at once (the batch kernels'
        # file takes most of the time)
        with concurrent.futures.ThreadPoolExecutor(max_workers=4) as pool:
            for r in pool.map(lambda c: subprocess.run(c, check=True), cmds):
                pass
        out_tmp = LIB_PATH + ".tmp"
        cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", out_tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out_tmp, LIB_PATH)
    return LIB_PATH


CPP_DIR = os.path.join(REPO_DIR, "tests", "cpp")
CPP_PROGRAMS = ["example_arrays", "example_slots", "verify_heap", "thread_malloc", "custom_policies",
                "udiv24_exhaustive"]


def build_cpp_programs(verbose: bool = False) -> list:
    """C++ programs written against the drop-in headers (include/mallocMC),
    linked with libmallocmc_hip.so; outputs in tests/cpp/_build/."""
    cc = hipcc()
    out_dir = os.path.join(CPP_DIR, "_build")
    os.makedirs(out_dir, exist_ok=True)
    outs = []
    for prog in CPP_PROGRAMS:
        src = os.path.join(CPP_DIR, prog + ".hip")
        out = os.path.join(out_dir, prog)
        cmd = [cc, *[f for f in HIPCC_FLAGS if f != "-fPIC"], "-I", INCLUDE, src, "-o", out + ".tmp",
               "-L", PKG_DIR, "-lmallocmc_hip", "-Wl,-rpath," + PKG_DIR + ",-rpath,$ORIGIN/../../../mallocmc_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_oracle(verbose: bool = False) -> None:
    oracle_dir = os.path.join(REPO_DIR, "oracle")
    cmd = ["make", "-s", "-C", oracle_dir]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    v = "--verbose" in sys.argv
    build_oracle(v)
    print(build(v))
    print(build_cpp_programs(v))
