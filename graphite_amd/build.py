"""Builds libgraphite_hip.so in-tree: HIP sources with hipcc for gfx950, host C++ with g++.

Usage: python -m graphite_amd.build [-v] [--sanitize]
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgraphite_hip.so")
CLI = os.path.join(HERE, "graphite")                 # the case-file driver executable
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
# No packed-FP32 instructions in device code (they have given wrong results on these GPUs);
# hipcc's host pass warns that it does not know the feature, which is harmless.
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                      "-munsafe-fp-atomics"]
CXX_FLAGS = COMMON + ["-march=x86-64-v3"]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    hip, cpp = _sources()
    jobs = []
    for f in hip:
        o = os.path.join(OBJ, f + ".o")
        jobs.append(([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, f), "-o", o] + HIP_FLAGS, o))
    for f in cpp:
        o = os.path.join(OBJ, f + ".o")
        jobs.append((["g++", "-c", os.path.join(CSRC, f), "-o", o] + CXX_FLAGS, o))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0], verbose), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + [o for _, o in jobs], verbose)
    os.replace(tmp, LIB)
    _run(["g++", os.path.join(HERE, "cli", "graphite_main.cpp"), "-o", CLI + ".tmp", f"-I{INCLUDE}"] + CXX_FLAGS +
         [f"-L{HERE}", "-lgraphite_hip", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(CLI + ".tmp", CLI)
    return LIB


SAN_DIR = os.path.join(OBJ, "san")
SAN_LIB = os.path.join(SAN_DIR, "libgraphite_hip.so")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]


def build_sanitized(verbose: bool = False) -> str:
    """The library with its host C++ sources (connectivity, generators, file I/O, case driver)
    compiled by g++ under AddressSanitizer and UndefinedBehaviorSanitizer, linked with the
    regular HIP objects; for CPU test runs with libasan preloaded.  Returns its path."""
    build(verbose)
    os.makedirs(SAN_DIR, exist_ok=True)
    hip, cpp = _sources()
    objs = [os.path.join(OBJ, f + ".o") for f in hip]
    for f in cpp:
        o = os.path.join(SAN_DIR, f + ".o")
        _run(["g++", "-c", os.path.join(CSRC, f), "-o", o] + CXX_FLAGS + SAN_FLAGS, verbose)
        objs.append(o)
    rocm_lib = os.path.join(ROCM, "lib")
    _run(["g++", "-shared", "-o", SAN_LIB + ".tmp"] + objs + SAN_FLAGS +
         [f"-L{rocm_lib}", "-lamdhip64", f"-Wl,-rpath,{rocm_lib}"], verbose)
    os.replace(SAN_LIB + ".tmp", SAN_LIB)
    return SAN_LIB


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized(verbose="-v" in sys.argv))
    else:
        print(build(verbose="-v" in sys.argv))
