"""Runs the CPU test suite against AddressSanitizer + UndefinedBehaviorSanitizer builds of the
oracle (gcc) and of the library's host C++ sources (g++), with libasan preloaded into Python.

  python tests/run_sanitized.py [pytest args...]

Leak detection is off (the Python interpreter itself is not leak-clean); any other report
aborts the test process, which the run reports as a failure.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main(argv):
    from graphite_amd.build import build_sanitized
    lib = build_sanitized()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_san/liboracle.so"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True, check=True).stdout.strip()
    env = dict(os.environ, GRAPHITE_HIP_LIB=lib, GRAPHITE_ORACLE_LIB=os.path.join(ROOT, "oracle", "_san", "liboracle.so"),
               LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
    return subprocess.run(cmd + argv, env=env, cwd=ROOT).returncode


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
