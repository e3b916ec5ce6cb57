"""Sanitizer runs of the library's host code (CPU only, no GPU needed).

Usage: python axia-fem_amd/sanitize.py address|thread [pytest args...]

Builds build/lib_<kind>/libaxia_gpu_<kind>.so with the host half of every source
instrumented (AddressSanitizer + UndefinedBehaviorSanitizer, or ThreadSanitizer; the
device code is compiled as usual), then runs the CPU tests that drive the host-side entry points
-- the symbolic phase on host threads (axg_pattern_host), the INP reader with its
threaded NODE / ELEMENT parsing (axg_deck_parse, axg_deck_arrays), the face-pressure
loads and the threaded DAT / FRD writers -- with that build loaded (AXIA_GPU_LIB)
and the sanitizer runtime preloaded into the Python interpreter. Any report fails
the run (halt_on_error).
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
import build as B  # noqa: E402

# the host pass only: each sanitizer flag follows -Xarch_host
FLAGS = {
    "address": (["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all"],
                ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"],
                "libclang_rt.asan-x86_64.so"),
    "thread": (["-Xarch_host", "-fsanitize=thread", "-Xarch_host", "-fno-omit-frame-pointer"],
               ["-Xarch_host", "-fsanitize=thread"],
               "libclang_rt.tsan-x86_64.so"),
}
TESTS = ["tests/test_host_abi.py", "tests/test_host_deck.py", "tests/test_writers.py",
         "tests/test_oracle_pins.py"]


def runtime(name):
    hits = glob.glob(os.path.join(os.path.dirname(B.hipcc()), "..", "llvm", "lib", "clang", "*",
                                  "lib", "linux", name))
    hits += glob.glob(os.path.join("/opt/rocm/llvm/lib/clang/*/lib/linux", name))
    if not hits:
        raise SystemExit(f"sanitizer runtime {name} not found")
    return os.path.realpath(hits[0])


def main():
    kind = sys.argv[1] if len(sys.argv) > 1 else "address"
    if kind not in FLAGS:
        raise SystemExit("usage: sanitize.py address|thread [pytest args]")
    cflags, lflags, rt = FLAGS[kind]
    lib = os.path.join(HERE, "build", f"lib_{kind}", f"libaxia_gpu_{kind}.so")
    B.build(lib=lib, obj_dir=os.path.join(HERE, "build", f"obj_{kind}"), extra=cflags,
            link_extra=lflags)
    env = dict(os.environ)
    env["AXIA_GPU_LIB"] = lib
    env["LD_PRELOAD"] = runtime(rt)
    # the interpreter itself is not instrumented: no leak report for its allocations
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:report_signal_unsafe=0"
    cmd = [sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
    cmd += sys.argv[2:] or TESTS
    print(" ".join(cmd), flush=True)
    sys.exit(subprocess.run(cmd, cwd=ROOT, env=env).returncode)


if __name__ == "__main__":
    main()
