#!/bin/bash
# Host-side AddressSanitizer + UndefinedBehaviorSanitizer run (no GPU):
# builds the oracle and the runner/C++ API with -fsanitize=address,undefined
# (gcc / g++ only) and runs the CPU tests that exercise them against those
# builds.  Any sanitizer report fails the run.
set -e
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
make -s -C "$ROOT/oracle" sanitize
make -s -C "$ROOT/tomos_amd/csrc" sanitize
ASAN_LIB="$(gcc -print-file-name=libasan.so)"
UBSAN_LIB="$(gcc -print-file-name=libubsan.so)"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
# the oracle is loaded into python: preload the sanitizer runtimes
TOMOS_ORACLE_LIB="$ROOT/oracle/_san/libtomos_oracle_san.so" LD_PRELOAD="$ASAN_LIB:$UBSAN_LIB" \
    python -m pytest "$ROOT/tests" -q -m "not gpu" -p no:cacheprovider -k "oracle" "$@"
# the runner is its own (sanitized) executable
TOMOS_RUNNER="$ROOT/tomos_amd/csrc/_san/tomos_lsqr_runner_san" \
    python -m pytest "$ROOT/tests/test_runner_io.py" -q -p no:cacheprovider "$@"
