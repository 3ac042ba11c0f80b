#!/bin/bash
# Runs the CPU tests (pytest -m "not gpu") against the library's host code and the oracle built with
# AddressSanitizer and UndefinedBehaviorSanitizer. CPU only: no GPU is used.
set -euo pipefail
REPO="$(cd "$(dirname "$0")/.." && pwd)"
JOBS="${JOBS:-8}"
make -s -j"$JOBS" -C "$REPO/particle-stabilizer_amd" asan all
make -s -C "$REPO/oracle" asan
RT="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
export PSX_LIBRARY="$REPO/particle-stabilizer_amd/lib_asan/libparticle_stabilizer_hip.so"
export PSX_ORACLE_LIBRARY="$REPO/oracle/_build_asan/libpsx_oracle.so"
# leaks of the interpreter and of the GPU runtime's one-time set-up are not this code's; every
# error report fails the run
export ASAN_OPTIONS="detect_leaks=0:detect_odr_violation=0:abort_on_error=1:halt_on_error=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
cd "$REPO"
LD_PRELOAD="$RT" python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
