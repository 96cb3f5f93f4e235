#!/bin/bash
# Sanitizer run of the host code (CPU only, no device needed):
#  1. AddressSanitizer + UndefinedBehaviorSanitizer builds of the library's
#     host C++ and of the oracle, with the CPU test suite run on them;
#  2. a ThreadSanitizer build of the library's host C++ driven by
#     tests/sanitize_host.cpp (thread pool, concurrent simulator calls, the
#     request loop's host paths).
# Builds go to $OUT (default /tmp/polyq_sanitize).  ASan/UBSan use GCC's
# runtimes; TSan uses the ROCm clang's (GCC 11's libtsan misses the
# pthread_cond_clockwait interceptor that condition_variable::wait_for needs).
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=${OUT:-/tmp/polyq_sanitize}
LIBASAN=$(g++ -print-file-name=libasan.so)
LIBUBSAN=$(g++ -print-file-name=libubsan.so)

echo "== ASan + UBSan build"
python3 "$ROOT/polyq-v1.0_amd/build_lib.py" --sanitize address,undefined --out "$OUT/asan" > /dev/null
make -s -C "$ROOT/oracle" _san/liboracle.so
echo "== CPU tests under ASan + UBSan"
# Python itself is not instrumented: leak checking off, every UBSan report fatal
ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
POLYQ_AMD_PKG_DIR="$OUT/asan" POLYQ_ORACLE_LIB="$ROOT/oracle/_san/liboracle.so" LD_PRELOAD="$LIBASAN $LIBUBSAN" \
    python3 -m pytest "$ROOT/tests" -q -m "not gpu" -p no:cacheprovider

echo "== TSan build (ROCm clang and its TSan runtime)"
python3 "$ROOT/polyq-v1.0_amd/build_lib.py" --sanitize thread --out "$OUT/tsan" > /dev/null
/opt/rocm/llvm/bin/clang++ -O1 -g -std=c++17 -fsanitize=thread -pthread "$ROOT/tests/sanitize_host.cpp" -o "$OUT/tsan/sanitize_host" \
    -L "$OUT/tsan/polyq_amd" -lpolyq_amd -Wl,-rpath,"$OUT/tsan/polyq_amd" -Wl,-rpath,/opt/rocm/lib
echo "== host concurrency driver under TSan"
TSAN_OPTIONS=halt_on_error=1:second_deadlock_stack=1 "$OUT/tsan/sanitize_host"
echo "== sanitizer run clean"
