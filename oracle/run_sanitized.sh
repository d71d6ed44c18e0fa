#!/bin/bash
# Runs the oracle's CPU tests against the oracle built with AddressSanitizer
# and UndefinedBehaviorSanitizer (make -C oracle sanitize). Python itself is
# not instrumented: the ASan runtime is preloaded, and leak reports (which
# would be the interpreter's) are off. Any error aborts the run.
#   bash oracle/run_sanitized.sh [pytest arguments]
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(dirname "$HERE")"
make -C "$HERE" sanitize
ASAN="$(g++ -print-file-name=libasan.so)"
UBSAN="$(g++ -print-file-name=libubsan.so)"
cd "$ROOT"
LD_PRELOAD="$ASAN:$UBSAN" ASAN_OPTIONS=detect_leaks=0:abort_on_error=1 UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
  AVBD_ORACLE_LIB="$HERE/liboracle_san.so" \
  python -m pytest tests -q -m "not gpu" -k "oracle or sceneio or contact_model" -p no:cacheprovider "$@"
