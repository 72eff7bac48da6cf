#!/bin/bash
# The CPU test suite against AddressSanitizer + UndefinedBehaviorSanitizer builds of the
# library's host C++ (librkdg_san.so: mesh ingest and writer, basis, projection, config and
# .plt I/O) and of the oracle (liboracle_san.so).  CPU only; leaks are not checked (the
# Python interpreter keeps its allocations to the end by design).
set -e
ROOT=$(cd "$(dirname "$0")/.." && pwd)
make -s -C "$ROOT/rkdg_heter_calculate_amd/csrc" san
make -s -C "$ROOT/oracle" san
export LD_PRELOAD="$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so)"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
export RKDG_LIB="$ROOT/rkdg_heter_calculate_amd/librkdg_san.so"
export ORACLE_LIB="$ROOT/oracle/liboracle_san.so"
cd "$ROOT"
python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
