#!/bin/bash
# Runs the CPU test suite (pytest -m "not gpu") against AddressSanitizer +
# UndefinedBehaviorSanitizer builds of the oracle and of the host programs/helpers.
# The oracle is loaded into python through ctypes, so libasan is preloaded; leak
# detection is off (python itself does not free everything at exit).
set -e
ROOT=$(cd "$(dirname "$0")/.." && pwd)
make -s -C "$ROOT/oracle" sanitize
make -s -C "$ROOT/tti-fdmodeling_amd" sanitize-host
export ORACLE_BUILD_DIR="$ROOT/oracle/_build/san"
export FDMOD_HOST_BIN="$ROOT/tti-fdmodeling_amd/_build/san"
export FDMOD_HOST_LIB="$ROOT/tti-fdmodeling_amd/_build/san/libfdmod_host.so"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)" \
  python3 -m pytest "$ROOT/tests" -q -m "not gpu" -p no:cacheprovider "$@"
