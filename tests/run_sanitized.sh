#!/bin/sh
# Builds the oracle (oracle/oracle.c, oracle/adjoint_ref.cpp) and the host scene library
# (specrel_amd/csrc/scene/scene.cpp) with AddressSanitizer and UndefinedBehaviorSanitizer into
# tests/_san/ and runs the CPU test suite with those builds loaded in their place; before that,
# runs both under ThreadSanitizer in a driver of their own. Host code only: nothing here runs on
# a GPU. Usage: sh tests/run_sanitized.sh [extra pytest args]
set -e
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
OUT="$ROOT/tests/_san"
mkdir -p "$OUT"
SAN="-O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined"
FP="-ffp-contract=off -fno-fast-math"
gcc -std=c11 $SAN $FP -fPIC -pthread -c -o "$OUT/oracle.o" "$ROOT/oracle/oracle.c"
g++ -std=c++17 $SAN $FP -fPIC -pthread -c -o "$OUT/adjoint_ref.o" "$ROOT/oracle/adjoint_ref.cpp"
g++ $SAN -shared -pthread -o "$OUT/liboracle.so" "$OUT/oracle.o" "$OUT/adjoint_ref.o" -lm
g++ -std=c++17 $SAN $FP -fPIC -shared -pthread -I"$ROOT/include" -o "$OUT/libspecrel_scene.so" \
    "$ROOT/specrel_amd/csrc/scene/scene.cpp"
rm -f "$OUT/oracle.o" "$OUT/adjoint_ref.o"
# ThreadSanitizer over the same code's thread pools, in a driver of its own (tests/tsan_driver.cpp).
TSAN="-O1 -g -fno-omit-frame-pointer -fsanitize=thread"
gcc -std=c11 $TSAN $FP -pthread -c -o "$OUT/oracle_tsan.o" "$ROOT/oracle/oracle.c"
g++ -std=c++17 $TSAN $FP -pthread -I"$ROOT/include" -o "$OUT/tsan_driver" "$ROOT/tests/tsan_driver.cpp" \
    "$ROOT/specrel_amd/csrc/scene/scene.cpp" "$OUT/oracle_tsan.o" -lm
rm -f "$OUT/oracle_tsan.o"
TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1" "$OUT/tsan_driver"
# Python itself is not instrumented: the sanitizer runtime is loaded first, leak checking is off
# (the interpreter's own allocations at exit are not this code's), any report ends the run.
export LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)${LD_PRELOAD:+ $LD_PRELOAD}"
# Reports go to files as well, since a test's captured output is lost when the run stops.
rm -f "$OUT"/report.*
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1:detect_odr_violation=0:log_path=$OUT/report.asan"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:log_path=$OUT/report.ubsan"
export SPECREL_ORACLE_LIB="$OUT/liboracle.so"
export SPECREL_SCENE_LIB="$OUT/libspecrel_scene.so"
cd "$ROOT"
rc=0
python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@" || rc=$?
for f in "$OUT"/report.*; do
    [ -e "$f" ] || continue
    echo "== $f"; cat "$f"; rc=1
done
exit $rc
