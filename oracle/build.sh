#!/bin/sh
# Builds the oracle's shared library next to its source (test infrastructure only).
# -O3 without the loop vectorizer: on the star loop gcc's vectorized -O3 code ran 45% slower than
# -O2, and this build is the fastest of the three (it is also the CPU baseline's).
set -e
cd "$(dirname "$0")"
gcc -std=c11 -O3 -fno-tree-vectorize -fPIC -ffp-contract=off -fno-fast-math -Wall -Wextra -pthread -c -o oracle.o oracle.c
g++ -std=c++17 -O3 -fno-tree-vectorize -fPIC -ffp-contract=off -fno-fast-math -Wall -Wextra -pthread -c -o adjoint_ref.o adjoint_ref.cpp
g++ -shared -pthread -o liboracle.so oracle.o adjoint_ref.o -lm
rm -f oracle.o adjoint_ref.o
