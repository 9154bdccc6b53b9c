#!/bin/sh
# Builds the oracle's shared library next to its source (test infrastructure only).
set -e
cd "$(dirname "$0")"
gcc -std=c11 -O2 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wextra -pthread \
    -o liboracle.so oracle.c -lm
