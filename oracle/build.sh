#!/bin/sh
# Builds the oracle's shared library next to its source (test infrastructure only).
set -e
cd "$(dirname "$0")"
gcc -std=c11 -O2 -fPIC -ffp-contract=off -fno-fast-math -Wall -Wextra -pthread -c -o oracle.o oracle.c
g++ -std=c++17 -O2 -fPIC -ffp-contract=off -fno-fast-math -Wall -Wextra -pthread -c -o adjoint_ref.o adjoint_ref.cpp
g++ -shared -pthread -o liboracle.so oracle.o adjoint_ref.o -lm
rm -f oracle.o adjoint_ref.o
