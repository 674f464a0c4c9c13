#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/liboracle.so.
# Strict IEEE FP64: no contraction into FMA, no fast-math.
set -e
cd "$(dirname "$0")"
g++ -O2 -std=c++17 -ffp-contract=off -fno-fast-math \
    -fvisibility=hidden -fPIC -shared -pthread -Wall -Wextra -Wno-unused-parameter \
    -Wl,-Bsymbolic -o liboracle.so oracle.cpp
