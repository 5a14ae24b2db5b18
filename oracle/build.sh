#!/usr/bin/env bash
# Builds the FP64 CPU oracle (test infrastructure) into oracle/_build/.
# The reference itself cannot be built here (it needs Eigen3, Boost,
# gfortran and MPI headers), so there is no oracle/_ref build.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
g++ -std=c++17 -O3 -march=x86-64-v3 -fopenmp -fPIC -shared -Wall -Wno-sign-compare \
    "$here/oracle.cpp" -o "$here/_build/liboracle.so.tmp"
mv "$here/_build/liboracle.so.tmp" "$here/_build/liboracle.so"
