#!/bin/bash
# Test-only: the host build of the product core for the test suite (see core_harness.cpp), OpenMP over
# stencils.  (tests/test_core_sanitized.py builds the sanitized variants of the same code itself.)
set -e
cd "$(dirname "$0")"
mkdir -p ../_build
g++ -O2 -ffp-contract=off -fno-fast-math -std=c++17 -fPIC -shared -fopenmp -Wall -Wno-unknown-pragmas -o ../_build/core_harness.so.tmp core_harness.cpp
mv ../_build/core_harness.so.tmp ../_build/core_harness.so
