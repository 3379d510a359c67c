#!/bin/bash
# Test-only: host build of the product core for the CPU test suite (see core_harness.cpp).
set -e
cd "$(dirname "$0")"
mkdir -p ../_build
g++ -O2 -ffp-contract=off -fno-fast-math -std=c++17 -fPIC -shared -Wall -Wno-unknown-pragmas -o ../_build/core_harness.so.tmp core_harness.cpp
mv ../_build/core_harness.so.tmp ../_build/core_harness.so
