#!/bin/bash
# Test-only: the host build of the product core for the test suite (see core_harness.cpp), OpenMP over
# stencils.  With CCD_SANITIZE=1 it also builds tests/_build/core_harness_san.so: the same core with
# AddressSanitizer and UndefinedBehaviorSanitizer and a bounds check on every runtime index into the
# core's per-lane arrays (CCD_CHECK_INDEX).
set -e
cd "$(dirname "$0")"
mkdir -p ../_build
g++ -O2 -ffp-contract=off -fno-fast-math -std=c++17 -fPIC -shared -fopenmp -Wall -Wno-unknown-pragmas -o ../_build/core_harness.so.tmp core_harness.cpp
mv ../_build/core_harness.so.tmp ../_build/core_harness.so
if [ "${CCD_SANITIZE:-0}" = 1 ]; then
    g++ -O1 -g -ffp-contract=off -fno-fast-math -std=c++17 -fPIC -shared -Wall -Wno-unknown-pragmas \
        -fsanitize=address,undefined -fno-sanitize-recover=undefined -DCCD_CHECK_INDEX \
        -o ../_build/core_harness_san.so.tmp core_harness.cpp
    mv ../_build/core_harness_san.so.tmp ../_build/core_harness_san.so
fi
