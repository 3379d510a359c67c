#!/bin/bash
# Test-only: host builds of the product core for the CPU test suite (see core_harness.cpp): the core as
# shipped, and a variant whose reference-root solver uses the C library's log/exp (CCD_HOST_LIBM).
set -e
cd "$(dirname "$0")"
mkdir -p ../_build
g++ -O2 -ffp-contract=off -fno-fast-math -std=c++17 -fPIC -shared -Wall -Wno-unknown-pragmas -o ../_build/core_harness.so.tmp core_harness.cpp
mv ../_build/core_harness.so.tmp ../_build/core_harness.so
g++ -O2 -ffp-contract=off -fno-fast-math -std=c++17 -fPIC -shared -Wall -Wno-unknown-pragmas -DCCD_HOST_LIBM -o ../_build/core_harness_libm.so.tmp core_harness.cpp
mv ../_build/core_harness_libm.so.tmp ../_build/core_harness_libm.so
