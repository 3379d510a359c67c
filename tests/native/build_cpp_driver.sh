#!/bin/bash
# Builds the C++ interface driver against include/ccdamd and libccd_amd.so (no HIP headers needed).
set -e
cd "$(dirname "$0")"
mkdir -p ../_build
ROOT=$(cd ../.. && pwd)
g++ -O2 -std=c++17 -Wall -I"$ROOT/include/ccdamd" -o ../_build/cpp_interface_driver cpp_interface_driver.cpp \
    -L"$ROOT/collisiondetection_amd" -lccd_amd -Wl,-rpath,"$ROOT/collisiondetection_amd"
g++ -O2 -std=c++17 -Wall -I"$ROOT/include/ccdamd" -o ../_build/velocity_filter_driver velocity_filter_driver.cpp \
    -L"$ROOT/collisiondetection_amd" -lccd_amd -Wl,-rpath,"$ROOT/collisiondetection_amd"
g++ -O2 -std=c++17 -Wall -I"$ROOT/include/ccdamd" -o ../_build/objio_check objio_check.cpp
g++ -O2 -std=c++17 -Wall -ffp-contract=off -I"$ROOT/include/ccdamd" -o ../_build/history_check history_check.cpp
