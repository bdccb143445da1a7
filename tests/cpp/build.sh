#!/bin/sh
# Builds the drop-in C++ test against include/clbool.hpp and the library.
set -e
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
g++ -O2 -std=c++17 -Wall -I"$ROOT/include" -I"$ROOT/oracle" "$ROOT/tests/cpp/test_dropin.cpp" \
    -o "$ROOT/tests/cpp/test_dropin" -L"$ROOT/clbool_amd" -lclbool_amd -L"$ROOT/oracle" -loracle \
    -Wl,-rpath,"$ROOT/clbool_amd" -Wl,-rpath,"$ROOT/oracle" -Wl,-rpath,'$ORIGIN/../../clbool_amd' \
    -Wl,-rpath,'$ORIGIN/../../oracle'
