#!/bin/sh
# Builds the FP64 oracle (test infrastructure) into oracle/_build/liboracle64.so.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
gcc -O3 -fPIC -shared -fopenmp -std=c99 -Wall -o "$here/_build/liboracle64.so" "$here/leapfrog64.c"
