#!/bin/sh
# Builds the oracle (test infrastructure) as a shared library next to this
# script.  Contraction is off so float expressions round exactly as written.
set -e
here=$(cd "$(dirname "$0")" && pwd)
gcc -O2 -std=c11 -fPIC -shared -fopenmp -ffp-contract=off -fno-fast-math \
    -Wall -Wno-misleading-indentation \
    -o "$here/libdsim_oracle.so" "$here/dsim_oracle.c" -lm
