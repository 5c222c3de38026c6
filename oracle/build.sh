#!/bin/sh
# Builds the CPU restatement of CARE's correction path (test infrastructure):
#   oracle/_build/liboracle.so  (C interface for the tests)
#   oracle/_build/care-oracle   (care-cpu compatible command line)
set -e
cd "$(dirname "$0")"
mkdir -p _build
FLAGS="-O2 -std=c++17 -fopenmp -ffp-contract=off -fPIC -Wall -Wno-unused-function"
g++ $FLAGS -shared -o _build/liboracle.so care_oracle.cpp -lz
g++ $FLAGS -o _build/care-oracle care_oracle_main.cpp care_oracle.cpp -lz
