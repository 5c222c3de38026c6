#!/bin/sh
# Builds the engine's kernel sources against the CPU emulation of the HIP
# API (tests/emu/hip/hip_runtime.h) for the CPU tests:
#   tests/emu/_build/libcare_amd_emu.so  same C-ABI as libcare_amd.so
#   tests/emu/_build/care-gpu-emu        care-gpu linked against it
set -e
here="$(cd "$(dirname "$0")" && pwd)"
root="$(dirname "$(dirname "$here")")"
out="$here/_build"
mkdir -p "$out"
FLAGS="-std=c++20 -O2 -g -U_FORTIFY_SOURCE -fPIC -pthread -ffp-contract=off -w -I$here -Wl,-Bsymbolic"
g++ $FLAGS -shared -x c++ -o "$out/libcare_amd_emu.so" \
    "$root/care_amd/csrc/api.hip" "$root/care_amd/csrc/minhash.hip" "$root/care_amd/csrc/correct.hip" \
    "$root/care_amd/csrc/mlp.hip"
g++ -std=c++17 -O2 -o "$out/care-gpu-emu" "$root/care_amd/host/care_host.cpp" "$root/care_amd/host/care_gpu_main.cpp" \
    -L"$out" -lcare_amd_emu -Wl,-rpath,"$out" -lz -lpthread
