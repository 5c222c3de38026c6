#!/bin/sh
# Builds the engine's kernel sources against the CPU emulation of the HIP
# API (tests/emu/hip/hip_runtime.h) for the CPU tests:
#   tests/emu/_build/libcare_amd_emu.so  same C-ABI as libcare_amd.so
#   tests/emu/_build/care-gpu-emu        care-gpu linked against it
#
#   build_emu.sh checked   the same into tests/emu/_build/checked/, with the
#                          kernels' index checks (CARE_EMU_CHECKS) and the
#                          address and undefined-behaviour sanitizers (CPU
#                          code only: nothing here is built for a GPU)
set -e
here="$(cd "$(dirname "$0")" && pwd)"
root="$(dirname "$(dirname "$here")")"
out="$here/_build"
FLAGS="-std=c++20 -O2 -g -U_FORTIFY_SOURCE -fPIC -pthread -ffp-contract=off -w -I$here -Wl,-Bsymbolic"
HOSTFLAGS="-std=c++17 -O2"
if [ "$1" = "checked" ]; then
    out="$out/checked"
    SAN="-fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer"
    FLAGS="$FLAGS -O1 -DCARE_EMU_CHECKS $SAN"
    HOSTFLAGS="$HOSTFLAGS -g $SAN"
fi
mkdir -p "$out"
# up to date: both outputs newer than every source they are built from
lib="$out/libcare_amd_emu.so"
exe="$out/care-gpu-emu"
if [ -f "$lib" ] && [ -f "$exe" ] && [ -z "$(find "$root/care_amd/csrc" "$root/care_amd/host" "$root/include" \
        "$here/hip" "$here/build_emu.sh" -type f -newer "$lib" 2>/dev/null)" ] && \
        [ -z "$(find "$root/care_amd/csrc" "$root/care_amd/host" "$root/include" "$here/hip" "$here/build_emu.sh" \
        -type f -newer "$exe" 2>/dev/null)" ]; then
    exit 0
fi
g++ $FLAGS -shared -x c++ -o "$out/libcare_amd_emu.so" \
    "$root/care_amd/csrc/api.hip" "$root/care_amd/csrc/minhash.hip" "$root/care_amd/csrc/correct.hip" \
    "$root/care_amd/csrc/mlp.hip" "$root/care_amd/csrc/store.hip"
g++ $HOSTFLAGS -o "$out/care-gpu-emu" "$root/care_amd/host/care_host.cpp" "$root/care_amd/host/storage_files.cpp" \
    "$root/care_amd/host/care_gpu_main.cpp" \
    -L"$out" -lcare_amd_emu -Wl,-rpath,"$out" -lz -lpthread
