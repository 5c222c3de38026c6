// Device helpers shared by the engine's kernels (gfx950, wave64).
#pragma once

#include "engine.h"

// Index checks of the kernels: compiled only into the CPU emulation of the
// HIP API built with CARE_EMU_CHECKS (test infrastructure, where a failed
// check aborts with its location); nothing in the device build.
#ifdef CARE_EMU_CHECKS
#define CARE_CHECK(cond) emu::check((cond), #cond, __FILE__, __LINE__)
#else
#define CARE_CHECK(cond) ((void)0)
#endif

namespace care_engine {

constexpr int kWave = 64;

__device__ __forceinline__ u64 murmur64(u64 x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// base i of a 2-bit packed row (first base in bits 31..30)
__device__ __forceinline__ u32 base_at(const u32* enc, int i) {
    return (enc[i >> 4] >> (30 - 2 * (i & 15))) & 3u;
}

// base i of a read as it appears in the MSA: forward, or reverse complement
__device__ __forceinline__ u32 oriented_base(const u32* enc, int len, int i, bool rc) {
    return rc ? 3u - base_at(enc, len - 1 - i) : base_at(enc, i);
}

// character j of a binned quality row (8, 2 or 1 bits per base)
__device__ __forceinline__ char quality_char(const u32* enc, int j, int bits) {
    if (bits == 8) {
        const u32 w = enc[j >> 2];
        return char((w >> (8 * (j & 3))) & 0xFFu);  // little-endian byte order
    }
    if (bits == 2) {
        // the four bin characters '$', '(', '1', '9' packed in one constant:
        // a string-table lookup would be a memory load per call
        const u32 v = (enc[j >> 4] >> (30 - 2 * (j & 15))) & 3u;
        return char((0x39312824u >> (8u * v)) & 0xFFu);
    }
    const u32 v = (enc[j >> 5] >> (31 - (j & 31))) & 1u;
    return v ? '5' : '(';
}

// character of a 2-bit base code ('A', 'C', 'G', 'T'), without a table load
__device__ __forceinline__ char base_char(u32 b) { return char((0x54474341u >> (8u * b)) & 0xFFu); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Minimum over the 64 lanes, returned in every lane; all lanes must be
// active.  Within each row of 16 lanes by DPP moves (quad_perm [1,0,3,2] and
// [2,3,0,1], i.e. lane ^ 1 and lane ^ 2; row_half_mirror, lane i of an
// 8-lane half-row taking lane 7 - i; row_mirror, lane i of a row taking lane
// 15 - i), then across the four rows by readlane: no LDS round trips, unlike
// shuffles, which compile to ds_bpermute.
__device__ __forceinline__ u32 wave_min_u32(u32 v) {
    v = min(v, u32(__builtin_amdgcn_mov_dpp(int(v), 0xB1, 0xF, 0xF, true)));
    v = min(v, u32(__builtin_amdgcn_mov_dpp(int(v), 0x4E, 0xF, 0xF, true)));
    v = min(v, u32(__builtin_amdgcn_mov_dpp(int(v), 0x141, 0xF, 0xF, true)));
    v = min(v, u32(__builtin_amdgcn_mov_dpp(int(v), 0x140, 0xF, 0xF, true)));
    const u32 r0 = u32(__builtin_amdgcn_readlane(int(v), 0)), r1 = u32(__builtin_amdgcn_readlane(int(v), 16));
    const u32 r2 = u32(__builtin_amdgcn_readlane(int(v), 32)), r3 = u32(__builtin_amdgcn_readlane(int(v), 48));
    return min(min(r0, r1), min(r2, r3));
}

// Reduction over the 64 lanes with the associative, commutative `op`,
// returned in every lane; all lanes must be active.  Same DPP / readlane
// pattern as wave_min_u32.
template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
    v = op(v, T(__builtin_amdgcn_mov_dpp(int(v), 0xB1, 0xF, 0xF, true)));
    v = op(v, T(__builtin_amdgcn_mov_dpp(int(v), 0x4E, 0xF, 0xF, true)));
    v = op(v, T(__builtin_amdgcn_mov_dpp(int(v), 0x141, 0xF, 0xF, true)));
    v = op(v, T(__builtin_amdgcn_mov_dpp(int(v), 0x140, 0xF, 0xF, true)));
    const T r0 = T(__builtin_amdgcn_readlane(int(v), 0)), r1 = T(__builtin_amdgcn_readlane(int(v), 16));
    const T r2 = T(__builtin_amdgcn_readlane(int(v), 32)), r3 = T(__builtin_amdgcn_readlane(int(v), 48));
    return op(op(r0, r1), op(r2, r3));
}
__device__ __forceinline__ int wave_sum_int(int v) { return wave_reduce(v, [](int x, int y) { return x + y; }); }
__device__ __forceinline__ int wave_min_int(int v) { return wave_reduce(v, [](int x, int y) { return x < y ? x : y; }); }
__device__ __forceinline__ int wave_max_int(int v) { return wave_reduce(v, [](int x, int y) { return x > y ? x : y; }); }

// Exclusive prefix count of `flag` over the whole block, in thread order.
// `scratch` must hold blockDim.x/64 + 1 ints.  Returns the prefix for this
// thread; *total receives the block-wide count.  All threads must call.
__device__ __forceinline__ int block_exclusive_count(bool flag, int* scratch, int* total) {
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const int nwaves = (blockDim.x + kWave - 1) / kWave;
    const unsigned long long bal = __ballot(flag);
    const unsigned long long below = lane == 0 ? 0ull : (bal & (~0ull >> (64 - lane)));
    const int inWave = __popcll(below);
    __syncthreads();
    if (lane == 0) scratch[wave] = __popcll(bal);
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < nwaves; w++) {
        const int c = scratch[w];
        if (w < wave) before += c;
        all += c;
    }
    *total = all;
    return before + inWave;
}

__device__ __forceinline__ unsigned next_pow2(unsigned v) {
    if (v <= 1) return 1;
    return 1u << (32 - __clz(v - 1));
}

}  // namespace care_engine
