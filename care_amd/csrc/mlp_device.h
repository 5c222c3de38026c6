// Learned correction gate: a 32 -> 64 -> 32 -> 16 -> 1 MLP over MSA column
// features (ReLU hidden layers, sigmoid output), evaluated with BF16 MFMA
// (v_mfma_f32_16x16x32_bf16 / 16x16x16) and FP32 accumulation.  One wave
// evaluates a tile of 16 columns.  The gate has no counterpart in the
// reference; it replaces the forest decision of include/forest.hpp:217-226
// in the engine's gated correction mode.
#pragma once

#include "engine.h"

namespace care_engine {

constexpr int kMlpIn = 32, kMlpH1 = 64, kMlpH2 = 32, kMlpH3 = 16;

// Offsets (floats) of the FP32 master parameters.  Weight matrices are
// stored output-major: W[n][k] multiplies input k into output n.
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpH1 * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpH1;
constexpr int kMlpB2 = kMlpW2 + kMlpH2 * kMlpH1;
constexpr int kMlpW3 = kMlpB2 + kMlpH2;
constexpr int kMlpB3 = kMlpW3 + kMlpH3 * kMlpH2;
constexpr int kMlpW4 = kMlpB3 + kMlpH3;
constexpr int kMlpB4 = kMlpW4 + kMlpH3;
constexpr int kMlpParams = kMlpB4 + 1;  // 4737

typedef unsigned short bf16raw;  // bits of a bfloat16
typedef bf16raw ushort8 __attribute__((vector_size(16)));
typedef bf16raw ushort4v __attribute__((vector_size(8)));
typedef float floatx4 __attribute__((vector_size(16)));
typedef __bf16 bf16x8 __attribute__((vector_size(16)));
typedef short short4v __attribute__((vector_size(8)));

// float -> bfloat16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ bf16raw f2bf(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return bf16raw(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(bf16raw b) {
    const unsigned u = unsigned(b) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// D = A x B + C for 16x16 tiles, K = 32.  Lane l holds A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16] and D[4(l/16)..+3][l%16].
__device__ __forceinline__ floatx4 mfma32(ushort8 a, ushort8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}
// K = 16: lane l holds A[l%16][4(l/16)..+3], B[4(l/16)..+3][l%16].
__device__ __forceinline__ floatx4 mfma16(ushort4v a, ushort4v b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4v, a), __builtin_bit_cast(short4v, b),
                                                     c, 0, 0, 0);
}

// BF16 weights and FP32 biases of the gate, as laid out in LDS or HBM.
struct MlpWeights {
    const bf16raw* w1;  // [64][32]
    const bf16raw* w2;  // [32][64]
    const bf16raw* w3;  // [16][32]
    const float* b1;    // [64]
    const float* b2;    // [32]
    const float* b3;    // [16]
    const float* w4;    // [16]
    float b4;
};

// Per-wave LDS scratch of one tile: H1 (bf16 [16][64]), H2 (bf16 [16][32]),
// H3 (float [16][16]).
constexpr int kMlpTileScratch = 16 * kMlpH1 * 2 + 16 * kMlpH2 * 2 + 16 * kMlpH3 * 4;

__device__ __forceinline__ ushort8 load8(const bf16raw* p) { return *reinterpret_cast<const ushort8*>(p); }

// Forward pass of one tile of 16 rows x (bf16 [16][32], row-major) by one
// wave.  Writes the pre-sigmoid outputs z[16] (lanes 0..15 write z[lane]).
// Hidden activations are rounded to bf16 between layers; layer 4 runs in
// FP32.  All 64 lanes must call.
__device__ __forceinline__ void mlp_tile_forward(const MlpWeights& W, const bf16raw* x, char* scratch, float* z) {
    // the lane index is made opaque here so that the compiler computes the
    // lane-dependent addresses below at the call, instead of hoisting them
    // out of the caller's loops where they would hold registers throughout
    int lane = int(__lane_id());
    asm volatile("" : "+v"(lane));
    const int r = lane & 15, g = lane >> 4;
    bf16raw* h1 = reinterpret_cast<bf16raw*>(scratch);
    bf16raw* h2 = h1 + 16 * kMlpH1;
    float* h3 = reinterpret_cast<float*>(h2 + 16 * kMlpH2);
    const ushort8 a1 = load8(x + r * kMlpIn + 8 * g);
    // loops unrolled so that the weight loads of a layer are in flight
    // together (the weights come from L2; one load per MFMA in sequence
    // would serialise their latencies); with the lane index opaque the
    // unrolled loads do not raise the correction kernel's register peak
#pragma unroll
    for (int t = 0; t < kMlpH1 / 16; t++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma32(a1, load8(W.w1 + (16 * t + r) * kMlpIn + 8 * g), acc);
        const float b = W.b1[16 * t + r];
        for (int i = 0; i < 4; i++) {
            const float v = acc[i] + b;
            h1[(4 * g + i) * kMlpH1 + 16 * t + r] = f2bf(v > 0.f ? v : 0.f);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int t = 0; t < kMlpH2 / 16; t++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < kMlpH1 / 32; kk++)
            acc = mfma32(load8(h1 + r * kMlpH1 + 32 * kk + 8 * g), load8(W.w2 + (16 * t + r) * kMlpH1 + 32 * kk + 8 * g),
                         acc);
        const float b = W.b2[16 * t + r];
        for (int i = 0; i < 4; i++) {
            const float v = acc[i] + b;
            h2[(4 * g + i) * kMlpH2 + 16 * t + r] = f2bf(v > 0.f ? v : 0.f);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma32(load8(h2 + r * kMlpH2 + 8 * g), load8(W.w3 + r * kMlpH2 + 8 * g), acc);
        const float b = W.b3[r];
        for (int i = 0; i < 4; i++) {
            const float v = acc[i] + b;
            h3[(4 * g + i) * kMlpH3 + r] = v > 0.f ? v : 0.f;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < 16) {
        float s = W.b4;
#pragma unroll
        for (int j = 0; j < kMlpH3; j++) s = fmaf(h3[lane * kMlpH3 + j], W.w4[j], s);
        z[lane] = s;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace care_engine
