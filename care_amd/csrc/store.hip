// Gathers from the read store in HBM into caller device buffers, by read id,
// on the caller's stream (the device side of GpuReadStorage::gather*,
// include/gpu/gpureadstorage.cuh:22-126, include/gpu/multigpureadstorage.cuh:
// 669,921,1040,1341,1529, and areSequencesAmbiguous :637).  Rows are copied
// word by word with one thread per word, so a wavefront reads whole rows
// (pitch 7-16 words) in a few 64-byte lines.

#include "device_common.h"

namespace care_engine {

__global__ void k_gather_rows(const u32* __restrict__ src, int srcPitch, u64 numReads, const u32* __restrict__ ids,
                              int n, u32* __restrict__ out, int outPitch) {
    const u64 total = u64(n) * u64(outPitch);
    for (u64 i = blockIdx.x * u64(blockDim.x) + threadIdx.x; i < total; i += u64(gridDim.x) * blockDim.x) {
        const u64 r = i / u64(outPitch);
        const int w = int(i - r * u64(outPitch));
        const u32 id = ids[r];
        // words beyond the stored pitch, and rows of ids outside the store,
        // are written as zero
        out[i] = (w < srcPitch && id < numReads) ? src[u64(id) * u64(srcPitch) + u64(w)] : 0u;
    }
}

__global__ void k_gather_lengths(const int* __restrict__ lens, u64 numReads, const u32* __restrict__ ids, int n,
                                 int* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const u32 id = ids[i];
        out[i] = id < numReads ? lens[id] : 0;
    }
}

__global__ void k_gather_ambiguous(const u8* __restrict__ ambig, u64 numReads, const u32* __restrict__ ids, int n,
                                   u8* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const u32 id = ids[i];
        out[i] = id < numReads ? ambig[id] : u8(0);
    }
}

// decoded quality characters (the bins' representative characters for 1-
// and 2-bit storage), outPitch characters per row; positions past the
// store's maximum length are left as they are
__global__ void k_gather_qualities(DevReadStore rs, const u32* __restrict__ ids, int n, char* __restrict__ out,
                                   int outPitch) {
    const int width = min(outPitch, rs.maxLen);
    const u64 total = u64(n) * u64(width);
    for (u64 i = blockIdx.x * u64(blockDim.x) + threadIdx.x; i < total; i += u64(gridDim.x) * blockDim.x) {
        const u64 r = i / u64(width);
        const int j = int(i - r * u64(width));
        const u32 id = ids[r];
        out[r * u64(outPitch) + u64(j)] = id < rs.n ? quality_char(rs.quals + u64(id) * rs.qpitch, j, rs.qbits) : 0;
    }
}

static int grid_of(u64 work) {
    const u64 b = (work + 255) / 256;
    return int(b == 0 ? 1 : (b < 8192 ? b : 8192));
}

void launch_gather_sequences(const DevReadStore& rs, const u32* ids, int n, u32* out, int outPitch, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_of(u64(n) * outPitch)), dim3(256), 0, s, rs.seqs, rs.pitch, rs.n, ids,
                       n, out, outPitch);
}
void launch_gather_lengths(const DevReadStore& rs, const u32* ids, int n, int* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_lengths, dim3(grid_of(u64(n))), dim3(256), 0, s, rs.lens, rs.n, ids, n, out);
}
void launch_gather_ambiguous(const DevReadStore& rs, const u32* ids, int n, u8* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_ambiguous, dim3(grid_of(u64(n))), dim3(256), 0, s, rs.ambig, rs.n, ids, n, out);
}
void launch_gather_qualities(const DevReadStore& rs, const u32* ids, int n, char* out, int outPitch, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_qualities, dim3(grid_of(u64(n) * u64(outPitch))), dim3(256), 0, s, rs, ids, n, out,
                       outPitch);
}

}  // namespace care_engine
