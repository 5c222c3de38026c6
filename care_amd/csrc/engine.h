// Internal declarations shared by the HIP translation units of the engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace care_engine {

using u8 = std::uint8_t;
using u16 = std::uint16_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

constexpr u64 kEmptyKey = ~u64(0);
// dynamic LDS a kernel of the engine may request per workgroup
constexpr int kMaxLdsBytes = 64 * 1024;
// default longest hit list a query wavefront merges in LDS
constexpr int kDefaultQueryCap = 512;

// Read store resident in HBM.  Rows are padded to `pitch` words (2-bit
// sequences) and `qpitch` words (binned qualities).
struct DevReadStore {
    const u32* seqs = nullptr;
    const int* lens = nullptr;
    const u8* ambig = nullptr;
    const u32* quals = nullptr;
    u64 n = 0;
    int pitch = 0;
    int qpitch = 0;
    int qbits = 8;
    int maxLen = 0;
};

// One finished minhash map: open-addressing table (linear probing on
// murmur64(key) & mask) holding keys with 2..cap values, CSR values.
struct DevMap {
    const u64* keys;
    const uint2* offcnt;  // (offset into values, count)
    const u32* values;
    u64 mask;
};

struct DevTables {
    const DevMap* maps = nullptr;
    int h = 0;
    int k = 0;
    u64 kmask = 0;
    const u64* sigs = nullptr;  // [n][h] signatures of the read store, kept from the build
};

// Per-candidate record of the correction kernel (16 bytes).
struct CandMeta {
    u32 id;
    short shift;
    u16 len;
    union {
        struct {
            u16 nOps;
            u16 overlap;
        } al;
        float weight;
    } u;
    u8 flag;    // 1 forward, 2 reverse complement
    u8 paired;
    u16 pad;
};
static_assert(sizeof(CandMeta) == 16, "CandMeta layout");

// Random-forest node (include/forest.hpp:22-37): a split on feature `att`
// at `split`; flag bit 1 / bit 0 mark the left / right child as a leaf, in
// which case lhs / rhs hold the leaf probability's float bits, otherwise
// the child's index within the tree.
struct ForestNode {
    double split;
    u32 lhs;
    u32 rhs;
    u8 att;
    u8 flag;
    u8 pad[6];
};
static_assert(sizeof(ForestNode) == 24, "ForestNode layout");

struct DevForest {
    const ForestNode* nodes = nullptr;
    const u32* treeStart = nullptr;
    int numTrees = 0;
    float threshold = 0.5f;
};

// Scalar parameters of the correction kernel, precomputed on the host in
// the reference's float arithmetic.
struct CorrectParams {
    DevReadStore rs;
    u64 firstAnchor;
    int numAnchors;
    const u32* candIds;     // [numAnchors][maxCand]
    const int* candCount;   // [numAnchors]
    int maxCand;
    CandMeta* metaA;        // [numAnchors][maxCand]
    CandMeta* metaB;
    u8* outStatus;
    char* outSeq;
    int outPitch;
    u32* candOutIds;
    char* candOutSeq;
    int* candOutCount;
    int candOutCap;
    int* queueHead;
    const float* qweights;  // 256 entries
    // options
    int pairedEnd;
    int useQuality;
    int correctCandidates;
    int newColumns;
    int minOverlap;
    float maxErrorRate;
    float minOverlapRatio;
    float maxErr4;              // 4 * estimatedErrorrate
    float mmBase2, mmBase3, mmBase4;  // 2,3,4 * estimatedErrorrate
    float goodAlignmentsThreshold;    // estimatedCoverage * m_coverage
    float pairedThreshold;
    int sigCountThreshold;     // int(estimatedCoverage)
    float avgSupportThreshold;  // 1 - err
    float minSupportThreshold;  // fma(-3, err, 1)
    float minCoverageThreshold; // m_coverage / 6 * coverage
    int errPercent;             // ceil(err * 100)
    float errPercentHalf;       // errPercent * 0.5
    // forest / learned gate
    int correctionType;
    int correctionTypeCands;
    DevForest forestA;
    DevForest forestC;
    float estimatedCoverage;
    int maxRefinements;         // different-region refinement rounds (the reference: 5)
    int gatedRefine;            // correctionType 3 recording the plain method's columns: refine first
    // learned gate (correctionType 3): BF16 weight matrices and the FP32
    // parameter vector of the gate, logit threshold of acceptance
    const unsigned short* gateW1;
    const unsigned short* gateW2;
    const unsigned short* gateW3;
    const float* gateParams;
    float gateLogit;
    // optional record of every gate decision column (features, read id,
    // position, original and consensus base); null when not recording
    float* gateFeat;  // [gateCap][32]
    u32* gateIds;
    unsigned short* gatePos;
    u8* gateOrig;
    u8* gateCons;
    int* gateCount;
    int gateCap;
    int gateRecordCands;  // record candidate decision columns instead of the anchor's
    // learned gate on candidate columns (correctionTypeCands 3 in the gated
    // mode): weights, parameters and logit threshold
    const unsigned short* gateCW1;
    const unsigned short* gateCW2;
    const unsigned short* gateCW3;
    const float* gateCParams;
    float gateCLogit;
};

// launchers (all asynchronous on `stream`)
void launch_signatures(const DevReadStore& rs, u64 first, int n, int k, int h, u64* out, hipStream_t stream);
// LDS of one query wavefront: header (hit-list prefix counts and list
// pointers), two merge buffers of capL ids, capL hit counts with candLimit
__host__ __device__ inline size_t query_wave_lds_bytes(int capL, int candLimit) {
    return (size_t(800) + 8 * size_t(capL) + (candLimit > 0 ? size_t(capL) : 0) + 15) / 16 * 16;
}
// Where a query takes its anchors from: by default anchor a is read
// first + a of the read store with its stored signature row; a caller's own
// sequences bring their signature rows (sigs, [n][h]), lengths, and the read
// ids used to drop the anchor itself and, under exclude_ambiguous, ambiguous
// anchors (ids may be null: no read is dropped as the anchor itself).
struct QueryAnchors {
    const u64* sigs = nullptr;
    const u32* ids = nullptr;
    const int* lens = nullptr;
};
// spill: 4 * maxCand words per anchor of the batch (sort/merge buffers of
// anchors with more than capL hits); regMax: longest hit list sorted in
// registers (0 = the kernel's maximum, 1024), longer lists are merged;
// rawCount (may be null): table hits per anchor before de-duplication
void launch_query(const DevReadStore& rs, const DevTables& t, u64 first, int n, int maxCand, int capL, int regMax,
                  int excludeAmbiguous, int candLimit, u32* candIds, int* candCount, u32* spill, int* rawCount,
                  hipStream_t stream, const QueryAnchors& anchors = QueryAnchors());
// read-store gathers by read id into device buffers (store.hip)
void launch_gather_sequences(const DevReadStore& rs, const u32* ids, int n, u32* out, int outPitch, hipStream_t s);
void launch_gather_lengths(const DevReadStore& rs, const u32* ids, int n, int* out, hipStream_t s);
void launch_gather_qualities(const DevReadStore& rs, const u32* ids, int n, char* out, int outPitch, hipStream_t s);
void launch_gather_ambiguous(const DevReadStore& rs, const u32* ids, int n, u8* out, hipStream_t s);
// totals[0] += sum(rawCount[0..n)), totals[1] += sum(candCount[0..n))
void launch_sum_counts(const int* rawCount, const int* candCount, int n, unsigned long long* totals, hipStream_t s);
// chunk: candidates staged per MSA pass (0 = automatic); gated: bit 0 the
// anchor gate (correction type 3), bit 1 the candidate gate
int correct_gate_mode(int correctionType, int correctCandidates, int correctionTypeCands);
void launch_correct(const CorrectParams& p, int blocks, int threads, int chunk, hipStream_t stream);
size_t correct_lds_bytes(int maxLen, int useQuality, int qpitch, int threads, int gated, int chunk);
int correct_msa_chunk(int maxLen, int useQuality, int qpitch, int threads, int gated, int chunk);
// correction workgroups resident per compute unit at once (occupancy)
int correct_resident_blocks(int threads, size_t ldsBytes, int gated);
int query_lds_bytes(int h, int cap, int maxLen, int threads, int candLimit, int capL);

// learned gate (mlp.hip)
void launch_mlp_grad(const unsigned short* X, const u8* Y, const u32* perm, u64 first, int rows,
                     const unsigned short* w1b, const unsigned short* w2b, const unsigned short* w3b,
                     const float* master, float invBatch, float* partial, hipStream_t s);
int mlp_partials(int rows);
int mlp_partial_stride();
void launch_mlp_reduce(const float* partial, int nPartials, int mode, float* out, float* master, float* m1, float* m2,
                       float lr, float beta1, float beta2, float eps, float bc1, float bc2, unsigned short* w1b,
                       unsigned short* w2b, unsigned short* w3b, float* lossAcc, hipStream_t s);
void launch_mlp_cast(const float* master, unsigned short* w1b, unsigned short* w2b, unsigned short* w3b,
                     hipStream_t s);
void launch_mlp_forward(const unsigned short* X, u64 n, const unsigned short* w1b, const unsigned short* w2b,
                        const unsigned short* w3b, const float* master, float* zOut, hipStream_t s);

}  // namespace care_engine
