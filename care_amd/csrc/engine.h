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
};

// launchers (all asynchronous on `stream`)
void launch_signatures(const DevReadStore& rs, u64 first, int n, int k, int h, u64* out, hipStream_t stream);
void launch_build_map(const DevReadStore& rs, const u64* sigs, int h, int map, int cap, u64* tkeys, u32* tcount,
                      u64 tmask, u32* tfill, u64* scanTmp, u32* scanOut, hipStream_t stream);
void launch_query(const DevReadStore& rs, const DevTables& t, u64 first, int n, int maxCand, int excludeAmbiguous,
                  int candLimit, u32* candIds, int* candCount, unsigned long long* rawHits, hipStream_t stream);
void launch_correct(const CorrectParams& p, int blocks, int threads, hipStream_t stream);
size_t correct_lds_bytes(int maxLen, int useQuality, int qpitch, int threads, int gated);
int query_lds_bytes(int h, int cap, int maxLen, int threads, int candLimit);

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
