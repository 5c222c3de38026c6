// C-ABI of the engine: contexts, read store upload, table construction,
// queries and batch correction.  Declared in include/care_amd.h.

#include "../../include/care_amd.h"
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace care_engine {
// launchers of minhash.hip
void launch_fill_u64(u64* p, u64 n, u64 v, hipStream_t s);
void launch_fill_u32(u32* p, u64 n, u32 v, hipStream_t s);
void launch_count_keys(const u64* sigs, u64 n, int h, int map, u64* tkeys, u32* tcount, u64 tmask, int* overflow,
                       hipStream_t s);
// value offsets of the keys with 2..cap values in slot order; tileSums holds
// assign_offsets_scratch(tsize) words
u64 assign_offsets_scratch(u64 tsize);
void launch_assign_offsets(const u64* tkeys, const u32* tcount, u64 tsize, int cap, u32* toff, u64* tileSums,
                           unsigned long long* counters, hipStream_t s);
void launch_insert_final(const u64* tkeys, const u32* tcount, const u32* toff, u64 tsize, u64* fkeys, uint2* foffcnt,
                         u64 fmask, int* overflow, hipStream_t s);
void launch_scatter_values(const u64* sigs, u64 n, int h, int map, const u64* tkeys, const u32* toff, u64 tmask,
                           u32* tfill, u32* values, hipStream_t s);
void launch_sort_buckets(const u32* toff, const u32* tcount, u64 tsize, u32* values, hipStream_t s);
void launch_insert_triples(const u64* keys, const u32* offs, const u32* counts, u64 numKeys, u64* fkeys,
                           uint2* foffcnt, u64 fmask, int* overflow, hipStream_t s);
void launch_reset_counters(int* queue, int* candOut, hipStream_t s);
}  // namespace care_engine

using namespace care_engine;

namespace {

thread_local std::string g_lastError;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_lastError = buf;
    return code;
}

struct HipError {
    hipError_t err;
    const char* what;
};

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) throw HipError{e_, #expr}; \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        HIP_TRY(hipMalloc(&p, sizeof(T) * count));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

template <class T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count, hipHostMallocDefault));
        n = count;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    ~PinnedBuf() { release(); }
};

struct MapStore {
    DevBuf<u64> keys;
    DevBuf<uint2> offcnt;
    DevBuf<u32> values;
    u64 mask = 0;
    u64 numKeys = 0;
    u64 numValues = 0;
};

int resultsPerMapThreshold(float coverage) {
    // include/minhasherlimit.hpp:10-15
    int r = int(coverage * 2.5f);
    r = std::min(r, 65535);
    return std::max(10, r);
}

}  // namespace

struct care_context {
    int device = 0;
    care_options opt{};
    hipStream_t stream = nullptr;
    hipEvent_t ev[4]{};
    // device-to-host copies of each device batch's results, overlapping the
    // next batches' kernels; copyDone joins it back into `stream`
    hipStream_t copyStream = nullptr;
    hipEvent_t copyDone = nullptr;
    int numCUs = 0;
    size_t freeAtCreate = 0;  // device memory free when the context was created

    // read store
    DevBuf<u32> seqs;
    DevBuf<int> lens;
    DevBuf<u8> ambig;
    DevBuf<u32> quals;
    DevReadStore rs{};
    bool hasReads = false;

    // tables
    DevBuf<u64> sigs;  // [n][h] signatures of all reads
    std::vector<std::unique_ptr<MapStore>> maps;
    DevBuf<DevMap> devMaps;
    DevTables tables{};
    bool hasTables = false;
    int cap = 0;

    // batch buffers
    int batchCap = 0;
    int maxCand = 0;
    int queryCap = 0;  // hit lists sorted in LDS up to this length, longer ones in global scratch
    int queryRegMax = 0;  // hit lists sorted in registers up to this length (0 = kernel maximum)
    int blockThreads = 256;  // correction workgroup
    int blocksPerCU = 0;     // persistent correction workgroups per CU (0 = resident maximum)
    int msaChunk = 0;        // candidates per MSA staging pass (0 = automatic)
    int requestedBatch = 0;
    DevBuf<u32> candIds;
    DevBuf<int> candCount;
    DevBuf<int> rawCount;  // table hits per anchor of a batch (statistics)
    DevBuf<CandMeta> metaA, metaB;
    DevBuf<u8> outStatus;
    DevBuf<char> outSeq;
    int outPitch = 0;
    int outCap = 0;                     // anchors the output buffers hold
    std::vector<hipEvent_t> batchEv;    // 3 per device batch of a call
    DevBuf<u32> candOutIds;
    DevBuf<char> candOutSeq;
    int candOutCap = 0;
    DevBuf<int> counters;  // [0] queue head, [1] candidate output count
    DevBuf<unsigned long long> rawHits;
    DevBuf<float> qweights;
    PinnedBuf<int> hInts;
    PinnedBuf<unsigned long long> hU64;

    care_batch_stats stats{};

    // care_correct_batch_graph: the device work of the last call, captured
    // into a graph and replayed while calls repeat it (same range, outputs,
    // device buffers, options and launch shapes)
    struct GraphCache {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<unsigned char> key;
        int launches = 0;
        void release() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr;
            graph = nullptr;
            key.clear();
        }
    } graphCache;

    struct ForestStore {
        DevBuf<ForestNode> nodes;
        DevBuf<u32> treeStart;
        DevForest dev;
        bool loaded = false;
    } forest[2];

    // learned gate: FP32 parameters and BF16 weight copies used by the
    // correction kernel, plus training state
    struct GateStore {
        DevBuf<float> params;
        DevBuf<unsigned short> w1, w2, w3;
        bool loaded = false;
    } gate, gateCands;  // anchor columns; candidate columns (defaults to the anchor gate)
    struct GateRecord {
        DevBuf<float> feat;
        DevBuf<u32> ids;
        DevBuf<unsigned short> pos;
        DevBuf<u8> orig, cons;
        int cap = 0;
        bool on = false;
        bool cands = false;  // record candidate decision columns instead of the anchors'
        bool refine = false; // the plain method's columns: all candidates, refinement first
    } record;
};

namespace {

// Batch buffers sized for `batch` anchors.
void ensure_batch(care_context* ctx, int batch) {
    if (batch <= ctx->batchCap) return;
    const size_t mc = size_t(ctx->maxCand);
    ctx->candIds.alloc(size_t(batch) * mc);
    ctx->candCount.alloc(batch);
    ctx->rawCount.alloc(batch);
    ctx->metaA.alloc(size_t(batch) * mc);
    ctx->metaB.alloc(size_t(batch) * mc);
    ctx->outStatus.alloc(batch);
    ctx->outPitch = (ctx->rs.maxLen + 15) / 16 * 16;
    ctx->outSeq.alloc(size_t(batch) * ctx->outPitch);
    ctx->outCap = batch;
    if (ctx->opt.correct_candidates) {
        // every HQ anchor corrects at most its candidates within the window;
        // a generous average bound, checked per batch
        ctx->candOutCap = int(std::min<size_t>(size_t(batch) * 64, size_t(1) << 26));
        ctx->candOutIds.alloc(ctx->candOutCap);
        ctx->candOutSeq.alloc(size_t(ctx->candOutCap) * ctx->outPitch);
    }
    ctx->batchCap = batch;
}

// Anchors per device batch: the batch scratch (candidate ids and two
// records per candidate slot, h x cap slots per anchor) gets a quarter of the
// device memory free when the context was created, between 2 and 24 GiB, and
// a batch holds at most 2^18 anchors.  Fewer, larger batches mean fewer
// launches and fewer drains of the persistent correction kernel.
int default_batch(care_context* ctx) {
    if (ctx->requestedBatch > 0) return ctx->requestedBatch;
    const size_t perAnchor = size_t(ctx->maxCand) * (sizeof(u32) + 2 * sizeof(CandMeta)) + ctx->rs.maxLen + 64;
    const size_t budget = std::min(std::max(ctx->freeAtCreate / 4, size_t(2) << 30), size_t(24) << 30);
    size_t b = budget / perAnchor;
    b = std::max<size_t>(256, std::min<size_t>(b, size_t(1) << 18));
    return int(b & ~size_t(1));
}

}  // namespace

namespace care_engine {
__global__ void k_reset_counters(int* queue, int* candOut) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        queue[0] = 0;
        if (candOut) candOut[0] = 0;
    }
}
void launch_reset_counters(int* queue, int* candOut, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, queue, candOut);
}
}  // namespace care_engine

extern "C" {

void care_options_default(care_options* o) {
    if (!o) return;
    o->kmer_length = 20;
    o->num_hash_maps = 48;
    o->coverage = 1.0f;
    o->paired_end = 0;
    o->use_quality = 0;
    o->exclude_ambiguous = 0;
    o->correct_candidates = 0;
    o->new_columns_to_correct = 15;
    o->min_overlap = 30;
    o->max_mismatch_ratio = 0.2f;
    o->min_overlap_ratio = 0.30f;
    o->error_factor = 0.06f;
    o->coverage_factor = 0.6f;
    o->paired_filter_threshold = 0.06f;
    o->correction_type = 0;
    o->correction_type_cands = 0;
    o->threshold_anchor = 0.5f;
    o->threshold_cands = 0.5f;
    o->candidate_limit = 0;
    o->max_refinements = 5;
}

const char* care_last_error(void) { return g_lastError.c_str(); }

int care_context_create(int device, const care_options* opt, care_context** out) {
    if (!out || !opt) return fail(CARE_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (opt->kmer_length < 1 || opt->kmer_length > 31)
        return fail(CARE_ERR_ARGUMENT, "kmer_length %d outside 1..31", opt->kmer_length);
    if (opt->num_hash_maps < 1 || opt->num_hash_maps > 64)
        return fail(CARE_ERR_ARGUMENT, "num_hash_maps %d outside 1..64", opt->num_hash_maps);
    if ((opt->correction_type != 0 && opt->correction_type != 1 && opt->correction_type != 3) ||
        (opt->correction_type_cands != 0 && opt->correction_type_cands != 1 && opt->correction_type_cands != 3))
        return fail(CARE_ERR_ARGUMENT, "correction types: 0 classic, 1 forest, 3 learned gate");
    if (opt->correction_type_cands == 3 && opt->correction_type != 3)
        return fail(CARE_ERR_ARGUMENT, "correction_type_cands 3 (learned gate) needs correction_type 3");
    if (opt->candidate_limit < 0) return fail(CARE_ERR_ARGUMENT, "candidate_limit must be >= 0");
    if (opt->max_refinements < 0 || opt->max_refinements > 5)
        return fail(CARE_ERR_ARGUMENT, "max_refinements must be in 0..5");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CARE_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(CARE_ERR_ARGUMENT, "device %d outside 0..%d", device, count - 1);
    auto* ctx = new (std::nothrow) care_context();
    if (!ctx) return fail(CARE_ERR_DEVICE, "out of host memory");
    try {
        ctx->device = device;
        ctx->opt = *opt;
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&ctx->copyStream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ctx->copyDone, hipEventDisableTiming));
        for (auto& e : ctx->ev) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipDeviceGetAttribute(&ctx->numCUs, hipDeviceAttributeMultiprocessorCount, device));
        size_t freeB = 0, totalB = 0;
        HIP_TRY(hipMemGetInfo(&freeB, &totalB));
        ctx->freeAtCreate = freeB;
        ctx->cap = resultsPerMapThreshold(opt->coverage);
        ctx->maxCand = opt->num_hash_maps * ctx->cap;
        ctx->queryCap = std::min((ctx->maxCand + 15) / 16 * 16, kDefaultQueryCap);
        // quality weights: include/qualityscoreweights.hpp:15-30 (host powf)
        float w[256];
        for (int i = 0; i < 256; i++) {
            float p = 1.0f;
            if (i >= 33) p = std::pow(10.0f, float(-(i - 33)) / 10.0f);
            w[i] = std::max(0.001f, 1.0f - p);
        }
        ctx->qweights.alloc(256);
        HIP_TRY(hipMemcpy(ctx->qweights.p, w, sizeof(w), hipMemcpyHostToDevice));
        ctx->counters.alloc(4);
        ctx->rawHits.alloc(2);
        ctx->hInts.alloc(8);
        ctx->hU64.alloc(4);
    } catch (const HipError& e) {
        delete ctx;
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    } catch (const std::bad_alloc&) {
        delete ctx;
        return fail(CARE_ERR_DEVICE, "out of host memory");
    }
    *out = ctx;
    return CARE_OK;
}

int care_context_destroy(care_context* ctx) {
    if (!ctx) return CARE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copyStream) (void)hipStreamSynchronize(ctx->copyStream);
    ctx->graphCache.release();
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->batchEv) (void)hipEventDestroy(e);
    if (ctx->copyDone) (void)hipEventDestroy(ctx->copyDone);
    if (ctx->copyStream) (void)hipStreamDestroy(ctx->copyStream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CARE_OK;
}

int care_set_launch_config(care_context* ctx, int batch_anchors, int block_threads) {
    if (!ctx) return fail(CARE_ERR_ARGUMENT, "null context");
    if (batch_anchors < 0 ||
        (block_threads != 0 && block_threads != 128 && block_threads != 192 && block_threads != 256))
        return fail(CARE_ERR_ARGUMENT, "invalid launch configuration (block_threads 128, 192 or 256)");
    if (batch_anchors > 0 && (batch_anchors & 1)) return fail(CARE_ERR_ARGUMENT, "batch_anchors must be even");
    ctx->requestedBatch = batch_anchors;
    if (block_threads) ctx->blockThreads = block_threads;
    return CARE_OK;
}

int care_set_tuning(care_context* ctx, const care_tuning* t) {
    if (!ctx || !t) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (t->batch_anchors < 0 || (t->batch_anchors & 1)) return fail(CARE_ERR_ARGUMENT, "batch_anchors must be even and >= 0");
    if (t->correct_threads != 0 && t->correct_threads != 128 && t->correct_threads != 192 && t->correct_threads != 256)
        return fail(CARE_ERR_ARGUMENT, "correct_threads must be 128, 192 or 256");
    if (t->correct_blocks_per_cu < 0 || t->correct_blocks_per_cu > 64)
        return fail(CARE_ERR_ARGUMENT, "correct_blocks_per_cu outside 0..64");
    if (t->msa_chunk != 0 && (t->msa_chunk < 16 || t->msa_chunk > 256))
        return fail(CARE_ERR_ARGUMENT, "msa_chunk outside 16..256");
    const int q = t->query_lds_hits ? t->query_lds_hits : kDefaultQueryCap;
    if (q < 16 || q > 4096 || (q & 15)) return fail(CARE_ERR_ARGUMENT, "query_lds_hits must be a multiple of 16 in 16..4096");
    if (4 * query_wave_lds_bytes(q, ctx->opt.candidate_limit) > size_t(kMaxLdsBytes))
        return fail(CARE_ERR_CAPACITY, "query_lds_hits %d exceeds the query workgroup's LDS", q);
    const int rm = t->query_register_hits;
    if (rm != 0 && (rm < 64 || rm > 2048 || (rm & (rm - 1))))
        return fail(CARE_ERR_ARGUMENT, "query_register_hits must be 0 or a power of two in 64..2048");
    ctx->requestedBatch = t->batch_anchors;
    ctx->blockThreads = t->correct_threads ? t->correct_threads : 256;
    ctx->blocksPerCU = t->correct_blocks_per_cu;
    ctx->msaChunk = t->msa_chunk;
    ctx->queryCap = q;
    ctx->queryRegMax = rm;
    return CARE_OK;
}

int care_get_tuning(care_context* ctx, care_tuning* out) {
    if (!ctx || !out) return fail(CARE_ERR_ARGUMENT, "null argument");
    const auto& o = ctx->opt;
    const int T = ctx->blockThreads;
    const int gatedK = correct_gate_mode(o.correction_type, o.correct_candidates, o.correction_type_cands);
    out->batch_anchors = ctx->rs.n ? default_batch(ctx) : ctx->requestedBatch;
    out->correct_threads = T;
    out->msa_chunk = correct_msa_chunk(ctx->rs.maxLen ? ctx->rs.maxLen : 1, o.use_quality, ctx->rs.qpitch, T, gatedK,
                                       ctx->msaChunk);
    out->query_lds_hits = ctx->queryCap;
    // the query kernel sorts up to 1024 hits in registers, 2048 when lists can be longer
    const int regLimit = ctx->maxCand > 1024 ? 2048 : 1024;
    out->query_register_hits = ctx->queryRegMax ? std::min(ctx->queryRegMax, regLimit) : regLimit;
    out->correct_blocks_per_cu = ctx->blocksPerCU;
    if (ctx->blocksPerCU == 0 && ctx->rs.maxLen) {
        try {
            HIP_TRY(hipSetDevice(ctx->device));
            const size_t lds = correct_lds_bytes(ctx->rs.maxLen, o.use_quality, ctx->rs.qpitch, T, gatedK, ctx->msaChunk);
            out->correct_blocks_per_cu = correct_resident_blocks(T, lds, gatedK);
        } catch (const HipError& e) {
            return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
        }
    }
    return CARE_OK;
}

int care_readstore_upload(care_context* ctx, uint64_t num_reads, int max_length, const uint32_t* seqs, int seq_pitch,
                          const int32_t* lengths, const uint8_t* ambiguous, const uint32_t* quals, int qual_pitch,
                          int qual_bits) {
    if (!ctx || !seqs || !lengths) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (num_reads == 0 || num_reads >= 0xFFFFFFFFull) return fail(CARE_ERR_ARGUMENT, "num_reads out of range");
    if (max_length < 1 || max_length > 1000) return fail(CARE_ERR_ARGUMENT, "max_length %d outside 1..1000", max_length);
    if (seq_pitch < (max_length + 15) / 16) return fail(CARE_ERR_ARGUMENT, "seq_pitch too small");
    if (ctx->opt.use_quality) {
        if (!quals) return fail(CARE_ERR_ARGUMENT, "quality scores required with use_quality");
        if (qual_bits != 1 && qual_bits != 2 && qual_bits != 8) return fail(CARE_ERR_ARGUMENT, "qual_bits must be 1, 2 or 8");
        if (qual_pitch < (qual_bits * max_length + 31) / 32) return fail(CARE_ERR_ARGUMENT, "qual_pitch too small");
    }
    for (uint64_t i = 0; i < num_reads; i++)
        if (lengths[i] < 1 || lengths[i] > max_length) return fail(CARE_ERR_ARGUMENT, "read %llu has length %d", (unsigned long long)i, lengths[i]);
    // The device store always uses the minimal pitches (the kernels size
    // their LDS rows from max_length); wider caller rows are repacked.
    const int minPitch = (max_length + 15) / 16;
    const int minQPitch = ctx->opt.use_quality ? (qual_bits * max_length + 31) / 32 : 0;
    std::vector<u32> packedSeqs, packedQuals;
    if (seq_pitch != minPitch) {
        packedSeqs.resize(size_t(num_reads) * minPitch);
        for (uint64_t i = 0; i < num_reads; i++)
            std::memcpy(packedSeqs.data() + i * minPitch, seqs + i * size_t(seq_pitch), sizeof(u32) * minPitch);
        seqs = packedSeqs.data();
        seq_pitch = minPitch;
    }
    if (ctx->opt.use_quality && qual_pitch != minQPitch) {
        packedQuals.resize(size_t(num_reads) * minQPitch);
        for (uint64_t i = 0; i < num_reads; i++)
            std::memcpy(packedQuals.data() + i * minQPitch, quals + i * size_t(qual_pitch), sizeof(u32) * minQPitch);
        quals = packedQuals.data();
        qual_pitch = minQPitch;
    }
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        ctx->seqs.alloc(num_reads * seq_pitch);
        ctx->lens.alloc(num_reads);
        ctx->ambig.alloc(num_reads);
        HIP_TRY(hipMemcpy(ctx->seqs.p, seqs, sizeof(u32) * num_reads * seq_pitch, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->lens.p, lengths, sizeof(int) * num_reads, hipMemcpyHostToDevice));
        if (ambiguous) {
            HIP_TRY(hipMemcpy(ctx->ambig.p, ambiguous, num_reads, hipMemcpyHostToDevice));
        } else {
            std::vector<u8> z(num_reads, 0);
            HIP_TRY(hipMemcpy(ctx->ambig.p, z.data(), num_reads, hipMemcpyHostToDevice));
        }
        if (ctx->opt.use_quality) {
            ctx->quals.alloc(num_reads * qual_pitch);
            HIP_TRY(hipMemcpy(ctx->quals.p, quals, sizeof(u32) * num_reads * qual_pitch, hipMemcpyHostToDevice));
        } else {
            ctx->quals.alloc(1);
            qual_pitch = 0;
        }
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    ctx->rs.seqs = ctx->seqs.p;
    ctx->rs.lens = ctx->lens.p;
    ctx->rs.ambig = ctx->ambig.p;
    ctx->rs.quals = ctx->quals.p;
    ctx->rs.n = num_reads;
    ctx->rs.pitch = seq_pitch;
    ctx->rs.qpitch = qual_pitch;
    ctx->rs.qbits = ctx->opt.use_quality ? qual_bits : 8;
    ctx->rs.maxLen = max_length;
    ctx->hasReads = true;
    ctx->hasTables = false;
    ctx->batchCap = 0;
    return CARE_OK;
}

int care_readstore_gather(care_context* ctx, const uint32_t* ids, int n, uint32_t* out_seqs, int out_pitch,
                          int32_t* out_lengths) {
    if (!ctx || (n > 0 && !ids)) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    if (out_seqs && out_pitch < ctx->rs.pitch) return fail(CARE_ERR_ARGUMENT, "out_pitch too small");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        for (int i = 0; i < n; i++) {
            if (ids[i] >= ctx->rs.n) return fail(CARE_ERR_ARGUMENT, "read id %u out of range", ids[i]);
            if (out_seqs)
                HIP_TRY(hipMemcpy(out_seqs + size_t(i) * out_pitch, ctx->seqs.p + size_t(ids[i]) * ctx->rs.pitch,
                                  sizeof(u32) * ctx->rs.pitch, hipMemcpyDeviceToHost));
            if (out_lengths)
                HIP_TRY(hipMemcpy(out_lengths + i, ctx->lens.p + ids[i], sizeof(int), hipMemcpyDeviceToHost));
        }
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

// ---- stream-ordered device-buffer interface ------------------------------

int care_readstore_gather_sequences_async(care_context* ctx, const uint32_t* d_ids, int n, uint32_t* d_out,
                                          int out_pitch, struct ihipStream_t* stream) {
    if (!ctx || (n > 0 && (!d_ids || !d_out))) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    if (n < 0 || out_pitch < 1) return fail(CARE_ERR_ARGUMENT, "n must be >= 0 and out_pitch >= 1");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        launch_gather_sequences(ctx->rs, d_ids, n, d_out, out_pitch, stream);
        HIP_TRY(hipGetLastError());
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_readstore_gather_lengths_async(care_context* ctx, const uint32_t* d_ids, int n, int32_t* d_out,
                                        struct ihipStream_t* stream) {
    if (!ctx || (n > 0 && (!d_ids || !d_out))) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    if (n < 0) return fail(CARE_ERR_ARGUMENT, "n must be >= 0");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        launch_gather_lengths(ctx->rs, d_ids, n, d_out, stream);
        HIP_TRY(hipGetLastError());
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_readstore_gather_qualities_async(care_context* ctx, const uint32_t* d_ids, int n, char* d_out, int out_pitch,
                                          struct ihipStream_t* stream) {
    if (!ctx || (n > 0 && (!d_ids || !d_out))) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    if (!ctx->opt.use_quality) return fail(CARE_ERR_STATE, "the read store holds no quality scores (use_quality 0)");
    if (n < 0 || out_pitch < 1) return fail(CARE_ERR_ARGUMENT, "n must be >= 0 and out_pitch >= 1");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        launch_gather_qualities(ctx->rs, d_ids, n, d_out, out_pitch, stream);
        HIP_TRY(hipGetLastError());
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_readstore_ambiguous_async(care_context* ctx, const uint32_t* d_ids, int n, uint8_t* d_out,
                                   struct ihipStream_t* stream) {
    if (!ctx || (n > 0 && (!d_ids || !d_out))) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    if (n < 0) return fail(CARE_ERR_ARGUMENT, "n must be >= 0");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        launch_gather_ambiguous(ctx->rs, d_ids, n, d_out, stream);
        HIP_TRY(hipGetLastError());
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_minhash_max_candidates(care_context* ctx) { return ctx ? ctx->maxCand : 0; }

int care_device_alloc(care_context* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(CARE_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(CARE_ERR_DEVICE, "hipSetDevice failed");
    const hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return CARE_OK;
}

int care_device_free(care_context* ctx, void* p) {
    if (!ctx) return fail(CARE_ERR_ARGUMENT, "null context");
    if (!p) return CARE_OK;
    (void)hipSetDevice(ctx->device);
    const hipError_t e = hipFree(p);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipFree: %s", hipGetErrorString(e));
    return CARE_OK;
}

int care_device_copy(care_context* ctx, void* dst, const void* src, size_t bytes, int to_device) {
    if (!ctx || (bytes && (!dst || !src))) return fail(CARE_ERR_ARGUMENT, "null argument");
    (void)hipSetDevice(ctx->device);
    const hipError_t e = hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
    return CARE_OK;
}

int care_stream_create(care_context* ctx, struct ihipStream_t** out) {
    if (!ctx || !out) return fail(CARE_ERR_ARGUMENT, "null argument");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    *out = reinterpret_cast<struct ihipStream_t*>(s);
    return CARE_OK;
}

int care_stream_destroy(care_context* ctx, struct ihipStream_t* stream) {
    if (!ctx || !stream) return fail(CARE_ERR_ARGUMENT, "null argument");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(stream);
    const hipError_t e = hipStreamDestroy(stream);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipStreamDestroy: %s", hipGetErrorString(e));
    return CARE_OK;
}

int care_stream_synchronize(care_context* ctx, struct ihipStream_t* stream) {
    if (!ctx) return fail(CARE_ERR_ARGUMENT, "null context");
    (void)hipSetDevice(ctx->device);
    const hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipStreamSynchronize: %s", hipGetErrorString(e));
    return CARE_OK;
}

namespace {
// workspace of care_minhash_query_async: signature rows (8 h bytes per
// anchor) and the query's spill rows for long hit lists (16 maxCand bytes
// per anchor), 256-byte aligned parts
size_t query_workspace_sig_bytes(const care_context* ctx, int n) {
    return (size_t(n) * size_t(ctx->opt.num_hash_maps) * sizeof(u64) + 255) / 256 * 256;
}
size_t query_workspace_bytes(const care_context* ctx, int n) {
    return query_workspace_sig_bytes(ctx, n) + size_t(n) * size_t(ctx->maxCand) * 4 * sizeof(u32);
}
}  // namespace

size_t care_minhash_query_workspace_bytes(care_context* ctx, int n) {
    return ctx && n > 0 ? query_workspace_bytes(ctx, n) : 0;
}

int care_minhash_query_async(care_context* ctx, const uint32_t* d_seqs, int seq_pitch, const int32_t* d_lengths,
                             const uint32_t* d_read_ids, int n, uint32_t* d_candidates, int32_t* d_counts,
                             void* d_workspace, size_t workspace_bytes, struct ihipStream_t* stream) {
    if (!ctx || (n > 0 && (!d_seqs || !d_lengths || !d_candidates || !d_counts || !d_workspace)))
        return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasTables) return fail(CARE_ERR_STATE, "tables not built");
    if (n < 0 || seq_pitch < 1) return fail(CARE_ERR_ARGUMENT, "n must be >= 0 and seq_pitch >= 1");
    if (n == 0) return CARE_OK;
    if (reinterpret_cast<uintptr_t>(d_workspace) % 256 != 0) return fail(CARE_ERR_ARGUMENT, "workspace must be 256-byte aligned");
    if (workspace_bytes < query_workspace_bytes(ctx, n))
        return fail(CARE_ERR_CAPACITY, "workspace of %zu bytes, %zu needed", workspace_bytes, query_workspace_bytes(ctx, n));
    const auto& o = ctx->opt;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        u64* sigs = static_cast<u64*>(d_workspace);
        u32* spill = reinterpret_cast<u32*>(static_cast<char*>(d_workspace) + query_workspace_sig_bytes(ctx, n));
        // the caller's sequences as a read store of n rows: lengths are cut to
        // what the pitch holds
        DevReadStore in{};
        in.seqs = d_seqs;
        in.lens = d_lengths;
        in.n = u64(n);
        in.pitch = seq_pitch;
        in.maxLen = 16 * seq_pitch;
        launch_signatures(in, 0, n, o.kmer_length, o.num_hash_maps, sigs, stream);
        QueryAnchors qa;
        qa.sigs = sigs;
        qa.ids = d_read_ids;
        qa.lens = d_lengths;
        launch_query(ctx->rs, ctx->tables, 0, n, ctx->maxCand, ctx->queryCap, ctx->queryRegMax, o.exclude_ambiguous,
                     o.candidate_limit, d_candidates, d_counts, spill, nullptr, stream, qa);
        HIP_TRY(hipGetLastError());
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_minhash_signatures(care_context* ctx, uint64_t first, int n, uint64_t* out) {
    if (!ctx || !out) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    if (n < 0 || first + uint64_t(n) > ctx->rs.n) return fail(CARE_ERR_ARGUMENT, "range outside read store");
    if (n == 0) return CARE_OK;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        DevBuf<u64> sig;
        sig.alloc(size_t(n) * ctx->opt.num_hash_maps);
        launch_signatures(ctx->rs, first, n, ctx->opt.kmer_length, ctx->opt.num_hash_maps, sig.p, ctx->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, sig.p, sizeof(u64) * size_t(n) * ctx->opt.num_hash_maps, hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_minhash_build(care_context* ctx) {
    if (!ctx) return fail(CARE_ERR_ARGUMENT, "null context");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    const int h = ctx->opt.num_hash_maps;
    const u64 n = ctx->rs.n;
    // one query workgroup holds up to h * cap hits twice in LDS
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DevBuf<u64>& sigs = ctx->sigs;  // kept: queries read the anchors' signatures
        sigs.alloc(n * h);
        launch_signatures(ctx->rs, 0, int(n), ctx->opt.kmer_length, h, sigs.p, s);
        HIP_TRY(hipGetLastError());
        // scratch table at load <= 0.5
        u64 tsize = 1;
        while (tsize < 2 * n) tsize <<= 1;
        DevBuf<u64> tkeys;
        DevBuf<u32> tcount, tfill, toff;
        tkeys.alloc(tsize);
        tcount.alloc(tsize);
        tfill.alloc(tsize);
        toff.alloc(tsize);
        DevBuf<unsigned long long> counters;
        counters.alloc(2);
        DevBuf<u64> tileSums;
        tileSums.alloc(assign_offsets_scratch(tsize));
        DevBuf<int> overflow;
        overflow.alloc(1);
        ctx->maps.clear();
        std::vector<DevMap> hostMaps(h);
        for (int m = 0; m < h; m++) {
            launch_fill_u64(tkeys.p, tsize, kEmptyKey, s);
            launch_fill_u32(tcount.p, tsize, 0, s);
            launch_fill_u32(tfill.p, tsize, 0, s);
            launch_fill_u64(reinterpret_cast<u64*>(counters.p), 2, 0, s);
            launch_fill_u32(reinterpret_cast<u32*>(overflow.p), 1, 0, s);
            launch_count_keys(sigs.p, n, h, m, tkeys.p, tcount.p, tsize - 1, overflow.p, s);
            launch_assign_offsets(tkeys.p, tcount.p, tsize, ctx->cap, toff.p, tileSums.p, counters.p, s);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(ctx->hU64.p, counters.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            const u64 numValues = ctx->hU64.p[0];
            const u64 numKeys = ctx->hU64.p[1];
            auto ms = std::make_unique<MapStore>();
            u64 fsize = 1;
            while (fsize < 2 * numKeys + 2) fsize <<= 1;
            ms->keys.alloc(fsize);
            ms->offcnt.alloc(fsize);
            ms->values.alloc(numValues);
            ms->mask = fsize - 1;
            ms->numKeys = numKeys;
            ms->numValues = numValues;
            launch_fill_u64(ms->keys.p, fsize, kEmptyKey, s);
            launch_insert_final(tkeys.p, tcount.p, toff.p, tsize, ms->keys.p, ms->offcnt.p, ms->mask, overflow.p, s);
            launch_scatter_values(sigs.p, n, h, m, tkeys.p, toff.p, tsize - 1, tfill.p, ms->values.p, s);
            launch_sort_buckets(toff.p, tcount.p, tsize, ms->values.p, s);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(ctx->hInts.p, overflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (ctx->hInts.p[0]) return fail(CARE_ERR_DEVICE, "hash table overflow while building map %d", m);
            hostMaps[m] = DevMap{ms->keys.p, ms->offcnt.p, ms->values.p, ms->mask};
            ctx->maps.push_back(std::move(ms));
        }
        ctx->devMaps.alloc(h);
        HIP_TRY(hipMemcpy(ctx->devMaps.p, hostMaps.data(), sizeof(DevMap) * h, hipMemcpyHostToDevice));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    ctx->tables.maps = ctx->devMaps.p;
    ctx->tables.h = h;
    ctx->tables.k = ctx->opt.kmer_length;
    ctx->tables.kmask = ~u64(0) >> (2 * (32 - ctx->opt.kmer_length));
    ctx->tables.sigs = ctx->sigs.p;
    ctx->hasTables = true;
    return CARE_OK;
}

int care_minhash_import(care_context* ctx, const care_map_view* maps, int num_maps, int kmer_length,
                        int results_per_map) {
    if (!ctx || !maps) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasReads) return fail(CARE_ERR_STATE, "no read store uploaded");
    if (num_maps < 1 || num_maps > 64) return fail(CARE_ERR_ARGUMENT, "num_maps %d outside 1..64", num_maps);
    if (kmer_length < 1 || kmer_length > 31) return fail(CARE_ERR_ARGUMENT, "kmer_length %d outside 1..31", kmer_length);
    if (results_per_map < 1 || results_per_map > 65535)
        return fail(CARE_ERR_ARGUMENT, "results_per_map %d outside 1..65535", results_per_map);
    const u64 n = ctx->rs.n;
    // every value range inside the map's values, every read id inside the
    // store, keys distinct: the query kernel relies on all three
    for (int m = 0; m < num_maps; m++) {
        const care_map_view& v = maps[m];
        if ((v.num_keys && (!v.keys || !v.offsets || !v.counts)) || (v.num_values && !v.values))
            return fail(CARE_ERR_ARGUMENT, "map %d: null array", m);
        if (v.num_values > 0xFFFFFFFFull) return fail(CARE_ERR_ARGUMENT, "map %d: more than 2^32-1 values", m);
        for (u64 i = 0; i < v.num_keys; i++) {
            if (v.counts[i] < 1 || v.counts[i] > u32(results_per_map))
                return fail(CARE_ERR_ARGUMENT, "map %d: key %llu has %u values, outside 1..%d", m,
                            (unsigned long long)i, v.counts[i], results_per_map);
            if (u64(v.offsets[i]) + v.counts[i] > v.num_values)
                return fail(CARE_ERR_ARGUMENT, "map %d: values of key %llu end beyond num_values", m,
                            (unsigned long long)i);
        }
        for (u64 j = 0; j < v.num_values; j++)
            if (v.values[j] >= n)
                return fail(CARE_ERR_ARGUMENT, "map %d: read id %u outside the store of %llu reads", m, v.values[j],
                            (unsigned long long)n);
        std::vector<u64> sorted(v.keys, v.keys + v.num_keys);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(CARE_ERR_ARGUMENT, "map %d: repeated key", m);
        if (!sorted.empty() && sorted.back() == kEmptyKey)
            return fail(CARE_ERR_ARGUMENT, "map %d: key 2^64-1 is reserved", m);
    }
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        ctx->hasTables = false;
        ctx->maps.clear();
        std::vector<DevMap> hostMaps(num_maps);
        DevBuf<int> overflow;
        overflow.alloc(1);
        for (int m = 0; m < num_maps; m++) {
            const care_map_view& v = maps[m];
            auto ms = std::make_unique<MapStore>();
            u64 fsize = 1;
            while (fsize < 2 * v.num_keys + 2) fsize <<= 1;
            ms->keys.alloc(fsize);
            ms->offcnt.alloc(fsize);
            ms->values.alloc(v.num_values);
            ms->mask = fsize - 1;
            ms->numKeys = v.num_keys;
            ms->numValues = v.num_values;
            DevBuf<u64> dkeys;
            DevBuf<u32> doffs, dcounts;
            dkeys.alloc(v.num_keys);
            doffs.alloc(v.num_keys);
            dcounts.alloc(v.num_keys);
            if (v.num_keys) {
                HIP_TRY(hipMemcpy(dkeys.p, v.keys, sizeof(u64) * v.num_keys, hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(doffs.p, v.offsets, sizeof(u32) * v.num_keys, hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(dcounts.p, v.counts, sizeof(u32) * v.num_keys, hipMemcpyHostToDevice));
            }
            if (v.num_values)
                HIP_TRY(hipMemcpy(ms->values.p, v.values, sizeof(u32) * v.num_values, hipMemcpyHostToDevice));
            launch_fill_u64(ms->keys.p, fsize, kEmptyKey, s);
            launch_fill_u32(reinterpret_cast<u32*>(overflow.p), 1, 0, s);
            if (v.num_keys) {
                launch_insert_triples(dkeys.p, doffs.p, dcounts.p, v.num_keys, ms->keys.p, ms->offcnt.p, ms->mask,
                                      overflow.p, s);
                // the query merges each key's values as ascending lists
                launch_sort_buckets(doffs.p, dcounts.p, v.num_keys, ms->values.p, s);
            }
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(ctx->hInts.p, overflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (ctx->hInts.p[0]) return fail(CARE_ERR_DEVICE, "hash table overflow while importing map %d", m);
            hostMaps[m] = DevMap{ms->keys.p, ms->offcnt.p, ms->values.p, ms->mask};
            ctx->maps.push_back(std::move(ms));
        }
        ctx->devMaps.alloc(num_maps);
        HIP_TRY(hipMemcpy(ctx->devMaps.p, hostMaps.data(), sizeof(DevMap) * num_maps, hipMemcpyHostToDevice));
        ctx->opt.kmer_length = kmer_length;
        ctx->opt.num_hash_maps = num_maps;
        ctx->cap = results_per_map;
        ctx->maxCand = num_maps * results_per_map;
        ctx->queryCap = std::min((ctx->maxCand + 15) / 16 * 16, kDefaultQueryCap);
        ctx->queryRegMax = 0;
        ctx->sigs.alloc(n * num_maps);
        launch_signatures(ctx->rs, 0, int(n), kmer_length, num_maps, ctx->sigs.p, s);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    ctx->tables.maps = ctx->devMaps.p;
    ctx->tables.h = num_maps;
    ctx->tables.k = kmer_length;
    ctx->tables.kmask = ~u64(0) >> (2 * (32 - kmer_length));
    ctx->tables.sigs = ctx->sigs.p;
    ctx->hasTables = true;
    return CARE_OK;
}

int care_minhash_geometry(care_context* ctx, int* kmer_length, int* num_maps, int* results_per_map) {
    if (!ctx || !kmer_length || !num_maps || !results_per_map) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasTables) return fail(CARE_ERR_STATE, "tables not built");
    *kmer_length = ctx->tables.k;
    *num_maps = ctx->tables.h;
    *results_per_map = ctx->cap;
    return CARE_OK;
}

int care_minhash_map_size(care_context* ctx, int map, uint64_t* num_keys, uint64_t* num_values) {
    if (!ctx || !num_keys || !num_values) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasTables) return fail(CARE_ERR_STATE, "tables not built");
    if (map < 0 || map >= int(ctx->maps.size())) return fail(CARE_ERR_ARGUMENT, "map %d outside 0..%d", map,
                                                             int(ctx->maps.size()) - 1);
    *num_keys = ctx->maps[size_t(map)]->numKeys;
    *num_values = ctx->maps[size_t(map)]->numValues;
    return CARE_OK;
}

int care_minhash_export_map(care_context* ctx, int map, uint64_t* keys, uint32_t* offsets, uint32_t* counts,
                            uint32_t* values) {
    uint64_t nk = 0, nv = 0;
    if (const int rc = care_minhash_map_size(ctx, map, &nk, &nv)) return rc;
    if ((nk && (!keys || !offsets || !counts)) || (nv && !values)) return fail(CARE_ERR_ARGUMENT, "null array");
    const MapStore& ms = *ctx->maps[size_t(map)];
    const u64 fsize = ms.mask + 1;
    std::vector<u64> slotKeys(fsize);
    std::vector<uint2> slotOffcnt(fsize);
    std::vector<u32> slotValues(nv);
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(slotKeys.data(), ms.keys.p, sizeof(u64) * fsize, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(slotOffcnt.data(), ms.offcnt.p, sizeof(uint2) * fsize, hipMemcpyDeviceToHost));
        if (nv) HIP_TRY(hipMemcpy(slotValues.data(), ms.values.p, sizeof(u32) * nv, hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    std::vector<u64> order;
    order.reserve(nk);
    for (u64 sl = 0; sl < fsize; sl++)
        if (slotKeys[sl] != kEmptyKey) order.push_back(sl);
    if (order.size() != nk) return fail(CARE_ERR_STATE, "map %d holds %zu keys, expected %llu", map, order.size(),
                                        (unsigned long long)nk);
    std::sort(order.begin(), order.end(), [&](u64 a, u64 b) { return slotKeys[a] < slotKeys[b]; });
    u64 next = 0;
    for (u64 i = 0; i < nk; i++) {
        const u64 sl = order[i];
        const uint2 oc = slotOffcnt[sl];
        if (u64(oc.x) + oc.y > nv || next + oc.y > nv)
            return fail(CARE_ERR_STATE, "map %d: inconsistent value range", map);
        keys[i] = slotKeys[sl];
        offsets[i] = u32(next);
        counts[i] = oc.y;
        std::memcpy(values + next, slotValues.data() + oc.x, sizeof(u32) * oc.y);
        next += oc.y;
    }
    return CARE_OK;
}

int care_minhash_query(care_context* ctx, uint64_t first, int n, uint32_t* ids, int32_t* offsets,
                       uint64_t ids_capacity) {
    if (!ctx || !offsets || (ids_capacity > 0 && !ids)) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasTables) return fail(CARE_ERR_STATE, "tables not built");
    if (n < 0 || first + uint64_t(n) > ctx->rs.n) return fail(CARE_ERR_ARGUMENT, "range outside read store");
    offsets[0] = 0;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const int B = default_batch(ctx);
        ensure_batch(ctx, B);
        std::vector<int> counts;
        std::vector<u32> buf;
        uint64_t written = 0;
        for (int done = 0; done < n; done += B) {
            const int nb = std::min(B, n - done);
            launch_query(ctx->rs, ctx->tables, first + done, nb, ctx->maxCand, ctx->queryCap, ctx->queryRegMax,
                         ctx->opt.exclude_ambiguous, ctx->opt.candidate_limit,
                         ctx->candIds.p, ctx->candCount.p, reinterpret_cast<u32*>(ctx->metaA.p), nullptr,
                         ctx->stream);
            HIP_TRY(hipGetLastError());
            counts.resize(nb);
            buf.resize(size_t(nb) * ctx->maxCand);
            HIP_TRY(hipMemcpyAsync(counts.data(), ctx->candCount.p, sizeof(int) * nb, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipMemcpyAsync(buf.data(), ctx->candIds.p, sizeof(u32) * size_t(nb) * ctx->maxCand,
                                   hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            for (int i = 0; i < nb; i++) {
                const int c = counts[i];
                if (written + c > ids_capacity) return fail(CARE_ERR_CAPACITY, "ids_capacity exceeded");
                std::memcpy(ids + written, buf.data() + size_t(i) * ctx->maxCand, sizeof(u32) * c);
                written += c;
                offsets[done + i + 1] = int32_t(written);
            }
        }
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

}  // extern "C"

namespace {
int correct_batch_impl(care_context* ctx, uint64_t first, int n, care_batch_result* r, bool graph);

// rows of `width` bytes from a device buffer with pitch spitch to host rows
// with pitch dpitch: one contiguous copy when the pitches agree
void copy_rows_to_host(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t rows,
                       hipStream_t s) {
    if (rows == 0) return;
    if (dpitch == spitch) {
        HIP_TRY(hipMemcpyAsync(dst, src, (rows - 1) * spitch + width, hipMemcpyDeviceToHost, s));
    } else {
        HIP_TRY(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToHost, s));
    }
}

// Page-locked host memory (hipHostMalloc): an asynchronous copy into it runs
// on a copy engine without blocking the host; a copy into pageable memory
// blocks the host until the copy is done.
bool is_page_locked(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory may be reported as an error
        return false;
    }
    return a.type == hipMemoryTypeHost;
}
}

extern "C" int care_correct_batch(care_context* ctx, uint64_t first, int n, care_batch_result* r) {
    return correct_batch_impl(ctx, first, n, r, false);
}

extern "C" int care_correct_batch_graph(care_context* ctx, uint64_t first, int n, care_batch_result* r) {
    return correct_batch_impl(ctx, first, n, r, true);
}

namespace {
// Scalar parameters and model pointers of the correction kernel from the
// context's options (device buffers are filled in by the caller).
int prepare_params(care_context* ctx, CorrectParams& P) {
    const auto& o = ctx->opt;
    // MSA base counts are 16-bit: an anchor has at most maxCand candidates
    if (ctx->maxCand > 65534)
        return fail(CARE_ERR_ARGUMENT, "%d candidates per anchor (maps x results per map) exceed 65534",
                    ctx->maxCand);
    P.rs = ctx->rs;
    P.maxCand = ctx->maxCand;
    P.qweights = ctx->qweights.p;
    P.pairedEnd = o.paired_end;
    P.useQuality = o.use_quality;
    P.correctCandidates = o.correct_candidates;
    P.newColumns = o.new_columns_to_correct;
    P.minOverlap = o.min_overlap;
    P.maxErrorRate = o.max_mismatch_ratio;
    P.minOverlapRatio = o.min_overlap_ratio;
    P.maxErr4 = o.error_factor * 4.0f;
    const float base = o.error_factor * 1.0f;
    P.mmBase2 = 2 * base;
    P.mmBase3 = 3 * base;
    P.mmBase4 = 4 * base;
    P.goodAlignmentsThreshold = o.coverage * o.coverage_factor;
    P.pairedThreshold = o.paired_filter_threshold;
    P.sigCountThreshold = int(float(int(o.coverage)) * 0.3f);  // significance count, src/msa.cpp:537-541
    P.avgSupportThreshold = 1.0f - o.error_factor;
    P.minSupportThreshold = std::fma(-3.0f, o.error_factor, 1.0f);
    P.minCoverageThreshold = o.coverage_factor / 6.0f * o.coverage;
    P.errPercent = int(std::ceil(o.error_factor * 100.0f));
    P.errPercentHalf = float(P.errPercent) * 0.5f;
    P.correctionType = o.correction_type;
    P.correctionTypeCands = o.correction_type_cands;
    P.estimatedCoverage = o.coverage;
    P.maxRefinements = o.max_refinements;
    if (o.correction_type == 1) {
        if (!ctx->forest[0].loaded) return fail(CARE_ERR_STATE, "correction_type 1 needs an anchor forest (care_forest_load)");
        P.forestA = ctx->forest[0].dev;
    }
    if (o.correction_type == 3) {
        if (!ctx->gate.loaded && !ctx->record.on)
            return fail(CARE_ERR_STATE, "correction_type 3 needs gate weights (care_mlp_load)");
        if (ctx->gate.loaded) {
            P.gateW1 = ctx->gate.w1.p;
            P.gateW2 = ctx->gate.w2.p;
            P.gateW3 = ctx->gate.w3.p;
            P.gateParams = ctx->gate.params.p;
        }
        const float t = std::min(std::max(o.threshold_anchor, 1e-6f), 1.0f - 1e-6f);
        P.gateLogit = std::log(t / (1.0f - t));
        if (ctx->record.on) {
            P.gatedRefine = ctx->record.refine ? 1 : 0;
            P.gateFeat = ctx->record.feat.p;
            P.gateIds = ctx->record.ids.p;
            P.gatePos = ctx->record.pos.p;
            P.gateOrig = ctx->record.orig.p;
            P.gateCons = ctx->record.cons.p;
            P.gateCount = ctx->counters.p + 2;
            P.gateCap = ctx->record.cap;
            P.gateRecordCands = ctx->record.cands ? 1 : 0;
        }
        if (o.correct_candidates && o.correction_type_cands == 3) {
            // the candidate gate's own weights if loaded, else the anchor gate's
            const auto& gc = ctx->gateCands.loaded ? ctx->gateCands : ctx->gate;
            if (gc.loaded) {
                P.gateCW1 = gc.w1.p;
                P.gateCW2 = gc.w2.p;
                P.gateCW3 = gc.w3.p;
                P.gateCParams = gc.params.p;
            }
            const float tc = std::min(std::max(o.threshold_cands, 1e-6f), 1.0f - 1e-6f);
            P.gateCLogit = std::log(tc / (1.0f - tc));
        }
    }
    if (o.correct_candidates && o.correction_type_cands == 1) {
        if (!ctx->forest[1].loaded) return fail(CARE_ERR_STATE, "correction_type_cands 1 needs a candidate forest (care_forest_load)");
        P.forestC = ctx->forest[1].dev;
    }
    return CARE_OK;
}

// bytes of every value in `vals`, appended to key
template <class... T>
void key_append(std::vector<unsigned char>& key, const T&... vals) {
    auto one = [&](const auto& v) {
        const auto* b = reinterpret_cast<const unsigned char*>(&v);
        key.insert(key.end(), b, b + sizeof(v));
    };
    (one(vals), ...);
}

// Size of the next device batch of a call over n anchors with rem still to
// go: batches of B, except that a call of more than B anchors ends in a
// batch of B / 8, so the result copy that no later batch overlaps is short.
static int next_batch(int B, int n, int rem) {
    const int tail = (B / 8) & ~1;
    if (n > B && tail >= 2 && rem > tail && rem <= B + tail) return rem - tail;
    return std::min(B, rem);
}

int correct_batch_impl(care_context* ctx, uint64_t first, int n, care_batch_result* r, bool graph) {
    if (!ctx || !r || !r->anchor_status || !r->anchor_sequences) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasTables) return fail(CARE_ERR_STATE, "tables not built");
    if (n < 0 || first + uint64_t(n) > ctx->rs.n) return fail(CARE_ERR_ARGUMENT, "range outside read store");
    if (r->sequence_pitch < ctx->rs.maxLen) return fail(CARE_ERR_ARGUMENT, "sequence_pitch < max read length");
    if (ctx->opt.paired_end && ((first & 1) || (n & 1)))
        return fail(CARE_ERR_ARGUMENT, "paired-end batches must start at an even id and hold whole pairs");
    if (ctx->opt.correct_candidates && (!r->candidate_read_ids || !r->candidate_sequences))
        return fail(CARE_ERR_ARGUMENT, "candidate output buffers required with correct_candidates");
    r->num_candidates = 0;
    ctx->stats = care_batch_stats{};
    const auto& o = ctx->opt;
    CorrectParams P{};
    if (const int rc = prepare_params(ctx, P)) return rc;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const int B = default_batch(ctx);
        ensure_batch(ctx, B);
        const int T = ctx->blockThreads;
        const int gatedK = correct_gate_mode(o.correction_type, o.correct_candidates, o.correction_type_cands);
        const size_t lds = correct_lds_bytes(ctx->rs.maxLen, o.use_quality, ctx->rs.qpitch, T, gatedK, ctx->msaChunk);
        if (lds > size_t(kMaxLdsBytes))
            return fail(CARE_ERR_CAPACITY, "reads of length %d need %zu bytes of LDS per correction workgroup (limit %d)",
                        ctx->rs.maxLen, lds, kMaxLdsBytes);
        const int blocksPerCU = ctx->blocksPerCU > 0 ? ctx->blocksPerCU
                                                     : correct_resident_blocks(T, lds, gatedK | (P.gatedRefine ? 4 : 0));
        hipStream_t s = ctx->stream;
        P.candIds = ctx->candIds.p;
        P.candCount = ctx->candCount.p;
        P.metaA = ctx->metaA.p;
        P.metaB = ctx->metaB.p;
        P.outStatus = ctx->outStatus.p;
        P.outSeq = ctx->outSeq.p;
        P.outPitch = ctx->outPitch;
        P.candOutIds = ctx->candOutIds.p;
        P.candOutSeq = ctx->candOutSeq.p;
        P.candOutCount = ctx->counters.p + 1;
        P.candOutCap = ctx->candOutCap;
        P.queueHead = ctx->counters.p;
        // Every device batch writes its slice of call-wide anchor outputs and
        // appends its candidate corrections to a call-wide candidate buffer
        // (one counter over the call), so the batches run back to back with
        // no host synchronisation and the results come back in a few copies
        // at the end.  Candidate results come in no particular order: the
        // per-read combination of results does not depend on it.
        const int nBatches = (n + B - 1) / B + 1;  // at most, with the short final batch
        if (n > ctx->outCap) {
            ctx->outStatus.alloc(size_t(n));
            ctx->outSeq.alloc(size_t(n) * ctx->outPitch);
            ctx->outCap = n;
        }
        const int candCap = o.correct_candidates ? std::max(0, r->candidate_capacity) : 0;
        if (candCap > ctx->candOutCap) {
            ctx->candOutIds.alloc(size_t(candCap));
            ctx->candOutSeq.alloc(size_t(candCap) * ctx->outPitch);
            ctx->candOutCap = candCap;
        }
        P.candOutIds = ctx->candOutIds.p;
        P.candOutSeq = ctx->candOutSeq.p;
        P.candOutCap = candCap;  // results beyond are dropped and reported below
        while (int(ctx->batchEv.size()) < 3 * nBatches) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            ctx->batchEv.push_back(e);
        }
        // into page-locked destinations each batch's anchor results are copied
        // as soon as the batch is done, overlapping the next batches; into
        // pageable ones once at the end (a pageable copy would stall the
        // host between batches)
        const bool overlapCopies = is_page_locked(r->anchor_status) && is_page_locked(r->anchor_sequences);
        // graph mode: the device work below (kernels, the anchor result
        // copies into the caller's page-locked arrays and the counter copies)
        // is captured once and replayed while the calls repeat it; the
        // per-kernel timing events are left out of the graph
        auto& gc = ctx->graphCache;
        std::vector<unsigned char> key;
        if (graph) {
            if (!overlapCopies)
                return fail(CARE_ERR_ARGUMENT, "care_correct_batch_graph needs page-locked anchor_status and "
                                               "anchor_sequences (care_host_buffer_alloc)");
            key_append(key, first, n, r->anchor_status, r->anchor_sequences, r->sequence_pitch, candCap, B, T,
                       blocksPerCU, ctx->msaChunk, ctx->queryCap, ctx->queryRegMax, ctx->numCUs, ctx->outPitch, o,
                       P, ctx->tables, ctx->rs, ctx->candIds.p, ctx->candCount.p, ctx->metaA.p, ctx->rawCount.p,
                       ctx->rawHits.p, ctx->counters.p, ctx->outStatus.p, ctx->outSeq.p);
        }
        const bool replay = graph && gc.exec && gc.key == key;
        HIP_TRY(hipEventRecord(ctx->ev[2], s));
        int bi = 0;
        if (replay) {
            bi = gc.launches;
            ctx->stats.launches = gc.launches;
        } else {
            if (graph) {
                gc.release();
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            }
            try {
                launch_fill_u64(reinterpret_cast<u64*>(ctx->rawHits.p), 2, 0, s);  // totals over the call
                launch_reset_counters(ctx->counters.p, ctx->counters.p + 1, s);  // candidate results over the call
                for (int done = 0, nb = 0; done < n; done += nb, bi++) {
                    nb = next_batch(B, n, n - done);
                    hipEvent_t* e = &ctx->batchEv[size_t(3 * bi)];
                    if (!graph) HIP_TRY(hipEventRecord(e[0], s));
                    launch_query(ctx->rs, ctx->tables, first + done, nb, ctx->maxCand, ctx->queryCap,
                                 ctx->queryRegMax, o.exclude_ambiguous, o.candidate_limit, ctx->candIds.p,
                                 ctx->candCount.p, reinterpret_cast<u32*>(ctx->metaA.p), ctx->rawCount.p, s);
                    launch_sum_counts(ctx->rawCount.p, ctx->candCount.p, nb, ctx->rawHits.p, s);
                    if (!graph) HIP_TRY(hipEventRecord(e[1], s));
                    launch_reset_counters(ctx->counters.p, nullptr, s);  // work queue of this batch
                    P.firstAnchor = first + done;
                    P.numAnchors = nb;
                    P.outStatus = ctx->outStatus.p + size_t(done);
                    P.outSeq = ctx->outSeq.p + size_t(done) * ctx->outPitch;
                    const int blocks = std::max(1, std::min(nb, ctx->numCUs * blocksPerCU));
                    launch_correct(P, blocks, T, ctx->msaChunk, s);
                    HIP_TRY(hipGetLastError());
                    HIP_TRY(hipEventRecord(e[2], s));
                    ctx->stats.launches++;
                    if (overlapCopies) {
                        hipStream_t cs = ctx->copyStream;
                        HIP_TRY(hipStreamWaitEvent(cs, e[2], 0));
                        HIP_TRY(hipMemcpyAsync(r->anchor_status + done, ctx->outStatus.p + done, size_t(nb),
                                               hipMemcpyDeviceToHost, cs));
                        copy_rows_to_host(r->anchor_sequences + size_t(done) * size_t(r->sequence_pitch),
                                          size_t(r->sequence_pitch),
                                          ctx->outSeq.p + size_t(done) * size_t(ctx->outPitch), size_t(ctx->outPitch),
                                          size_t(ctx->rs.maxLen), size_t(nb), cs);
                    }
                }
                if (overlapCopies) {
                    HIP_TRY(hipEventRecord(ctx->copyDone, ctx->copyStream));
                    HIP_TRY(hipStreamWaitEvent(s, ctx->copyDone, 0));
                } else {
                    HIP_TRY(hipMemcpyAsync(r->anchor_status, ctx->outStatus.p, size_t(n), hipMemcpyDeviceToHost, s));
                    copy_rows_to_host(r->anchor_sequences, size_t(r->sequence_pitch), ctx->outSeq.p,
                                      size_t(ctx->outPitch), size_t(ctx->rs.maxLen), size_t(n), s);
                }
                if (graph) {
                    HIP_TRY(hipMemcpyAsync(ctx->hInts.p, ctx->counters.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
                    HIP_TRY(hipMemcpyAsync(ctx->hU64.p, ctx->rawHits.p, 2 * sizeof(unsigned long long),
                                           hipMemcpyDeviceToHost, s));
                    HIP_TRY(hipStreamEndCapture(s, &gc.graph));
                    HIP_TRY(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
                    gc.key = key;
                    gc.launches = bi;
                }
            } catch (const HipError&) {
                if (graph) {
                    hipGraph_t g = nullptr;
                    if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
                    gc.release();
                }
                throw;
            }
        }
        if (graph) HIP_TRY(hipGraphLaunch(gc.exec, s));
        if (o.correct_candidates) {
            if (!graph) HIP_TRY(hipMemcpyAsync(ctx->hInts.p, ctx->counters.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            const int nc = ctx->hInts.p[1];
            if (nc > candCap) return fail(CARE_ERR_CAPACITY, "candidate_capacity exceeded (%d results)", nc);
            if (nc > 0) {
                HIP_TRY(hipMemcpyAsync(r->candidate_read_ids, ctx->candOutIds.p, sizeof(u32) * size_t(nc),
                                       hipMemcpyDeviceToHost, s));
                copy_rows_to_host(r->candidate_sequences, size_t(r->sequence_pitch), ctx->candOutSeq.p,
                                  size_t(ctx->outPitch), size_t(ctx->rs.maxLen), size_t(nc), s);
            }
            r->num_candidates = nc;
        }
        if (!graph)
            HIP_TRY(hipMemcpyAsync(ctx->hU64.p, ctx->rawHits.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   s));
        HIP_TRY(hipEventRecord(ctx->ev[3], s));
        HIP_TRY(hipEventSynchronize(ctx->ev[3]));
        float qms = 0, cms = 0;
        for (int i = 0; i < (graph ? 0 : bi); i++) {
            float t1 = 0, t2 = 0;
            HIP_TRY(hipEventElapsedTime(&t1, ctx->batchEv[size_t(3 * i)], ctx->batchEv[size_t(3 * i + 1)]));
            HIP_TRY(hipEventElapsedTime(&t2, ctx->batchEv[size_t(3 * i + 1)], ctx->batchEv[size_t(3 * i + 2)]));
            qms += t1;
            cms += t2;
        }
        ctx->stats.raw_hits = ctx->hU64.p[0];
        ctx->stats.candidates = ctx->hU64.p[1];
        float tot = 0;
        HIP_TRY(hipEventElapsedTime(&tot, ctx->ev[2], ctx->ev[3]));
        ctx->stats.query_ms = qms;
        ctx->stats.correct_ms = cms;
        ctx->stats.total_ms = tot;
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

}  // namespace

extern "C" {

int care_last_batch_stats(care_context* ctx, care_batch_stats* stats) {
    if (!ctx || !stats) return fail(CARE_ERR_ARGUMENT, "null argument");
    *stats = ctx->stats;
    return CARE_OK;
}

int care_host_buffer_alloc(size_t bytes, void** out) {
    if (!out) return fail(CARE_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    const hipError_t e = hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return CARE_OK;
}

int care_host_buffer_free(void* p) {
    if (!p) return CARE_OK;
    const hipError_t e = hipHostFree(p);
    if (e != hipSuccess) return fail(CARE_ERR_DEVICE, "hipHostFree: %s", hipGetErrorString(e));
    return CARE_OK;
}

int care_forest_load(care_context* ctx, int which, const char* path, uint32_t max_trees) {
    // file format: include/forest.hpp:38-55 and 126-150, ml/care.py:198-282
    if (!ctx || !path) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (which != 0 && which != 1) return fail(CARE_ERR_ARGUMENT, "which must be 0 (anchor) or 1 (candidates)");
    const char* expected = which == 0 ? "13 extract_anchor_v3" : "14 extract_cands_v3";
    const int nfeat = which == 0 ? 13 : 14;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CARE_ERR_ARGUMENT, "Loading classifier file failed! %s", path);
    std::vector<unsigned char> buf;
    {
        unsigned char tmp[65536];
        size_t r;
        while ((r = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + r);
        std::fclose(f);
    }
    size_t pos = 0;
    bool bad = false;
    auto take = [&](void* dst, size_t n) {
        if (pos + n > buf.size()) {
            bad = true;
            std::memset(dst, 0, n);
            return;
        }
        std::memcpy(dst, buf.data() + pos, n);
        pos += n;
    };
    uint64_t dlen = 0;
    take(&dlen, 8);
    if (bad || dlen > 4096) return fail(CARE_ERR_ARGUMENT, "%s: truncated classifier header", path);
    std::string desc(size_t(dlen), '\0');
    take(&desc[0], size_t(dlen));
    if (bad || desc != expected)
        return fail(CARE_ERR_ARGUMENT, "Classifier and extractor descriptors do not match! Expected: %s Received: %s",
                    expected, desc.c_str());
    uint32_t ntrees = 0;
    take(&ntrees, 4);
    if (bad) return fail(CARE_ERR_ARGUMENT, "%s: truncated classifier", path);
    ntrees = std::min(ntrees, max_trees);
    std::vector<ForestNode> nodes;
    std::vector<u32> starts;
    // iterative preorder parse with an explicit stack of (node, side)
    for (uint32_t t = 0; t < ntrees && !bad; t++) {
        uint32_t reserve = 0;
        take(&reserve, 4);
        const u32 base = u32(nodes.size());
        starts.push_back(base);
        struct Pending {
            u32 node;
            int side;  // 0: left pending, 1: right pending
        };
        std::vector<Pending> stack;
        auto readNode = [&]() -> u32 {
            ForestNode nd{};
            take(&nd.att, 1);
            take(&nd.split, 8);
            take(&nd.flag, 1);
            nodes.push_back(nd);
            return u32(nodes.size() - 1 - base);
        };
        u32 root = readNode();
        stack.push_back({root, 0});
        while (!stack.empty() && !bad) {
            Pending& pd = stack.back();
            ForestNode& nd = nodes[base + pd.node];
            if (nd.att >= nfeat) {
                bad = true;
                break;
            }
            const bool leaf = pd.side == 0 ? (nd.flag & 2) : (nd.flag & 1);
            if (leaf) {
                float p;
                take(&p, 4);
                u32 bits;
                std::memcpy(&bits, &p, 4);
                if (pd.side == 0) nd.lhs = bits;
                else nd.rhs = bits;
                if (pd.side == 0) pd.side = 1;
                else stack.pop_back();
            } else {
                const u32 child = u32(nodes.size() - base);
                if (pd.side == 0) {
                    nodes[base + pd.node].lhs = child;
                    pd.side = 1;
                } else {
                    nodes[base + pd.node].rhs = child;
                    stack.pop_back();
                }
                readNode();
                stack.push_back({child, 0});
            }
        }
    }
    if (bad) return fail(CARE_ERR_ARGUMENT, "%s: malformed or truncated classifier", path);
    auto& st = ctx->forest[which];
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        st.nodes.alloc(std::max<size_t>(1, nodes.size()));
        st.treeStart.alloc(std::max<size_t>(1, starts.size()));
        if (!nodes.empty())
            HIP_TRY(hipMemcpy(st.nodes.p, nodes.data(), sizeof(ForestNode) * nodes.size(), hipMemcpyHostToDevice));
        if (!starts.empty())
            HIP_TRY(hipMemcpy(st.treeStart.p, starts.data(), sizeof(u32) * starts.size(), hipMemcpyHostToDevice));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    st.dev.nodes = st.nodes.p;
    st.dev.treeStart = st.treeStart.p;
    st.dev.numTrees = int(starts.size());
    st.dev.threshold = which == 0 ? ctx->opt.threshold_anchor : ctx->opt.threshold_cands;
    st.loaded = true;
    return CARE_OK;
}

// ---------------------------------------------------------------------------
// Learned gate: parameters, training, inference, recorded decision columns
// ---------------------------------------------------------------------------

namespace {
int load_gate(care_context* ctx, const float* params, bool cands);
}

int care_mlp_load(care_context* ctx, const float* params) { return load_gate(ctx, params, false); }

int care_mlp_load_candidates(care_context* ctx, const float* params) { return load_gate(ctx, params, true); }

}  // extern "C"

namespace {
int load_gate(care_context* ctx, const float* params, bool cands) {
    if (!ctx || !params) return fail(CARE_ERR_ARGUMENT, "null argument");
    for (int i = 0; i < CARE_MLP_PARAMS; i++)
        if (!std::isfinite(params[i])) return fail(CARE_ERR_ARGUMENT, "parameter %d is not finite", i);
    auto& g = cands ? ctx->gateCands : ctx->gate;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        g.params.alloc(CARE_MLP_PARAMS);
        g.w1.alloc(64 * 32);
        g.w2.alloc(32 * 64);
        g.w3.alloc(16 * 32);
        HIP_TRY(hipMemcpy(g.params.p, params, sizeof(float) * CARE_MLP_PARAMS, hipMemcpyHostToDevice));
        launch_mlp_cast(g.params.p, g.w1.p, g.w2.p, g.w3.p, ctx->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    g.loaded = true;
    return CARE_OK;
}
}  // namespace

extern "C" {

extern "C++" {
namespace {

// features rounded to bf16 on the host, exactly as the device rounds
std::vector<unsigned short> to_bf16(const float* x, uint64_t n) {
    std::vector<unsigned short> out(size_t(n) * CARE_MLP_FEATURES);
    for (size_t i = 0; i < out.size(); i++) {
        unsigned u;
        std::memcpy(&u, &x[i], 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        out[i] = (unsigned short)(u >> 16);
    }
    return out;
}

bool all_finite(const float* x, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(x[i])) return false;
    return true;
}

}  // namespace
}  // extern "C++"

int care_mlp_gradient(care_context* ctx, const float* features, const uint8_t* labels, int n, const float* params,
                      float* grad_out, float* loss_out) {
    if (!ctx || !features || !labels || !params || !grad_out) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (n < 1 || n > (1 << 20)) return fail(CARE_ERR_ARGUMENT, "n %d outside 1..2^20", n);
    if (!all_finite(features, size_t(n) * CARE_MLP_FEATURES) || !all_finite(params, CARE_MLP_PARAMS))
        return fail(CARE_ERR_ARGUMENT, "non-finite input");
    for (int i = 0; i < n; i++)
        if (labels[i] > 1) return fail(CARE_ERR_ARGUMENT, "label %d is not 0 or 1", i);
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const std::vector<unsigned short> xb = to_bf16(features, uint64_t(n));
        std::vector<u32> perm(static_cast<size_t>(n));
        for (int i = 0; i < n; i++) perm[size_t(i)] = u32(i);
        DevBuf<unsigned short> X, w1, w2, w3;
        DevBuf<u8> Y;
        DevBuf<u32> P;
        DevBuf<float> master, partial, out;
        X.alloc(xb.size());
        Y.alloc(size_t(n));
        P.alloc(size_t(n));
        master.alloc(CARE_MLP_PARAMS);
        w1.alloc(64 * 32);
        w2.alloc(32 * 64);
        w3.alloc(16 * 32);
        const int np = mlp_partials(n);
        partial.alloc(size_t(np) * mlp_partial_stride());
        out.alloc(CARE_MLP_PARAMS + 1);
        HIP_TRY(hipMemcpy(X.p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(Y.p, labels, size_t(n), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P.p, perm.data(), perm.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(master.p, params, sizeof(float) * CARE_MLP_PARAMS, hipMemcpyHostToDevice));
        launch_mlp_cast(master.p, w1.p, w2.p, w3.p, s);
        launch_mlp_grad(X.p, Y.p, P.p, 0, n, w1.p, w2.p, w3.p, master.p, 1.0f / float(n), partial.p, s);
        launch_mlp_reduce(partial.p, np, 0, out.p, nullptr, nullptr, nullptr, 0, 0, 0, 0, 1, 1, nullptr, nullptr,
                          nullptr, nullptr, s);
        HIP_TRY(hipGetLastError());
        std::vector<float> res(CARE_MLP_PARAMS + 1);
        HIP_TRY(hipMemcpyAsync(res.data(), out.p, res.size() * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        std::memcpy(grad_out, res.data(), sizeof(float) * CARE_MLP_PARAMS);
        if (loss_out) *loss_out = res[CARE_MLP_PARAMS];
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_mlp_train(care_context* ctx, const float* features, const uint8_t* labels, uint64_t n,
                   const care_mlp_train_options* o, float* params, float* epoch_loss) {
    if (!ctx || !features || !labels || !o || !params) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (n < 1 || n >= 0xFFFFFFFFull) return fail(CARE_ERR_ARGUMENT, "n out of range");
    if (o->epochs < 1 || o->batch < 16 || o->batch > (1 << 16) || !(o->lr > 0.f))
        return fail(CARE_ERR_ARGUMENT, "invalid training options");
    if (!all_finite(params, CARE_MLP_PARAMS)) return fail(CARE_ERR_ARGUMENT, "non-finite initial parameters");
    if (!all_finite(features, size_t(n) * CARE_MLP_FEATURES)) return fail(CARE_ERR_ARGUMENT, "non-finite features");
    for (uint64_t i = 0; i < n; i++)
        if (labels[i] > 1) return fail(CARE_ERR_ARGUMENT, "label %llu is not 0 or 1", (unsigned long long)i);
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const std::vector<unsigned short> xb = to_bf16(features, n);
        DevBuf<unsigned short> X, w1, w2, w3;
        DevBuf<u8> Y;
        DevBuf<u32> P;
        DevBuf<float> master, m1, m2, partial, loss;
        X.alloc(xb.size());
        Y.alloc(size_t(n));
        P.alloc(size_t(n));
        master.alloc(CARE_MLP_PARAMS);
        m1.alloc(CARE_MLP_PARAMS);
        m2.alloc(CARE_MLP_PARAMS);
        w1.alloc(64 * 32);
        w2.alloc(32 * 64);
        w3.alloc(16 * 32);
        const int np = mlp_partials(o->batch);
        partial.alloc(size_t(np) * mlp_partial_stride());
        loss.alloc(1);
        HIP_TRY(hipMemcpy(X.p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(Y.p, labels, size_t(n), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(master.p, params, sizeof(float) * CARE_MLP_PARAMS, hipMemcpyHostToDevice));
        std::vector<float> zeros(CARE_MLP_PARAMS, 0.f);
        HIP_TRY(hipMemcpy(m1.p, zeros.data(), zeros.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(m2.p, zeros.data(), zeros.size() * 4, hipMemcpyHostToDevice));
        launch_mlp_cast(master.p, w1.p, w2.p, w3.p, s);
        std::vector<u32> perm(static_cast<size_t>(n));
        uint64_t rng = o->seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        auto next = [&]() {
            // splitmix64
            uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        int step = 0;
        for (int e = 0; e < o->epochs; e++) {
            for (uint64_t i = 0; i < n; i++) perm[size_t(i)] = u32(i);
            for (uint64_t i = n - 1; i > 0; i--) std::swap(perm[size_t(i)], perm[size_t(next() % (i + 1))]);
            HIP_TRY(hipMemcpyAsync(P.p, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, s));
            const float zero = 0.f;
            HIP_TRY(hipMemcpyAsync(loss.p, &zero, 4, hipMemcpyHostToDevice, s));
            int batches = 0;
            for (uint64_t first = 0; first < n; first += uint64_t(o->batch)) {
                const int rows = int(std::min<uint64_t>(uint64_t(o->batch), n - first));
                step++;
                const float bc1 = 1.0f - std::pow(o->beta1, float(step));
                const float bc2 = 1.0f - std::pow(o->beta2, float(step));
                launch_mlp_grad(X.p, Y.p, P.p, first, rows, w1.p, w2.p, w3.p, master.p, 1.0f / float(rows), partial.p,
                                s);
                launch_mlp_reduce(partial.p, mlp_partials(rows), 1, nullptr, master.p, m1.p, m2.p, o->lr, o->beta1,
                                  o->beta2, o->eps, bc1, bc2, w1.p, w2.p, w3.p, loss.p, s);
                batches++;
            }
            HIP_TRY(hipGetLastError());
            float l = 0.f;
            HIP_TRY(hipMemcpyAsync(&l, loss.p, 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (epoch_loss) epoch_loss[e] = l / float(batches);
        }
        HIP_TRY(hipMemcpy(params, master.p, sizeof(float) * CARE_MLP_PARAMS, hipMemcpyDeviceToHost));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

int care_mlp_infer(care_context* ctx, const float* features, uint64_t n, const float* params, float* logits) {
    if (!ctx || !features || !params || !logits) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (n < 1 || n > (uint64_t(1) << 32)) return fail(CARE_ERR_ARGUMENT, "n out of range");
    if (!all_finite(features, size_t(n) * CARE_MLP_FEATURES) || !all_finite(params, CARE_MLP_PARAMS))
        return fail(CARE_ERR_ARGUMENT, "non-finite input");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const std::vector<unsigned short> xb = to_bf16(features, n);
        DevBuf<unsigned short> X, w1, w2, w3;
        DevBuf<float> master, z;
        X.alloc(xb.size());
        master.alloc(CARE_MLP_PARAMS);
        w1.alloc(64 * 32);
        w2.alloc(32 * 64);
        w3.alloc(16 * 32);
        z.alloc(size_t(n));
        HIP_TRY(hipMemcpy(X.p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(master.p, params, sizeof(float) * CARE_MLP_PARAMS, hipMemcpyHostToDevice));
        launch_mlp_cast(master.p, w1.p, w2.p, w3.p, s);
        launch_mlp_forward(X.p, n, w1.p, w2.p, w3.p, master.p, z.p, s);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(logits, z.p, size_t(n) * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}

namespace {
int gate_columns_impl(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                      uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity, uint64_t* count,
                      bool cands, bool plain = false);
}

int care_gate_columns(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                      uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity, uint64_t* count) {
    return gate_columns_impl(ctx, first, n, features, read_ids, positions, orig, cons, capacity, count, false);
}

int care_gate_columns_candidates(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                                 uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity,
                                 uint64_t* count) {
    return gate_columns_impl(ctx, first, n, features, read_ids, positions, orig, cons, capacity, count, true);
}

int care_gate_columns_plain(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                            uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity, uint64_t* count) {
    return gate_columns_impl(ctx, first, n, features, read_ids, positions, orig, cons, capacity, count, false, true);
}

}  // extern "C"

namespace {
int gate_columns_impl(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                      uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity, uint64_t* count,
                      bool cands, bool plain) {
    if (!ctx || !features || !read_ids || !positions || !orig || !cons || !count)
        return fail(CARE_ERR_ARGUMENT, "null argument");
    if (ctx->opt.correction_type != 3) return fail(CARE_ERR_STATE, "gate columns need correction_type 3");
    if (capacity < 1 || capacity > (uint64_t(1) << 31)) return fail(CARE_ERR_ARGUMENT, "capacity out of range");
    *count = 0;
    auto& r = ctx->record;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        if (r.cap < int(capacity)) {
            r.feat.alloc(size_t(capacity) * CARE_MLP_FEATURES);
            r.ids.alloc(size_t(capacity));
            r.pos.alloc(size_t(capacity));
            r.orig.alloc(size_t(capacity));
            r.cons.alloc(size_t(capacity));
            r.cap = int(capacity);
        }
        launch_fill_u32(reinterpret_cast<u32*>(ctx->counters.p + 2), 1, 0, ctx->stream);
        HIP_TRY(hipGetLastError());
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    const int cap = int(std::min<uint64_t>(capacity, uint64_t(r.cap)));
    const int savedCap = r.cap;
    r.cap = cap;
    r.on = true;
    std::vector<u8> status(size_t(std::max(n, 1)));
    const int pitch = (ctx->rs.maxLen + 15) / 16 * 16;
    std::vector<char> seqs(size_t(std::max(n, 1)) * size_t(pitch));
    care_batch_result res{};
    res.anchor_status = status.data();
    res.anchor_sequences = seqs.data();
    res.sequence_pitch = pitch;
    // anchor columns: no candidate correction; candidate columns: candidate
    // correction through the gate, results discarded
    const int cc = ctx->opt.correct_candidates, ctc = ctx->opt.correction_type_cands;
    ctx->opt.correct_candidates = cands ? 1 : 0;
    ctx->opt.correction_type_cands = cands ? 3 : ctc;
    std::vector<u32> candIds;
    std::vector<char> candSeqs;
    if (cands) {
        res.candidate_capacity = std::max(n, 1) * 64;
        candIds.resize(size_t(res.candidate_capacity));
        candSeqs.resize(size_t(res.candidate_capacity) * size_t(pitch));
        res.candidate_read_ids = candIds.data();
        res.candidate_sequences = candSeqs.data();
    }
    r.cands = cands;
    // the plain method's columns: every candidate, refinement before the decision
    const int limit = ctx->opt.candidate_limit;
    if (plain) ctx->opt.candidate_limit = 0;
    r.refine = plain;
    const int rc = correct_batch_impl(ctx, first, n, &res, false);
    ctx->opt.candidate_limit = limit;
    r.refine = false;
    ctx->opt.correct_candidates = cc;
    ctx->opt.correction_type_cands = ctc;
    r.cands = false;
    r.on = false;
    r.cap = savedCap;
    if (rc != CARE_OK) return rc;
    try {
        int got = 0;
        HIP_TRY(hipMemcpy(&got, ctx->counters.p + 2, sizeof(int), hipMemcpyDeviceToHost));
        if (got > cap) return fail(CARE_ERR_CAPACITY, "%d gate columns exceed capacity %d", got, cap);
        if (got > 0) {
            HIP_TRY(hipMemcpy(features, r.feat.p, sizeof(float) * size_t(got) * CARE_MLP_FEATURES, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(read_ids, r.ids.p, sizeof(u32) * size_t(got), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(positions, r.pos.p, sizeof(uint16_t) * size_t(got), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(orig, r.orig.p, size_t(got), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(cons, r.cons.p, size_t(got), hipMemcpyDeviceToHost));
        }
        *count = uint64_t(got);
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}
}  // namespace

extern "C" {

// Shape search: the query kernel over the LDS hit capacities that fit,
// then the correction kernel over workgroup sizes and MSA staging chunks,
// each timed three times (minimum kept) on the same sample of anchors.
int care_autotune(care_context* ctx, uint64_t first, int n, care_tuning* chosen) {
    if (!ctx) return fail(CARE_ERR_ARGUMENT, "null context");
    if (!ctx->hasTables) return fail(CARE_ERR_STATE, "tables not built");
    if (n <= 0 || first + uint64_t(n) > ctx->rs.n) return fail(CARE_ERR_ARGUMENT, "range outside read store");
    if (ctx->opt.paired_end && ((first & 1) || (n & 1)))
        return fail(CARE_ERR_ARGUMENT, "paired-end samples must start at an even id and hold whole pairs");
    const auto& o = ctx->opt;
    CorrectParams P{};
    if (const int rc = prepare_params(ctx, P)) return rc;
    // no gate recording while timing
    P.gateFeat = nullptr;
    P.gateIds = nullptr;
    P.gatePos = nullptr;
    P.gateOrig = nullptr;
    P.gateCons = nullptr;
    P.gateCount = nullptr;
    P.gateCap = 0;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const int B = default_batch(ctx);
        n = std::min(n, std::min(B, 65536)) & ~1;
        if (n == 0) n = 2 > int(ctx->rs.n - first) ? 1 : 2;
        ensure_batch(ctx, B);
        hipStream_t s = ctx->stream;
        const int candLimit = o.candidate_limit;
        auto timed = [&](auto&& launch) {
            float best = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                HIP_TRY(hipEventRecord(ctx->ev[0], s));
                launch();
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipEventRecord(ctx->ev[1], s));
                HIP_TRY(hipEventSynchronize(ctx->ev[1]));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
                best = std::min(best, ms);
            }
            return best;
        };
        auto query = [&](int capL, int regMax) {
            launch_query(ctx->rs, ctx->tables, first, n, ctx->maxCand, capL, regMax, o.exclude_ambiguous, candLimit,
                         ctx->candIds.p, ctx->candCount.p, reinterpret_cast<u32*>(ctx->metaA.p), nullptr, s);
        };
        // query: LDS hit capacities up to the longest possible list and, for
        // tables whose lists can exceed 1024 hits (unless pinned by
        // care_set_tuning), register sorts of up to 1024 (lists beyond are
        // merged) or up to 2048 hits
        const int full = (ctx->maxCand + 15) / 16 * 16;
        std::vector<int> regs{ctx->queryRegMax};
        if (ctx->queryRegMax == 0 && ctx->maxCand > 1024) regs = {1024, 0};
        int bestCap = ctx->queryCap, bestReg = ctx->queryRegMax;
        float bestQ = 1e30f;
        for (int regMax : regs) {
            for (int capL : {128, 256, 512, 1024, 1536, full}) {
                if (capL > full || 4 * query_wave_lds_bytes(capL, candLimit) > size_t(kMaxLdsBytes)) continue;
                const float ms = timed([&] { query(capL, regMax); });
                if (ms < bestQ) {
                    bestQ = ms;
                    bestCap = capL;
                    bestReg = regMax;
                }
            }
        }
        ctx->queryCap = bestCap;
        ctx->queryRegMax = bestReg;
        query(bestCap, bestReg);
        HIP_TRY(hipGetLastError());
        // correction: 2, 3 or 4 wavefronts per anchor, automatic and halved MSA chunks
        P.candIds = ctx->candIds.p;
        P.candCount = ctx->candCount.p;
        P.metaA = ctx->metaA.p;
        P.metaB = ctx->metaB.p;
        P.outStatus = ctx->outStatus.p;
        P.outSeq = ctx->outSeq.p;
        P.outPitch = ctx->outPitch;
        P.candOutIds = ctx->candOutIds.p;
        P.candOutSeq = ctx->candOutSeq.p;
        P.candOutCount = ctx->counters.p + 1;
        P.candOutCap = ctx->candOutCap;
        P.queueHead = ctx->counters.p;
        P.firstAnchor = first;
        P.numAnchors = n;
        const int gatedK = correct_gate_mode(o.correction_type, o.correct_candidates, o.correction_type_cands);
        int bestT = ctx->blockThreads, bestChunk = ctx->msaChunk;
        float bestC = 1e30f;
        for (int T : {256, 192, 128}) {
            const int autoChunk = correct_msa_chunk(ctx->rs.maxLen, o.use_quality, ctx->rs.qpitch, T, gatedK, 0);
            for (int chunk : {0, autoChunk / 2}) {
                if (chunk != 0 && (chunk < 16 || chunk == autoChunk)) continue;
                const size_t lds = correct_lds_bytes(ctx->rs.maxLen, o.use_quality, ctx->rs.qpitch, T, gatedK, chunk);
                if (lds > size_t(kMaxLdsBytes)) continue;
                const int bpc = ctx->blocksPerCU > 0 ? ctx->blocksPerCU : correct_resident_blocks(T, lds, gatedK);
                const int blocks = std::max(1, std::min(n, ctx->numCUs * bpc));
                const float ms = timed([&] {
                    launch_reset_counters(ctx->counters.p, ctx->counters.p + 1, s);
                    launch_correct(P, blocks, T, chunk, s);
                });
                if (ms < bestC) {
                    bestC = ms;
                    bestT = T;
                    bestChunk = chunk;
                }
            }
        }
        ctx->blockThreads = bestT;
        ctx->msaChunk = bestChunk;
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    } catch (const std::bad_alloc&) {
        return fail(CARE_ERR_DEVICE, "out of host memory");
    }
    if (chosen) return care_get_tuning(ctx, chosen);
    return CARE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Correction without allocation: care_correct_batch_noalloc
// ---------------------------------------------------------------------------
// The batch scratch (candidate lists, records, per-anchor results, counters,
// candidate results) of care_correct_batch_noalloc lives in this fixed
// device array.  Each device holds its own copy from the first load of the
// library until the process ends; host code reaches it only through the
// address hipGetSymbolAddress reports for the device.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void* g_scratchAddr[kMaxDevices] = {};

// the device address of `scratch` on the context's device (looked up once
// per device, with that device current)
int scratch_address(care_context* ctx, void** out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(CARE_ERR_DEVICE, "device index %d", ctx->device);
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    if (!g_scratchAddr[ctx->device]) {
        if (hipSetDevice(ctx->device) != hipSuccess) return fail(CARE_ERR_DEVICE, "hipSetDevice failed");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CARE_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratchAddr[ctx->device] = p;
    }
    *out = g_scratchAddr[ctx->device];
    return CARE_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// scratch bytes of a device batch of b anchors
size_t noalloc_bytes(const care_context* ctx, int b) {
    const size_t mc = size_t(ctx->maxCand);
    const size_t outPitch = size_t(ctx->rs.maxLen + 15) / 16 * 16;
    const size_t candCap = ctx->opt.correct_candidates ? size_t(b) * 64 : 0;
    return align256(sizeof(int) * 4) + align256(sizeof(u32) * size_t(b) * mc) + align256(sizeof(int) * size_t(b)) +
           2 * align256(sizeof(CandMeta) * size_t(b) * mc) + align256(size_t(b)) + align256(size_t(b) * outPitch) +
           align256(sizeof(u32) * std::max<size_t>(candCap, 1)) + align256(std::max<size_t>(candCap, 1) * outPitch);
}

}  // namespace

extern "C" int care_correct_batch_noalloc(care_context* ctx, uint64_t first, int n, care_batch_result* r) {
    if (!ctx || !r || !r->anchor_status || !r->anchor_sequences) return fail(CARE_ERR_ARGUMENT, "null argument");
    if (!ctx->hasTables) return fail(CARE_ERR_STATE, "tables not built");
    if (n < 0 || first + uint64_t(n) > ctx->rs.n) return fail(CARE_ERR_ARGUMENT, "range outside read store");
    if (r->sequence_pitch < ctx->rs.maxLen) return fail(CARE_ERR_ARGUMENT, "sequence_pitch < max read length");
    if (ctx->opt.paired_end && ((first & 1) || (n & 1)))
        return fail(CARE_ERR_ARGUMENT, "paired-end batches must start at an even id and hold whole pairs");
    if (ctx->opt.correct_candidates && (!r->candidate_read_ids || !r->candidate_sequences))
        return fail(CARE_ERR_ARGUMENT, "candidate output buffers required with correct_candidates");
    if (ctx->opt.correction_type == 3 && ctx->record.on)
        return fail(CARE_ERR_STATE, "gate column recording is not available without allocation");
    r->num_candidates = 0;
    if (n == 0) return CARE_OK;
    // the largest device batch whose scratch fits the fixed array
    const int step = ctx->opt.paired_end ? 2 : 1;
    int B = 0;
    for (int lo = 0, hi = std::min(n + (n & 1), 1 << 20); lo <= hi;) {
        const int mid = (lo + hi) / 2;
        if (noalloc_bytes(ctx, std::max(mid, 1)) <= kNoAllocScratchBytes) {
            B = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    B -= B % step;
    if (B < step)
        return fail(CARE_ERR_CAPACITY, "%s anchor needs %zu bytes of scratch, more than the fixed %zu",
                    step == 2 ? "a pair of" : "one", noalloc_bytes(ctx, step), kNoAllocScratchBytes);
    const auto& o = ctx->opt;
    CorrectParams P{};
    if (const int rc = prepare_params(ctx, P)) return rc;
    void* base = nullptr;
    if (const int rc = scratch_address(ctx, &base)) return rc;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const int T = ctx->blockThreads;
        const int gatedK = correct_gate_mode(o.correction_type, o.correct_candidates, o.correction_type_cands);
        const size_t lds = correct_lds_bytes(ctx->rs.maxLen, o.use_quality, ctx->rs.qpitch, T, gatedK, ctx->msaChunk);
        if (lds > size_t(kMaxLdsBytes))
            return fail(CARE_ERR_CAPACITY, "reads of length %d need %zu bytes of LDS per correction workgroup (limit %d)",
                        ctx->rs.maxLen, lds, kMaxLdsBytes);
        const int blocksPerCU = ctx->blocksPerCU > 0 ? ctx->blocksPerCU : correct_resident_blocks(T, lds, gatedK);
        // carve the scratch for batches of B anchors
        const size_t mc = size_t(ctx->maxCand);
        const int outPitch = (ctx->rs.maxLen + 15) / 16 * 16;
        const int candCapB = o.correct_candidates ? B * 64 : 0;
        char* p = static_cast<char*>(base);
        auto take = [&](size_t bytes) {
            char* q = p;
            p += align256(bytes);
            return q;
        };
        int* counters = reinterpret_cast<int*>(take(sizeof(int) * 4));
        u32* candIds = reinterpret_cast<u32*>(take(sizeof(u32) * size_t(B) * mc));
        int* candCount = reinterpret_cast<int*>(take(sizeof(int) * size_t(B)));
        CandMeta* metaA = reinterpret_cast<CandMeta*>(take(sizeof(CandMeta) * size_t(B) * mc));
        CandMeta* metaB = reinterpret_cast<CandMeta*>(take(sizeof(CandMeta) * size_t(B) * mc));
        u8* outStatus = reinterpret_cast<u8*>(take(size_t(B)));
        char* outSeq = take(size_t(B) * size_t(outPitch));
        u32* candOutIds = reinterpret_cast<u32*>(take(sizeof(u32) * std::max<size_t>(size_t(candCapB), 1)));
        char* candOutSeq = take(std::max<size_t>(size_t(candCapB), 1) * size_t(outPitch));
        hipStream_t s = ctx->stream;
        P.candIds = candIds;
        P.candCount = candCount;
        P.metaA = metaA;
        P.metaB = metaB;
        P.outPitch = outPitch;
        P.candOutIds = candOutIds;
        P.candOutSeq = candOutSeq;
        P.candOutCount = counters + 1;
        P.queueHead = counters;
        P.gateCount = nullptr;
        int done_cands = 0;
        for (int done = 0; done < n; done += B) {
            const int nb = std::min(B, n - done);
            launch_query(ctx->rs, ctx->tables, first + done, nb, ctx->maxCand, ctx->queryCap, ctx->queryRegMax,
                         o.exclude_ambiguous, o.candidate_limit, candIds, candCount, reinterpret_cast<u32*>(metaA),
                         nullptr, s);
            // queue head and candidate result count start at zero in every batch
            launch_reset_counters(counters, counters + 1, s);
            P.firstAnchor = first + done;
            P.numAnchors = nb;
            P.outStatus = outStatus;
            P.outSeq = outSeq;
            P.candOutCap = o.correct_candidates ? std::min(candCapB, std::max(0, r->candidate_capacity - done_cands)) : 0;
            const int blocks = std::max(1, std::min(nb, ctx->numCUs * blocksPerCU));
            launch_correct(P, blocks, T, ctx->msaChunk, s);
            HIP_TRY(hipGetLastError());
            int nc = 0;
            if (o.correct_candidates) HIP_TRY(hipMemcpyAsync(&nc, counters + 1, sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(r->anchor_status + done, outStatus, size_t(nb), hipMemcpyDeviceToHost, s));
            copy_rows_to_host(r->anchor_sequences + size_t(done) * size_t(r->sequence_pitch), size_t(r->sequence_pitch),
                              outSeq, size_t(outPitch), size_t(ctx->rs.maxLen), size_t(nb), s);
            HIP_TRY(hipStreamSynchronize(s));
            if (o.correct_candidates) {
                if (nc > P.candOutCap) return fail(CARE_ERR_CAPACITY, "candidate_capacity exceeded (%d results)", done_cands + nc);
                if (nc > 0) {
                    HIP_TRY(hipMemcpyAsync(r->candidate_read_ids + done_cands, candOutIds, sizeof(u32) * size_t(nc),
                                           hipMemcpyDeviceToHost, s));
                    copy_rows_to_host(r->candidate_sequences + size_t(done_cands) * size_t(r->sequence_pitch),
                                      size_t(r->sequence_pitch), candOutSeq, size_t(outPitch), size_t(ctx->rs.maxLen),
                                      size_t(nc), s);
                    HIP_TRY(hipStreamSynchronize(s));
                }
                done_cands += nc;
                r->num_candidates = done_cands;
            }
        }
    } catch (const HipError& e) {
        return fail(CARE_ERR_DEVICE, "%s: %s", e.what, hipGetErrorString(e.err));
    }
    return CARE_OK;
}
