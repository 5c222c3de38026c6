/*
 * care_amd.h - C-ABI of the MI355X correction engine for CARE.
 *
 * The engine replaces the device side of `care-gpu ... -g 0`: read storage in
 * HBM, the minhasher (signatures, tables, candidate queries) and the
 * correction pipeline (alignment, MSA, refinement, anchor and candidate
 * correction).  Every entry point takes plain pointers and sizes, returns
 * CARE_OK (0) on success or a CARE_ERR_* code, never throws, and records a
 * message retrievable with care_last_error().  Host buffers passed in are
 * only read during the call; host output buffers are filled before the call
 * returns.  A context is bound to one device and is not safe for concurrent
 * use from several threads; distinct contexts are independent.
 */
#ifndef CARE_AMD_H
#define CARE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CARE_OK 0
#define CARE_ERR_ARGUMENT 1  /* invalid argument or size out of range */
#define CARE_ERR_DEVICE 2    /* HIP runtime error (allocation, launch, copy) */
#define CARE_ERR_STATE 3     /* call made in the wrong order */
#define CARE_ERR_CAPACITY 4  /* an output buffer was too small */

/* Correction parameters.  Field meaning and defaults follow the reference's
 * ProgramOptions (include/options.hpp:36-98); care_options_default() fills
 * those defaults. */
typedef struct care_options {
    int kmer_length;             /* -k, 1..31 */
    int num_hash_maps;           /* -h, 1..64 */
    float coverage;              /* -c, estimated coverage */
    int paired_end;              /* --pairmode PE: mates are reads 2i, 2i+1 */
    int use_quality;             /* -q */
    int exclude_ambiguous;       /* --excludeAmbiguous */
    int correct_candidates;      /* --candidateCorrection */
    int new_columns_to_correct;  /* --candidateCorrectionNewColumns */
    int min_overlap;             /* --minalignmentoverlap */
    float max_mismatch_ratio;    /* --maxmismatchratio */
    float min_overlap_ratio;     /* --minalignmentoverlapratio */
    float error_factor;          /* --errorfactortuning (estimatedErrorrate) */
    float coverage_factor;       /* --coveragefactortuning (m_coverage) */
    float paired_filter_threshold; /* --pairedFilterThreshold */
    int correction_type;         /* --correctionType: 0 classic, 1 forest, 3 learned gate */
    int correction_type_cands;   /* --correctionTypeCands: 0 classic, 1 forest, 3 learned gate
                                    (with correction_type 3) */
    float threshold_anchor;      /* --thresholdAnchor */
    float threshold_cands;       /* --thresholdCands */
    int candidate_limit;         /* keep the candidates hit by the most maps, at most this many
                                    per anchor, ties to the smaller read id (0 = all, the
                                    reference's behaviour) */
    int max_refinements;         /* rounds of different-region refinement, 0..5 (5 = the
                                    reference, include/corrector.hpp:1309; correction_type 3
                                    always skips refinement) */
} care_options;

typedef struct care_context care_context;

void care_options_default(care_options* opt);

/* Thread-local text of the last error reported by this library. */
const char* care_last_error(void);

/* Creates a context on HIP device `device`.
 * Replaces the per-device setup of performCorrection
 * (src/gpu/dispatch_care_correct_gpu.cu:551-602). */
int care_context_create(int device, const care_options* opt, care_context** out);
int care_context_destroy(care_context* ctx);

/* Uploads the packed read store to HBM.
 *   seqs      : num_reads x seq_pitch u32, 2-bit A0 C1 G2 T3, 16 bases per
 *               word, first base in bits 31..30 (include/sequencehelpers.hpp:154-187)
 *   lengths   : num_reads read lengths (1..max_length)
 *   ambiguous : num_reads flags, 1 if the input read held a non-ACGT base
 *   quals     : num_reads x qual_pitch u32 binned qualities (1, 2 or 8 bits
 *               per base, include/qualityscorecompression.hpp:28-146), or NULL
 * Replaces MultiGpuReadStorage construction / rebuild
 * (include/gpu/multigpureadstorage.cuh, src/gpu/dispatch_care_correct_gpu.cu:773-806). */
int care_readstore_upload(care_context* ctx, uint64_t num_reads, int max_length, const uint32_t* seqs,
                          int seq_pitch, const int32_t* lengths, const uint8_t* ambiguous, const uint32_t* quals,
                          int qual_pitch, int qual_bits);

/* Gathers rows of the device read store into host buffers (sequences with
 * out_pitch words per row, lengths).  Replaces GpuReadStorage::gatherSequences /
 * gatherSequenceLengths (include/gpu/gpureadstorage.cuh:22-126). */
int care_readstore_gather(care_context* ctx, const uint32_t* ids, int n, uint32_t* out_seqs, int out_pitch,
                          int32_t* out_lengths);

/* ---- Stream-ordered interface on device buffers -------------------------
 * The calls below take device pointers and a HIP stream (hipStream_t is
 * `struct ihipStream_t*`, hip/hip_runtime_api.h; NULL is the null stream),
 * enqueue their kernels on it and return at once: outputs are valid when
 * the stream reaches that point.  They allocate nothing and never
 * synchronise the host.  Ids outside the read store give zero rows. */
struct ihipStream_t;

/* Rows of read ids d_ids[0..n) into d_out (out_pitch words per row; words
 * past the store's pitch are zero).  Replaces GpuReadStorage::gatherSequences
 * (include/gpu/gpureadstorage.cuh:22-126; MultiGpuReadStorage
 * include/gpu/multigpureadstorage.cuh:669). */
int care_readstore_gather_sequences_async(care_context* ctx, const uint32_t* d_ids, int n, uint32_t* d_out,
                                          int out_pitch, struct ihipStream_t* stream);
/* Lengths of read ids d_ids[0..n).  Replaces
 * GpuReadStorage::gatherSequenceLengths (include/gpu/multigpureadstorage.cuh:1341). */
int care_readstore_gather_lengths_async(care_context* ctx, const uint32_t* d_ids, int n, int32_t* d_out,
                                        struct ihipStream_t* stream);
/* Quality characters of read ids d_ids[0..n) (out_pitch chars per row;
 * binned qualities decode to their bins' characters).  Needs use_quality.
 * Replaces GpuReadStorage::gatherQualities (include/gpu/multigpureadstorage.cuh:1040). */
int care_readstore_gather_qualities_async(care_context* ctx, const uint32_t* d_ids, int n, char* d_out,
                                          int out_pitch, struct ihipStream_t* stream);
/* 1 for reads whose input held a non-ACGT base.  Replaces
 * GpuReadStorage::areSequencesAmbiguous (include/gpu/multigpureadstorage.cuh:637). */
int care_readstore_ambiguous_async(care_context* ctx, const uint32_t* d_ids, int n, uint8_t* d_out,
                                   struct ihipStream_t* stream);

/* Stride of the candidate lists written by care_minhash_query_async:
 * num_hash_maps x the per-map value limit (include/minhasherlimit.hpp:10-15). */
int care_minhash_max_candidates(care_context* ctx);
/* Device workspace bytes care_minhash_query_async needs for n sequences. */
size_t care_minhash_query_workspace_bytes(care_context* ctx, int n);
/* Candidates of n sequences given on the device (2-bit rows of seq_pitch
 * words, lengths up to 16 * seq_pitch): their minhash signatures, the table
 * hits of every map, sorted and de-duplicated, without the read d_read_ids[i]
 * itself and, under exclude_ambiguous, without ambiguous reads (an
 * ambiguous d_read_ids[i] gets no candidates; d_read_ids may be NULL).
 * Sequence i's candidates go to d_candidates[i * care_minhash_max_candidates
 * ..], their number to d_counts[i].  d_workspace: 256-byte aligned,
 * care_minhash_query_workspace_bytes(ctx, n) bytes, not used by another call
 * in flight.  Replaces GpuMinhasher::determineNumValues + retrieveValues
 * (include/gpu/gpuminhasher.cuh:20-110) followed by
 * GpuMinhashQueryFilter::keepDistinctAndNotMatching
 * (include/gpu/minhashqueryfilter.cuh:21-100). */
int care_minhash_query_async(care_context* ctx, const uint32_t* d_seqs, int seq_pitch, const int32_t* d_lengths,
                             const uint32_t* d_read_ids, int n, uint32_t* d_candidates, int32_t* d_counts,
                             void* d_workspace, size_t workspace_bytes, struct ihipStream_t* stream);

/* Device memory and streams of the context's device, for callers without a
 * HIP binding of their own (no counterpart in the reference, whose callers
 * hold rmm resources and CUDA streams).  to_device: 1 host to device, 0
 * device to host; copies are synchronous. */
int care_device_alloc(care_context* ctx, size_t bytes, void** out);
int care_device_free(care_context* ctx, void* p);
int care_device_copy(care_context* ctx, void* dst, const void* src, size_t bytes, int to_device);
int care_stream_create(care_context* ctx, struct ihipStream_t** out);
int care_stream_destroy(care_context* ctx, struct ihipStream_t* stream);
int care_stream_synchronize(care_context* ctx, struct ihipStream_t* stream);

/* Builds all minhash tables on the device from the uploaded read store.
 * Replaces constructGpuMinhasherFromGpuReadStorage / GpuMinhasher::insert +
 * compact (src/gpu/gpuminhasherconstruction.cu:37-253,
 * include/gpu/gpuhashtable.cuh:743). */
int care_minhash_build(care_context* ctx);

/* One minhash map in key-grouped form: num_keys distinct keys (never
 * UINT64_MAX), key i owning values[offsets[i] .. offsets[i] + counts[i]),
 * the read ids that share that signature value.  This is the content of
 * one CpuReadOnlyMultiValueHashTable (include/cpuhashtable.hpp:462-660):
 * its values array and, per key, the ValueIndex (offset, count) of its
 * lookup table. */
typedef struct care_map_view {
    uint64_t num_keys;
    const uint64_t* keys;
    const uint32_t* offsets;
    const uint32_t* counts;
    uint64_t num_values;
    const uint32_t* values;
} care_map_view;

/* Installs num_maps prebuilt maps over the uploaded read store, in place of
 * care_minhash_build (a hash-table file loaded by the host).  Counts must
 * lie in 1..results_per_map (<= 65535), values below the number of reads,
 * keys distinct within a map.  kmer_length and num_maps replace the
 * context's k and number of maps; the signatures the queries use are
 * computed on the device with them.
 * Replaces OrdinaryCpuMinhasher/FakeGpuMinhasher::loadFromStream
 * (include/ordinaryminhasher.hpp:318-339, include/gpu/fakegpuminhasher.cuh:505-525)
 * after the host has parsed the file. */
int care_minhash_import(care_context* ctx, const care_map_view* maps, int num_maps, int kmer_length,
                        int results_per_map);

/* k, number of maps and the per-key value limit (resultsPerMapThreshold)
 * of the context's tables.  Replaces GpuMinhasher::getKmerSize /
 * getNumberOfMaps / getNumResultsPerMapThreshold (include/gpu/gpuminhasher.cuh:90-94). */
int care_minhash_geometry(care_context* ctx, int* kmer_length, int* num_maps, int* results_per_map);

/* Size of map `map` in key-grouped form. */
int care_minhash_map_size(care_context* ctx, int map, uint64_t* num_keys, uint64_t* num_values);
/* Copies map `map` in key-grouped form, keys ascending and each key's
 * values (ascending read ids) contiguous in key order, into caller arrays
 * of the sizes care_minhash_map_size gives.  With care_minhash_map_size it
 * replaces the table walk of writeToStream
 * (include/ordinaryminhasher.hpp:303-316, include/cpuhashtable.hpp:221-232,627-636). */
int care_minhash_export_map(care_context* ctx, int map, uint64_t* keys, uint32_t* offsets, uint32_t* counts,
                            uint32_t* values);

/* Minhash signatures of reads [first, first+n): n x num_hash_maps keys.
 * Replaces minhashSignatures3264Kernel (include/gpu/gpusequencehasher.cuh:116). */
int care_minhash_signatures(care_context* ctx, uint64_t first, int n, uint64_t* out);

/* Candidate ids of anchors [first, first+n): for every anchor the sorted,
 * de-duplicated table hits without the anchor itself (and without ambiguous
 * reads under exclude_ambiguous).  offsets has n+1 entries.  Returns
 * CARE_ERR_CAPACITY if more than ids_capacity ids result.
 * Replaces determineNumValues + retrieveValues + keepDistinctAndNotMatching
 * (include/gpu/gpuminhasher.cuh:20-110, include/gpu/minhashqueryfilter.cuh:21-100). */
int care_minhash_query(care_context* ctx, uint64_t first, int n, uint32_t* ids, int32_t* offsets,
                       uint64_t ids_capacity);

/* Results of one correction batch, filled by care_correct_batch. */
typedef struct care_batch_result {
    /* per anchor: bit0 = corrected, bit1 = high-quality MSA */
    uint8_t* anchor_status;      /* [num_anchors] */
    char* anchor_sequences;      /* [num_anchors * sequence_pitch] corrected bases (first max_length
                                    bytes of a row; rows of uncorrected anchors are unspecified) */
    int sequence_pitch;          /* bytes per row, >= max_length */
    /* candidate corrections (only with correct_candidates) */
    uint32_t* candidate_read_ids;  /* [candidate_capacity] */
    char* candidate_sequences;     /* [candidate_capacity * sequence_pitch], forward strand */
    int candidate_capacity;
    int num_candidates;            /* out */
} care_batch_result;

/* Corrects anchors [first, first+n) (n even for paired-end, first even).
 * Result arrays may be any host memory; into page-locked ones
 * (care_host_buffer_alloc) each internal batch's anchor results are copied
 * while later batches still run.  Returns when all results are in place.
 * Replaces GpuErrorCorrector::correct (include/gpu/gpucorrector.cuh:1060)
 * with its kernels: alignment, filters, MSA build + refinement, anchor and
 * candidate correction (src/gpu/{alignment,msa,anchorcorrection,
 * candidatecorrection,correction}kernels.cu). */
int care_correct_batch(care_context* ctx, uint64_t first, int n, care_batch_result* result);

/* care_correct_batch through a HIP graph: the device work of the call (every
 * internal batch's query, count, queue reset and correction kernels, the
 * anchor result copies and the counter copies) is captured on the first
 * call and replayed by later calls with the same range, result arrays,
 * options, launch shapes and device buffers; any change captures anew.
 * anchor_status and anchor_sequences must be page-locked
 * (care_host_buffer_alloc).  Results equal care_correct_batch's; the batch
 * statistics give the total device time only.  Replaces the stage-by-stage
 * launches of the reference's correction pipeline with their host
 * synchronisations between stages (include/gpu/gpucorrector.cuh:336,433,3105). */
int care_correct_batch_graph(care_context* ctx, uint64_t first, int n, care_batch_result* result);

/* care_correct_batch without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as care_correct_batch.  Its batch scratch is a
 * fixed 256 MiB device array of the library (one per device, held until the
 * process ends): anchors go through in device batches that fit it (about
 * 1,460 anchors at -h 48 -c 40, 3,000 at -h 32 -c 30), each batch's results
 * copied to the result arrays before the next starts; CARE_ERR_CAPACITY if
 * one anchor (one pair) needs more scratch than the array holds.  Calls
 * that overlap on one device (other threads or contexts) race on the array.
 * Gate column recording (care_gate_columns) is not available here. */
int care_correct_batch_noalloc(care_context* ctx, uint64_t first, int n, care_batch_result* result);

/* Number of anchors per internal launch (even, 0 = automatic) and threads
 * per correction workgroup (128, 192 or 256, 0 = keep); the other launch shapes
 * stay as set by care_set_tuning / care_autotune.  The reference fixes these
 * at compile time (e.g. blocksize 128 in src/gpu/alignmentkernels.cu:652,
 * src/gpu/msakernels.cu:1186). */
int care_set_launch_config(care_context* ctx, int batch_anchors, int block_threads);

/* Launch shapes of the engine's kernels.  None of them changes a result:
 * every shape gives bit-identical candidate lists and corrections.  No
 * counterpart in the reference, which fixes its launch shapes at compile
 * time (src/gpu/alignmentkernels.cu:652, src/gpu/msakernels.cu:1186). */
typedef struct care_tuning {
    int batch_anchors;          /* anchors per internal launch, even (0 = automatic); a call of
                                   more than one launch ends in one of batch_anchors / 8 */
    int correct_threads;        /* threads per correction workgroup, i.e. 2, 3 or 4
                                   wavefronts per anchor: 128, 192 or 256 (0 = 256) */
    int correct_blocks_per_cu;  /* persistent correction workgroups per compute unit
                                   (0 = as many as are resident at once) */
    int msa_chunk;              /* candidates staged in LDS per MSA pass, 16..256
                                   (0 = automatic: about 16 KB of rows) */
    int query_lds_hits;         /* longest hit list a query wavefront sorts in LDS,
                                   multiple of 16 in 16..4096; longer lists use global
                                   scratch (0 = 512) */
    int query_register_hits;    /* longest hit list sorted in registers (bitonic network
                                   across the wavefront), power of two in 64..2048;
                                   longer lists are merged list by list (0 = 1024, or
                                   2048 when num_hash_maps * bucket cap exceeds 1024);
                                   up to 1024 runs the 16-per-lane query kernel (fewer
                                   registers), above it the 32-per-lane one */
} care_tuning;

int care_set_tuning(care_context* ctx, const care_tuning* tuning);
/* Current shapes with automatic values resolved. */
int care_get_tuning(care_context* ctx, care_tuning* out);

/* Times candidate shapes of the query and correction kernels on anchors
 * [first, first+n) of the built tables (n even for paired-end, first even;
 * at most 65536 anchors are used), keeps the fastest as the context's
 * tuning and writes it to `chosen` (may be NULL).  Costs a few query and
 * correction launches over the sample; results of later calls do not
 * depend on the outcome. */
int care_autotune(care_context* ctx, uint64_t first, int n, care_tuning* chosen);

/* Forest classifier for correction_type / correction_type_cands 1, in the
 * reference's .rf file format (include/forest.hpp:38-55, ml/care.py:198-282).
 * which: 0 anchor, 1 candidates.  Replaces GpuForest
 * (include/gpu/forest_gpu.cuh:133). */
int care_forest_load(care_context* ctx, int which, const char* path, uint32_t max_trees);

/* ---- Learned correction gate (correction_type 3) ----------------------
 * No counterpart in the reference: a 32-64-32-16-1 MLP (ReLU, sigmoid
 * output) over MSA column features replaces the forest decision of
 * correctAnchorClf (include/corrector.hpp:1448-1472).  In correction_type 3
 * the engine also keeps only the candidate_limit candidates hit by the most
 * maps and skips MSA refinement; every anchor column whose consensus
 * differs from the anchor base in a non-HQ MSA is corrected only if the
 * gate's output reaches threshold_anchor.  Parameters are one FP32 vector of
 * CARE_MLP_PARAMS values: W1[64][32], b1[64], W2[32][64], b2[32],
 * W3[16][32], b3[16], w4[16], b4 (matrices output-major).  The device
 * evaluates it with BF16 inputs and weights and FP32 accumulation. */
#define CARE_MLP_FEATURES 32
#define CARE_MLP_PARAMS 4737

typedef struct care_mlp_train_options {
    int epochs;      /* passes over the data */
    int batch;       /* rows per Adam step, 16..65536 */
    float lr, beta1, beta2, eps;  /* Adam */
    uint64_t seed;   /* shuffling */
} care_mlp_train_options;

/* Sets the gate used by care_correct_batch under correction_type 3. */
int care_mlp_load(care_context* ctx, const float* params);
/* Parameters of the gate on candidate columns (correction_type_cands 3);
 * without them the anchor gate's parameters are used. */
int care_mlp_load_candidates(care_context* ctx, const float* params);

/* Gradient of the mean binary cross-entropy over n rows (features [n][32],
 * labels 0/1) at params, computed by the device's training kernels; loss_out
 * may be NULL. */
int care_mlp_gradient(care_context* ctx, const float* features, const uint8_t* labels, int n, const float* params,
                      float* grad_out, float* loss_out);

/* Trains on the device with Adam from the parameters in params (overwritten
 * with the result); epoch_loss (may be NULL) receives the mean batch loss of
 * every epoch. */
int care_mlp_train(care_context* ctx, const float* features, const uint8_t* labels, uint64_t n,
                   const care_mlp_train_options* opt, float* params, float* epoch_loss);

/* Gate outputs before the sigmoid (logits) for n feature rows. */
int care_mlp_infer(care_context* ctx, const float* features, uint64_t n, const float* params, float* logits);

/* Runs the correction_type 3 pipeline on anchors [first, first+n) and
 * returns every gate decision column: its 32 features, anchor read id,
 * position, original and consensus base (0..3).  Works without loaded gate
 * weights (then no column is corrected).  *count receives the number of
 * columns; CARE_ERR_CAPACITY if they exceed capacity. */
int care_gate_columns(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                      uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity, uint64_t* count);

/* Candidate counterpart of care_gate_columns: the decision columns of the
 * candidate corrections (candidates of HQ anchors within new_columns of the
 * anchor, positions where the consensus differs from the candidate) with
 * their 32 features, read id, forward-strand position and forward-strand
 * original and consensus bases.  Runs correction_type_cands 3 over the
 * batch; needs correction_type 3. */
int care_gate_columns_candidates(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                                 uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity,
                                 uint64_t* count);

/* care_gate_columns over the MSAs of the plain method: every candidate (the
 * context's candidate_limit is not applied) and max_refinements rounds of
 * refinement before the decision columns are taken, as the plain path builds
 * them; the training columns of a gate learned from the plain method.  Needs
 * correction_type 3; no counterpart in the reference. */
int care_gate_columns_plain(care_context* ctx, uint64_t first, int n, float* features, uint32_t* read_ids,
                            uint16_t* positions, uint8_t* orig, uint8_t* cons, uint64_t capacity, uint64_t* count);

/* Timing and work counters of the last care_correct_batch call. */
typedef struct care_batch_stats {
    double query_ms;       /* minhash query kernel */
    double correct_ms;     /* fused correction kernel */
    double total_ms;       /* whole call incl. copies */
    uint64_t raw_hits;     /* table values gathered before de-duplication */
    uint64_t candidates;   /* candidates after query filtering */
    uint32_t launches;     /* correction kernel launches (device batches) */
} care_batch_stats;
int care_last_batch_stats(care_context* ctx, care_batch_stats* stats);

/* Page-locked host memory for result buffers (care_batch_result arrays):
 * results are copied into it by DMA without staging through a bounce
 * buffer.  Any host memory works as a result buffer; this only makes the
 * copies faster.  The reference's counterpart is its pinned staging of
 * results, helpers::SimpleAllocationPinnedHost (include/gpu/gpucorrector.cuh:115-126). */
int care_host_buffer_alloc(size_t bytes, void** out);
int care_host_buffer_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* CARE_AMD_H */
