// C interface of the CPU restatement of CARE's correction path.
// Test infrastructure only: used by tests/ and by the benchmark's CPU
// baseline, never by the product.
#ifndef CARE_ORACLE_H
#define CARE_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int orc_main(int argc, char** argv);
void orc_encode_2bit(const char* seq, int len, uint32_t* out);
void orc_decode_2bit(const uint32_t* enc, int len, char* out);
void orc_revcomp_2bit(const uint32_t* enc, int len, uint32_t* out);
void orc_to_hilo(const uint32_t* enc, int len, uint32_t* out);
void orc_encode_quality(const char* q, int len, int bits, uint32_t* out);
void orc_decode_quality(const uint32_t* enc, int len, int bits, char* out);
uint64_t orc_murmur64(uint64_t x);
void orc_signature(const uint32_t* enc, int len, int k, int h, uint64_t* out);
float orc_quality_weight(int c);
int orc_align(const uint32_t* a, int la, const uint32_t* c, int lc, int minOverlap, float maxErrorRate,
              float minOverlapRatio, int* out4);
int orc_choose_orientation(const int* f4, int fvalid, const int* r4, int rvalid, int la, float minOverlapRatio,
                           int minOverlap, float estimatedErrorrate);
void* orc_tables_build(const uint32_t* seqs, int pitch, const int* lens, size_t n, int k, int h, float coverage);
void orc_tables_free(void* t);
int orc_tables_query(void* t, const uint32_t* enc, int len, uint32_t self, uint32_t* out, int cap);
int orc_tables_bucket(void* t, int map, uint64_t key, uint32_t* out, int cap);
int orc_tables_query_limited(void* t, const uint32_t* enc, int len, uint32_t self, int limit, uint32_t* out, int cap);
int orc_msa_build(int useQ, int la, const char* anchor, const char* anchorQ, int n, const char* cands, int candPitch,
                  const char* candQ, int qualPitch, const int* lens, const int* shifts, const float* weights, int cap,
                  int* dims, char* consensus, float* support, int* coverage, int* counts, float* wts, int* origCov);
int orc_msa_properties(const float* support, const int* coverage, int ncols, int first, int last, float err, float cov,
                       float mcov, float* out4);
int orc_msa_corrected_anchor(int useQ, int la, const char* anchor, const char* anchorQ, int n, const char* cands,
                             int candPitch, const char* candQ, int qualPitch, const int* lens, const int* shifts,
                             const float* weights, float err, float cov, float mcov, char* out);
int orc_msa_different_region(int la, const char* anchor, int n, const char* cands, int candPitch, const int* lens,
                             const int* shifts, const float* weights, int datasetCoverage, char* remove);
float orc_se_filter_threshold(const float* ratios, int n, float err, float cov, float mcov);
int orc_candidate_in_range(int qb, int qe, int ab, int ae, int ncc, int forestOrGate);
float orc_overlap_weight(int nOps, int overlap, float maxMismatchRatio);
int orc_make_edits(const char* corrected, const char* original, int len, int ambiguous, int* positions, char* bases);
int orc_forest_decide(const char* path, int kind, uint32_t maxTrees, float threshold, const float* features);
int orc_combine(const char* original, int len, int ambiguous, int n, const int* flags, const char* corrected, char* out);
#ifdef __cplusplus
}
#endif
#endif
