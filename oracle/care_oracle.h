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
#ifdef __cplusplus
}
#endif
#endif
