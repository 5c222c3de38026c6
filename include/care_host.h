/*
 * care_host.h - host side of the MI355X CARE corrector (C-ABI).
 *
 * care_host_main is the `care-gpu` command line: the same options, inputs
 * and ordered FASTQ/FASTA output as the reference's care-gpu
 * (src/gpu/main_correct_gpu.cu:82-203, src/gpu/dispatch_care_correct_gpu.cu:522-929),
 * with the device work done by the engine declared in care_amd.h.
 *
 * care_readset_* expose the host read store (STEP1 of the reference,
 * include/chunkedreadstorageconstruction.hpp:28-460) so that other hosts
 * (Python bindings, benchmarks) can feed the engine.
 */
#ifndef CARE_HOST_H
#define CARE_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the corrector; returns the process exit status. */
int care_host_main(int argc, char** argv);

typedef struct care_readset care_readset;

/* Parses FASTA/FASTQ (optionally gzip'ed) files into a packed read store:
 * 2-bit sequences with non-ACGT bases replaced cyclically by A,C,G,T (the
 * counter restarts every 65536 reads of a file), ambiguity flags and binned
 * qualities (qual_bits 1, 2 or 8; 0 = no qualities).  paired_end with two
 * files interleaves mates.  Returns NULL on error (see care_host_last_error). */
care_readset* care_readset_load(const char* const* files, int num_files, int paired_end, int qual_bits);
void care_readset_free(care_readset* rs);
const char* care_host_last_error(void);

uint64_t care_readset_size(const care_readset* rs);
int care_readset_max_length(const care_readset* rs);
int care_readset_seq_pitch(const care_readset* rs);
int care_readset_qual_pitch(const care_readset* rs);
const uint32_t* care_readset_seqs(const care_readset* rs);
const int32_t* care_readset_lengths(const care_readset* rs);
const uint8_t* care_readset_ambiguous(const care_readset* rs);
const uint32_t* care_readset_quals(const care_readset* rs);

/* ---- Reference file formats ----------------------------------------------
 * Files written here load in the reference's care-gpu and the other way
 * round (same fields, packed in host byte order).  Functions returning int
 * give 0 on success, 1 on error (see care_host_last_error). */

/* Writes the read set as a preprocessed-reads file.  qual_bits is the
 * bits-per-quality field written when the read set holds no qualities.
 * Replaces ChunkedReadStorage::saveToFile (include/chunkedreadstorage.hpp:246-400). */
int care_readset_save(const care_readset* rs, const char* path, int qual_bits);

/* Reads a preprocessed-reads file.  With use_quality the file must hold
 * qualities of qual_bits bits; without, its qualities are skipped.  Returns
 * NULL on error.  Replaces ChunkedReadStorage::loadFromFile
 * (include/chunkedreadstorage.hpp:158-243). */
care_readset* care_readset_load_preprocessed(const char* path, int use_quality, int qual_bits);

struct care_context;

/* Writes the context's minhash tables as a hash-table file: k, the per-key
 * value limit (results_per_map; 0 = the tables' own), the load factor of
 * the written lookups (0 < load_factor <= 1) and every map.  Replaces
 * FakeGpuMinhasher/OrdinaryCpuMinhasher::writeToStream
 * (include/gpu/fakegpuminhasher.cuh:491-503, include/ordinaryminhasher.hpp:303-316). */
int care_minhash_save_file(struct care_context* ctx, const char* path, int results_per_map, float load_factor);

/* Loads the first min(max_maps, maps in the file) tables of a hash-table
 * file into the context (after care_readstore_upload, in place of
 * care_minhash_build); the file's k and map count then replace the
 * context's.  *loaded_maps (may be NULL) gets the number loaded.  Replaces
 * loadFromStream (include/gpu/fakegpuminhasher.cuh:505-525,
 * src/gpu/gpuminhasherconstruction.cu:313-320). */
int care_minhash_load_file(struct care_context* ctx, const char* path, int max_maps, int* loaded_maps);

#ifdef __cplusplus
}
#endif

#endif /* CARE_HOST_H */
