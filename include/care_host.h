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

#ifdef __cplusplus
}
#endif

#endif /* CARE_HOST_H */
