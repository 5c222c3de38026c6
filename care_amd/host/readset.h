// Host read store shared by the command line (care_host.cpp) and the
// preprocessed-reads file format (storage_files.cpp).  Rows hold 2-bit
// bases, 16 per word, first base in bits 31..30; quality rows hold binned
// qualities (1 or 2 bits, first base in the top bits) or the raw bytes
// (8 bits), as include/qualityscorecompression.hpp:74-330 lays them out.
#ifndef CARE_AMD_HOST_READSET_H
#define CARE_AMD_HOST_READSET_H

#include <cstdint>
#include <vector>

struct care_readset {
    std::uint64_t n = 0;
    int maxLen = 0;
    int pitch = 0;   // words per sequence row
    int qbits = 0;   // 0 = no qualities
    int qpitch = 0;  // words per quality row
    std::vector<std::uint32_t> seqs;
    std::vector<std::int32_t> lens;
    std::vector<std::uint8_t> ambig;
    std::vector<std::uint32_t> quals;
};

#endif
